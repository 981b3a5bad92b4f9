"""The oracle's P-rank phase 2 (ref_phase2_ranks), pinned without the
reference:

  - selection over all ranks (scope 1) must give the one-rank run for any P:
    the shares, ghost columns, ghost labels and their refresh from the
    deltas are then the only moving parts, and any slip in them (a ghost
    label left stale, a wrong local or ghost index, a lost delta) changes
    the labels;
  - per-rank selection (scope 0) at P = 1 is the one-rank run;
  - per-rank selection at P > 1 against a simulation written here on global
    arrays (no shares or ghosts at all): rank r's candidates are its owned
    vertices in or next to its own previous boundary, each label's take is
    counted over rank r's boundary vertices, and decisions are the oracle's
    one-rank LP on the global labels (itself pinned elsewhere).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import graphs
import oracle_py as O
from dmolp_amd import api

GRAPHS = {
    "grid9x11": graphs.grid(9, 11),
    "rand400": graphs.random_graph(400, 1600, 3),
    "plaw600": graphs.powerlaw(600, 4000, 5),
    "cliques7": graphs.two_cliques(7),
}


def same_run(a, b):
    assert np.array_equal(a.labels, b.labels)
    assert a.m.edge_cut == b.m.edge_cut and a.m.rounds == b.m.rounds
    assert a.m.exit_reason == b.m.exit_reason and a.m.lp_edges == b.m.lp_edges
    assert np.array_equal(a.vc, b.vc) and np.array_equal(a.ic, b.ic)
    for f in ("b_count", "b_deg", "chg_count", "chg_deg", "cut", "bd_count"):
        assert np.array_equal(a.trace[f], b.trace[f]), f


@pytest.mark.parametrize("name", list(GRAPHS))
@pytest.mark.parametrize("nranks", [1, 2, 3, 4])
def test_global_selection_equals_one_rank(name, nranks):
    off, col = GRAPHS[name]
    rng = np.random.default_rng(nranks)
    for k in (2, 5, 8):
        labels0 = rng.integers(0, k, len(off) - 1).astype(np.int32)
        want = O.phase2(off, col, k, labels0, 12, 0.02)
        got = O.phase2_ranks(off, col, k, labels0, nranks, 1, 12, 0.02)
        same_run(want, got)


def test_rmat_global_selection_equals_one_rank():
    off, col = api.gen_host(api.rmat_spec(13, 16, seed=41))
    p1 = O.phase1(off, col, 16, 1000)
    want = O.phase2(off, col, 16, p1.labels, 15, 0.03)
    for nranks in (2, 4):
        same_run(want, O.phase2_ranks(off, col, 16, p1.labels, nranks, 1, 15, 0.03))


@pytest.mark.parametrize("name", list(GRAPHS))
def test_per_rank_selection_at_one_rank(name):
    off, col = GRAPHS[name]
    labels0 = np.random.default_rng(7).integers(0, 4, len(off) - 1).astype(np.int32)
    same_run(O.phase2(off, col, 4, labels0, 10, 0.02),
             O.phase2_ranks(off, col, 4, labels0, 1, 0, 10, 0.02))


def sim_per_rank(off, col, k, labels0, P, rounds, eps):
    """Per-rank selection on global arrays (no shares, no ghosts)."""
    n = len(off) - 1
    deg = np.diff(off)
    own = np.arange(n) % P
    L = labels0.astype(np.int32).copy()
    B = [np.arange(r, n, P) for r in range(P)]
    rc, vc, ic, prev = O.stats(off, col, k, L)
    cc, out_rounds, chg_counts = 0, 0, []
    for t in range(1, rounds + 1):
        out_rounds = t
        pen, re = O.penalties(k, vc, ic, t)
        ratio = math.exp(-t / 5.0)
        newB = []
        for r in range(P):
            inb = np.zeros(n, bool)
            inb[B[r]] = True
            cand = []
            for x in range(r, n, P):
                nb = col[off[x]:off[x + 1]]
                if inb[x] or any(inb[nb]):
                    cand.append(x)
            groups = {}
            for x in cand:
                nb = col[off[x]:off[x + 1]]
                if np.any(L[nb] != L[x]):
                    groups.setdefault(int(L[x]), []).append(x)
            sel = []
            for p, g in groups.items():
                desc = re[p] > 1.0
                g.sort(key=lambda x: ((-deg[x] if desc else deg[x]), x))
                sel += g[:min(len(g), math.ceil(ratio * len(g)))]
            newB.append(np.array(sorted(sel), np.int64))
        B = newB
        allb = np.concatenate(B).astype(np.int32)
        if len(allb) == 0:
            break
        L2 = O.lp_round(off, col, L, allb, pen, k)
        chg = int(np.sum(L2 != L))
        chg_counts.append(chg)
        L = L2
        if chg == 0:
            break
        rc, vc, ic, cur = O.stats(off, col, k, L)
        delta = abs((cur - prev) / prev) if prev > 0 else 1.0
        cc = cc + 1 if delta < eps else 0
        prev = cur
        if cc >= 10:
            break
    assert np.all(own >= 0)
    return L, out_rounds, chg_counts


@pytest.mark.parametrize("name", list(GRAPHS))
@pytest.mark.parametrize("nranks", [2, 3])
def test_per_rank_selection_matches_global_array_simulation(name, nranks):
    off, col = GRAPHS[name]
    labels0 = np.random.default_rng(11).integers(0, 5, len(off) - 1).astype(np.int32)
    got = O.phase2_ranks(off, col, 5, labels0, nranks, 0, 8, 0.02)
    L, rounds, chg = sim_per_rank(off, col, 5, labels0, nranks, 8, 0.02)
    assert np.array_equal(got.labels, L)
    assert got.m.rounds == rounds
    assert got.trace["chg_count"].tolist()[:len(chg)] == chg
    rc, vc, ic, cut = O.stats(off, col, 5, L)
    assert got.m.edge_cut == cut and np.array_equal(got.vc, vc) and np.array_equal(got.ic, ic)


def test_ghost_refresh_counts():
    """Every delta reaches each rank that holds its vertex as a ghost."""
    off, col = GRAPHS["rand400"]
    n = len(off) - 1
    labels0 = np.random.default_rng(2).integers(0, 3, n).astype(np.int32)
    P = 3
    got = O.phase2_ranks(off, col, 3, labels0, P, 0, 1, 0.0)
    # Ranks holding v as a ghost: owners of v's neighbours other than v's own.
    holders = np.zeros(n, np.int64)
    for v in range(n):
        owners = {int(u) % P for u in col[off[v]:off[v + 1]]} - {v % P}
        holders[v] = len(owners)
    changed = np.flatnonzero(got.labels != labels0)
    assert got.trace["ghost_updates"][0] == holders[changed].sum()
