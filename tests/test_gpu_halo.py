"""Rank shares with an on-device halo (dmolp_shares_*, dmolp_halo_refresh,
dmolp_phase2_shares) simulating P ranks on one GPU:

  - each share's layout equals the reference's (partition.cpp:182-262,
    restated below on the host): owned rows in global-id order, owned
    neighbours as local ids, ghosts numbered after them in order of first
    appearance; ghost labels are their owners' labels (partition.cpp:264-318);
  - the halo refresh rewrites exactly the ghosts named by a delta list;
  - phase 2 over the shares equals the oracle's P-rank mode for per-rank
    selection (the reference's) and the one-rank oracle and dmolp_phase2
    for selection over all ranks, labels and counts bit for bit, with the
    same ghost refresh counts.
"""
from __future__ import annotations

import numpy as np
import pytest

import graphs
import oracle_py as O
from dmolp_amd import api

pytestmark = pytest.mark.gpu

GRAPHS = {
    "grid9x11": graphs.grid(9, 11),
    "rand400": graphs.random_graph(400, 1600, 3),
    "plaw600": graphs.powerlaw(600, 4000, 5),
}


def reference_share(off, col, labels, P, r):
    """partition.cpp:182-262 and :264-318 for rank r, on the host."""
    n = len(off) - 1
    locals_ = list(range(r, n, P))
    g2l = {v: i for i, v in enumerate(locals_)}
    soff, scol, ghosts, slot = [0], [], [], {}
    for v in locals_:
        for u in col[off[v]:off[v + 1]]:
            u = int(u)
            if u % P == r:
                scol.append(g2l[u])
            else:
                if u not in slot:
                    slot[u] = len(ghosts)
                    ghosts.append(u)
                scol.append(len(locals_) + slot[u])
        soff.append(len(scol))
    lab = [int(labels[v]) for v in locals_] + [int(labels[u]) for u in ghosts]
    return (np.array(soff, np.int64), np.array(scol, np.int32), np.array(ghosts, np.int32),
            np.array(lab, np.int32))


@pytest.mark.parametrize("name", list(GRAPHS))
@pytest.mark.parametrize("P", [1, 2, 3, 4])
def test_share_layout_and_ghost_fetch(ctx, name, P):
    off, col = GRAPHS[name]
    n = len(off) - 1
    labels = np.random.default_rng(P).integers(0, 6, n).astype(np.int32)
    g = ctx.upload(off, col)
    g.set_labels(labels)
    sh = api.Shares(g, P)
    try:
        for r in range(P):
            want = reference_share(off, col, labels, P, r)
            got = sh.download(r)
            for w, x, what in zip(want, got, ("off", "col", "ghost ids", "labels")):
                assert np.array_equal(w, x), (r, what)
    finally:
        sh.close()
        g.free()


def test_halo_refresh_rewrites_named_ghosts(ctx):
    off, col = GRAPHS["plaw600"]
    n = len(off) - 1
    labels = np.zeros(n, np.int32)
    g = ctx.upload(off, col)
    g.set_labels(labels)
    sh = api.Shares(g, 3)
    rng = np.random.default_rng(4)
    try:
        for r in range(3):
            _, _, ghosts, lab0 = sh.download(r)
            nl = len(lab0) - len(ghosts)
            gid = rng.choice(n, size=200, replace=False).astype(np.int32)
            new = rng.integers(1, 9, 200).astype(np.int32)
            updated = sh.refresh(r, gid, new)
            want = lab0.copy()
            pos = {int(u): s for s, u in enumerate(ghosts)}
            hits = 0
            for u, l in zip(gid, new):
                if int(u) in pos:
                    want[nl + pos[int(u)]] = l
                    hits += 1
            assert updated == hits
            assert np.array_equal(sh.download(r)[3], want)
            assert sh.refresh(r, np.zeros(0, np.int32), np.zeros(0, np.int32)) == 0
    finally:
        sh.close()
        g.free()


def _same(want, got, gu=None):
    assert np.array_equal(want.labels, got.labels)
    assert want.m.edge_cut == got.metrics.edge_cut
    assert want.m.rounds == got.metrics.rounds
    assert want.m.exit_reason == got.metrics.exit_reason
    assert want.m.lp_edges == got.metrics.lp_edges
    assert want.m.initial_cut == got.metrics.initial_cut
    assert np.array_equal(want.vc, got.vc) and np.array_equal(want.ic, got.ic)
    assert want.m.vertex_balance == got.metrics.vertex_balance
    assert want.m.edge_balance == got.metrics.edge_balance
    for f in ("prev_count", "prev_deg", "bd_count", "b_count", "b_deg", "chg_count", "chg_deg",
              "cut"):
        assert np.array_equal(want.trace[f], got.trace[f]), f
    if gu is not None:
        assert np.array_equal(gu, got.trace["ghost_updates"])


CASES = [(name, P) for name in GRAPHS for P in (2, 4)] + [("rmat12", 2), ("rmat12", 4)]


def _graph(name):
    if name == "rmat12":
        return api.gen_host(api.rmat_spec(12, 16, seed=8))
    return GRAPHS[name]


@pytest.mark.parametrize("name,P", CASES)
def test_per_rank_selection_matches_oracle(ctx, name, P):
    off, col = _graph(name)
    n = len(off) - 1
    g = ctx.upload(off, col)
    rng = np.random.default_rng(P)
    try:
        for k in (3, 8):
            labels0 = rng.integers(0, k, n).astype(np.int32)
            want = O.phase2_ranks(off, col, k, labels0, P, 0, 10, 0.02)
            g.set_labels(labels0)
            sh = api.Shares(g, P)
            got = sh.phase2(k, 10, 0.02, 0)
            sh.close()
            _same(want, got, want.trace["ghost_updates"])
            assert np.array_equal(g.get_labels(), want.labels)
    finally:
        g.free()


@pytest.mark.parametrize("name,P", CASES)
def test_global_selection_matches_one_rank(ctx, name, P):
    off, col = _graph(name)
    n = len(off) - 1
    g = ctx.upload(off, col)
    try:
        labels0 = np.random.default_rng(9).integers(0, 6, n).astype(np.int32)
        want = O.phase2(off, col, 6, labels0, 12, 0.02)
        ranks = O.phase2_ranks(off, col, 6, labels0, P, 1, 12, 0.02)
        g.set_labels(labels0)
        sh = api.Shares(g, P)
        got = sh.phase2(6, 12, 0.02, 1)
        sh.close()
        _same(want, got, ranks.trace["ghost_updates"])
        g.set_labels(labels0)
        plain = g.phase2(6, 12, 0.02)
        assert np.array_equal(plain.labels, got.labels)
    finally:
        g.free()


def test_phase1_then_shares(ctx):
    off, col = api.gen_host(api.rmat_spec(13, 16, seed=12))
    g = ctx.upload(off, col)
    try:
        p1 = g.phase1(16, 500)
        want = O.phase2_ranks(off, col, 16, p1.labels, 4, 0, 15, 0.03)
        sh = api.Shares(g, 4)
        got = sh.phase2(16, 15, 0.03, 0)
        sh.close()
        _same(want, got, want.trace["ghost_updates"])
    finally:
        g.free()


def test_labels_outside_k_are_rejected(ctx):
    off, col = GRAPHS["grid9x11"]
    g = ctx.upload(off, col)
    labels = np.zeros(len(off) - 1, np.int32)
    labels[5] = 7
    g.set_labels(labels)
    sh = api.Shares(g, 2)
    try:
        with pytest.raises(api.DmolpError) as e:
            sh.phase2(4, 5, 0.02, 0)
        assert e.value.status == -3
    finally:
        sh.close()
        g.free()
