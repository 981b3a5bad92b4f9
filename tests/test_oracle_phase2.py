"""Pinning the oracle's phase-2 functions: the literal 32-lane shuffle-tree
of the reference kernel simulated in Python, penalty closed forms, a
brute-force boundary selection, numpy recounts, and an independent
pure-Python composition of a whole phase-2 run on small graphs."""
from __future__ import annotations

import math

import numpy as np
import pytest

import graphs
import oracle_py as O


# ---------------------------------------------------------------- LP decision

def warp_tree_decision(hist, pen, own, k):
    """gpu_lp_boundary.cu:386-426 simulated lane by lane: 32 lanes, label =
    lane, score = count*(1+pen) for non-zero counts, -1e300 beyond
    min(k,32); __shfl_down_sync tree with offsets 16..1 keeping the own value
    unless the other is strictly greater (lanes past 31 read their own
    value); lane 0 then compares with the raw own-label count."""
    E = min(k, 32)
    score = [-1e300] * 32
    label = list(range(32))
    for l in range(E):
        s = float(hist[l])
        if s > 0.0:
            s *= (1.0 + pen[l])
        score[l] = s
    off = 16
    while off > 0:
        new_s, new_l = score[:], label[:]
        for lane in range(32):
            src = lane + off
            os_, ol = (score[src], label[src]) if src < 32 else (score[lane], label[lane])
            if os_ > score[lane]:
                new_s[lane], new_l[lane] = os_, ol
        score, label = new_s, new_l
        off >>= 1
    base = float(hist[own]) if 0 <= own < E else -1e300
    return label[0] if score[0] > base else own


def test_tie_order_is_bit_reversed():
    # Ties among labels {1, 2, 3} go to 2 (bit-reversed 5-bit order).
    hist = [0, 5, 5, 5] + [0] * 28
    assert warp_tree_decision(hist, [0.0] * 32, 0, 32) == 2
    off, col = graphs.star(16)
    labels = np.array([0] + [1] * 5 + [2] * 5 + [3] * 5, np.int32)
    out = O.lp_round(off, col, labels, [0], np.zeros(32), 32)
    assert out[0] == 2


def _star_case(counts, own, k, pen):
    n = 1 + sum(counts)
    off, col = graphs.star(n)
    labels = np.zeros(n, np.int32)
    labels[0] = own
    i = 1
    for l, c in enumerate(counts):
        labels[i:i + c] = l
        i += c
    return off, col, labels


@pytest.mark.parametrize("seed", range(6))
def test_lp_decision_matches_warp_tree(seed):
    rng = np.random.default_rng(seed)
    for _ in range(60):
        k = int(rng.choice([2, 3, 8, 31, 32, 40, 64]))
        E = min(k, 32)
        counts = [int(x) for x in rng.integers(0, 4, min(k, 34))]
        own = int(rng.integers(0, k))
        pen = rng.choice([0.0, 0.5, -0.5, -1.0, -2.0, 1e-17], size=k)
        off, col, labels = _star_case(counts, own, k, pen)
        got = O.lp_round(off, col, labels, [0], pen, k)[0]
        hist = np.bincount(labels[1:], minlength=max(k, 34))[:32]
        hist = [int(h) if l < E else 0 for l, h in enumerate(hist)]
        assert got == warp_tree_decision(hist, list(pen) + [0] * 32, own, k)


def test_lp_quirks():
    # Own label >= 32 with k > 32 is always replaced (base is -1e300).
    off, col, labels = _star_case([0, 1], 40, 64, None)
    assert O.lp_round(off, col, labels, [0], np.zeros(64), 64)[0] == 1
    # A zero score never displaces an own count >= 0.
    off, col, labels = _star_case([0, 0, 3], 5, 8, None)
    pen = np.zeros(8)
    pen[2] = -1.0  # score 3 * (1 - 1) = 0 > own count 0 is false
    assert O.lp_round(off, col, labels, [0], pen, 8)[0] == 5
    # The own count is compared unpenalised: 2 neighbours of the own label
    # beat 2 neighbours of label 1 scaled by 0.9.
    off, col, labels = _star_case([2, 2], 0, 4, None)
    pen = np.array([5.0, -0.1, 0.0, 0.0])
    assert O.lp_round(off, col, labels, [0], pen, 4)[0] == 0


# ---------------------------------------------------------------- penalties

def test_penalties_special_cases():
    # t = 1: 1 + log(1) = 1.  vc = [3, 1], ic = [2, 0]: RV = [1.5, .5],
    # RE = [2, 0], var_V = .25, var_E = 1, weights .2 / .8.
    pen, re = O.penalties(2, [3, 1], [2, 0], 1)
    assert re.tolist() == [2.0, 0.0]
    assert pen[0] == pytest.approx(-0.9, abs=1e-15)
    assert pen[1] == pytest.approx(0.9, abs=1e-15)
    pen3, _ = O.penalties(2, [3, 1], [2, 0], 3)
    assert pen3 == pytest.approx(pen * (1 + math.log(3)), rel=1e-15)
    # Balanced: zero variance, zero penalties.
    pen, re = O.penalties(4, [5] * 4, [7] * 4, 9)
    assert np.all(pen == 0.0) and np.all(re == 1.0)
    # No intra edges: expected edges default to 1.0, RE = ic / 1.
    pen, re = O.penalties(3, [1, 2, 3], [0, 0, 0], 2)
    assert np.all(re == 0.0)
    assert pen == pytest.approx([1 - 0.5, 0.0, 1 - 1.5] * np.array(1 + math.log(2)), abs=1e-15)


# ---------------------------------------------------------------- selection

def brute_select(off, col, labels, prev, re, k, t):
    n = len(off) - 1
    cand = set()
    for v in prev:
        cand.add(int(v))
        cand.update(int(u) for u in col[off[v]:off[v + 1]])
    groups = {}
    for x in sorted(cand):
        nb = col[off[x]:off[x + 1]]
        if (labels[nb] != labels[x]).any():
            groups.setdefault(int(labels[x]), []).append(x)
    ratio = math.exp(-t / 5.0)
    out = []
    deg = np.diff(off)
    for p, xs in groups.items():
        if re[p] > 1.0:
            xs = sorted(xs, key=lambda x: (-deg[x], x))
        else:
            xs = sorted(xs, key=lambda x: (deg[x], x))
        take = min(len(xs), math.ceil(ratio * len(xs)))
        out += xs[:take]
    assert all(0 <= x < n for x in out)
    return sorted(out)


@pytest.mark.parametrize("seed", range(5))
def test_select_boundary_brute_force(seed):
    rng = np.random.default_rng(seed)
    off, col = graphs.powerlaw(400, 1500, seed)
    for t in (1, 2, 4, 7, 15):
        k = int(rng.integers(2, 9))
        labels = rng.integers(0, k, 400).astype(np.int32)
        prev = np.sort(rng.choice(400, 120, replace=False)).astype(np.int32)
        re = rng.choice([0.5, 1.0, 1.0000000000000002, 2.0], size=k)
        want = brute_select(off, col, labels, prev, re, k, t)
        got = O.select_boundary(off, col, labels, prev, re, k, t)
        assert got.tolist() == want


# ---------------------------------------------------------------- stats

def test_stats_recount_and_errors():
    off, col = graphs.random_graph(300, 1200, 4)
    rng = np.random.default_rng(0)
    lab = rng.integers(0, 6, 300).astype(np.int32)
    rc, vc, ic, cut = O.stats(off, col, 6, lab)
    src = np.repeat(np.arange(300), np.diff(off))
    d = lab[src] != lab[col]
    assert rc == 0 and cut == d.sum() // 2
    assert np.array_equal(vc, np.bincount(lab, minlength=6))
    assert np.array_equal(ic, np.bincount(lab[src][~d], minlength=6) // 2)
    assert cut + ic.sum() == len(col) // 2
    lab2 = lab.copy()
    lab2[int(col[0])] = -1
    assert O.stats(off, col, 6, lab2)[0] == -3  # a -1 label reached the cut
    # An asymmetric CSR gives an odd cut.
    off3 = np.array([0, 1, 1], np.int64)
    col3 = np.array([1], np.int32)
    assert O.stats(off3, col3, 2, np.array([0, 1], np.int32))[0] == -4


# ---------------------------------------------------------------- whole phase 2

def py_phase2(off, col, k, labels, R, eps):
    """phase2.cpp:256-449 composed from the Python pieces above."""
    L = np.array(labels, np.int64)
    n = len(off) - 1
    src = np.repeat(np.arange(n), np.diff(off))

    def stats(L):
        d = L[src] != L[col]
        vc = np.bincount(L, minlength=k)[:k]
        ic = np.bincount(L[src][~d], minlength=k)[:k] // 2
        return vc, ic, int(d.sum() // 2)

    vc, ic, prev = stats(L)
    cc = 0
    B = list(range(n))
    rounds, reason = 0, 0
    for t in range(1, R + 1):
        rounds = t
        pen, re = O.penalties(k, vc, ic, t)
        B = brute_select(off, col, L, B, re, k, t)
        if not B:
            reason = 1
            break
        newL = L.copy()
        for b in B:
            E = min(k, 32)
            nb = L[col[off[b]:off[b + 1]]]
            hist = [int((nb == l).sum()) if l < E else 0 for l in range(32)]
            newL[b] = warp_tree_decision(hist, list(pen) + [0] * 32, int(L[b]), k)
        changed = int((newL != L).sum())
        L = newL
        if changed == 0:
            reason = 2
            break
        vc, ic, cur = stats(L)
        delta = abs((cur - prev) / prev) if prev > 0 else 1.0
        cc = cc + 1 if delta < eps else 0
        prev = cur
        if cc >= 10:
            reason = 3
            break
    return L, prev, rounds, reason


@pytest.mark.parametrize("name,g,k", [
    ("grid", graphs.grid(9, 11), 3),
    ("plaw", graphs.powerlaw(300, 1200, 3), 4),
    ("rand", graphs.random_graph(250, 700, 8), 5),
    ("cliques", graphs.two_cliques(7), 2),
])
def test_phase2_composition(name, g, k):
    off, col = g
    rng = np.random.default_rng(2)
    lab0 = rng.integers(0, k, len(off) - 1).astype(np.int32)
    for R, eps in ((12, 0.03), (25, 0.5)):
        want_L, want_cut, want_rounds, want_reason = py_phase2(off, col, k, lab0, R, eps)
        r = O.phase2(off, col, k, lab0, R, eps)
        assert np.array_equal(r.labels, want_L)
        assert r.m.edge_cut == want_cut and r.m.rounds == want_rounds
        assert r.m.exit_reason == want_reason


def test_phase2_exits():
    off, col = graphs.grid(6, 6)
    # One label everywhere: no boundary in round 1.
    r = O.phase2(off, col, 2, np.zeros(36, np.int32), 10, 0.03)
    assert r.m.rounds == 1 and r.m.exit_reason == 1 and r.m.edge_cut == 0
    # Zero rounds: the initial cut is reported.
    lab = (np.arange(36) % 2).astype(np.int32)
    r = O.phase2(off, col, 2, lab, 0, 0.03)
    _, _, _, cut = O.stats(off, col, 2, lab)
    assert r.m.rounds == 0 and r.m.edge_cut == cut
