"""Brute-force pin of the oracle's finite-theta expansion (ref_expand).

`expand_bf` restates partition.cpp:320-604 for one rank directly with
Python lists and sets, written from the reference's round structure rather
than from the oracle:

  - seeds label their parts and form the first frontiers (:339-352);
  - when every frontier is empty, every unlabelled vertex takes the part of
    smallest size (ties: lowest part) and joins its frontier (:377-394);
  - all frontier members are ordered by (degree, part, id) and each part
    selects its first theta of them (:400-425); the rest carry over
    (:427-435);
  - the selected members, part by part, push their unlabelled neighbours
    into next frontiers, each vertex once over all parts (the shared
    `visited` set, :446-465), so a vertex goes to the lowest part that
    reaches it; with one rank each claim list holds that single part and
    it wins (:530-545);
  - the next frontier of a part is its carried members, then its claims
    (:583-592).

Random graphs (some with isolated components, so the fill runs) at
theta in {1, 2, 3, 7}, seeds from the oracle's seed search and random
seeds: labels, round count and fill round must all agree.
"""
from __future__ import annotations

import numpy as np
import pytest

import graphs
import oracle_py as O


def expand_bf(off, col, k, theta, seeds):
    n = len(off) - 1
    adj = [[int(x) for x in col[off[v]:off[v + 1]]] for v in range(n)]
    deg = [len(a) for a in adj]
    label = [-1] * n
    frontier = [[] for _ in range(k)]
    for p, s in enumerate(seeds):
        label[s] = p
        frontier[p].append(s)
    rounds, fill_round = 0, -1
    while sum(1 for x in label if x >= 0) < n:
        if all(not f for f in frontier):
            size = [label.count(p) for p in range(k)]
            best = min(range(k), key=lambda p: (size[p], p))
            for v in range(n):
                if label[v] < 0:
                    label[v] = best
                    frontier[best].append(v)
            if fill_round < 0:
                fill_round = rounds
        tasks = sorted((deg[u], p, u) for p in range(k) for u in frontier[p])
        picked = [0] * k
        selected = set()
        for _, p, u in tasks:
            if picked[p] < theta:
                selected.add((p, u))
                picked[p] += 1
        carry = [[u for u in frontier[p] if (p, u) not in selected] for p in range(k)]
        visited = set()
        claims = [[] for _ in range(k)]
        for p in range(k):
            for u in frontier[p]:
                if (p, u) not in selected:
                    continue
                for v in adj[u]:
                    if label[v] < 0 and v not in visited:
                        visited.add(v)
                        claims[p].append(v)
        for p in range(k):
            for v in claims[p]:
                label[v] = p
        frontier = [carry[p] + claims[p] for p in range(k)]
        rounds += 1
    return np.array(label, np.int32), rounds, fill_round


def with_islands(off, col, extra, seed):
    """The graph plus `extra` vertices in small paths and singletons."""
    n = len(off) - 1
    rng = np.random.default_rng(seed)
    edges = [(v, int(u)) for v in range(n) for u in col[off[v]:off[v + 1]] if v < u]
    v = n
    while v < n + extra:
        ln = int(rng.integers(1, 4))
        for i in range(ln - 1):
            if v + i + 1 < n + extra:
                edges.append((v + i, v + i + 1))
        v += ln
    return graphs.from_edges(n + extra, edges)


CASES = [
    ("random", lambda s: graphs.random_graph(120, 260, s)),
    ("powerlaw", lambda s: graphs.powerlaw(150, 420, s)),
    ("grid", lambda s: graphs.grid(7, 9)),
    ("islands", lambda s: with_islands(*graphs.random_graph(90, 150, s), 17, s)),
]


@pytest.mark.parametrize("theta", [1, 2, 3, 7])
@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_expand_matches_brute_force(theta, name, make, seed):
    off, col = make(seed)
    n = len(off) - 1
    rng = np.random.default_rng(100 + seed)
    for k in (2, 3, 5):
        for how in ("oracle", "random"):
            if how == "oracle":
                seeds, _, _ = O.seeds(off, col, k)
                seeds = [int(s) for s in seeds]
            else:
                seeds = [int(s) for s in rng.choice(n, k, replace=False)]
            want, wr, wf = expand_bf(off, col, k, theta, seeds)
            got, gr, gf = O.expand(off, col, k, theta, seeds)
            assert np.array_equal(got, want), (k, how, seeds)
            assert (gr, gf) == (wr, wf), (k, how, seeds)


def test_brute_force_hand_trace():
    # The 3x3 grid trace at theta = 1, derived by hand in test_oracle_phase1:
    # the brute force itself reproduces it.
    off, col = graphs.grid(3, 3)
    lab, rounds, fill = expand_bf(off, col, 2, 1, [0, 8])
    assert lab.tolist() == [0, 0, 0, 0, 0, 1, 1, 1, 1] and rounds == 3 and fill == -1
