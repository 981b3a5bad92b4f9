"""Small named test graphs as symmetric CSR (int64 offsets, int32 columns)."""
from __future__ import annotations

import numpy as np


def from_edges(n, edges):
    adj = [set() for _ in range(n)]
    for u, v in edges:
        if u != v:
            adj[u].add(v)
            adj[v].add(u)
    off = np.zeros(n + 1, np.int64)
    for i in range(n):
        off[i + 1] = off[i] + len(adj[i])
    col = np.array([v for i in range(n) for v in sorted(adj[i])], np.int32)
    return off, col


def path(n):
    return from_edges(n, [(i, i + 1) for i in range(n - 1)])


def star(n):
    return from_edges(n, [(0, i) for i in range(1, n)])


def grid(r, c):
    e = []
    for i in range(r):
        for j in range(c):
            v = i * c + j
            if j + 1 < c:
                e.append((v, v + 1))
            if i + 1 < r:
                e.append((v, v + c))
    return from_edges(r * c, e)


def two_cliques(m):
    e = [(i, j) for i in range(m) for j in range(i + 1, m)]
    e += [(m + i, m + j) for i in range(m) for j in range(i + 1, m)]
    e.append((m - 1, m))
    return from_edges(2 * m, e)


def random_graph(n, m, seed):
    rng = np.random.default_rng(seed)
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    return from_edges(n, list(zip(u.tolist(), v.tolist())))


def powerlaw(n, m, seed):
    """Skewed random graph: endpoints drawn with weight 1/(i+1)."""
    rng = np.random.default_rng(seed)
    w = 1.0 / np.arange(1, n + 1)
    w /= w.sum()
    u = rng.choice(n, m, p=w)
    v = rng.choice(n, m, p=w)
    return from_edges(n, list(zip(u.tolist(), v.tolist())))


def to_nx(off, col):
    import networkx as nx
    G = nx.Graph()
    n = len(off) - 1
    G.add_nodes_from(range(n))
    for u in range(n):
        for e in range(off[u], off[u + 1]):
            G.add_edge(u, int(col[e]))
    return G
