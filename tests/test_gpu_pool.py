"""Device memory is recycled by the context that allocated it (a block goes
back to the context's pool and is handed out again in the order of its
stream).  Results must not depend on what a recycled block held: graphs of
different sizes are made, partitioned and freed in turns, and every run is
checked against the oracle.  A graph used with a second context is freed
behind a device-wide synchronisation."""
from __future__ import annotations

import numpy as np
import pytest

import graphs
import oracle_py as O
from dmolp_amd import api

pytestmark = pytest.mark.gpu

CASES = [
    ("rand300", graphs.random_graph(300, 900, 7), 4),
    ("plaw500", graphs.powerlaw(500, 3000, 11), 8),
    ("grid20x30", graphs.grid(20, 30), 16),
    ("rand2000", graphs.random_graph(2000, 12000, 5), 32),
]


def run_and_check(ctx, off, col, k):
    g = ctx.upload(off, col)
    try:
        w1 = O.phase1(off, col, k, 10000)
        p1 = g.phase1(k, 10000)
        assert np.array_equal(p1.labels, w1.labels)
        w2 = O.phase2(off, col, k, w1.labels, 5, 0.03)
        p2 = g.phase2(k, 5, 0.03)
        assert np.array_equal(p2.labels, w2.labels)
        assert p2.metrics.edge_cut == w2.m.edge_cut
    finally:
        g.free()


def test_graphs_in_turns_reuse_memory(ctx):
    for _ in range(3):
        for _, (off, col), k in CASES:
            run_and_check(ctx, off, col, k)
        for _, (off, col), k in reversed(CASES):
            run_and_check(ctx, off, col, k)


def view_in(g, other):
    """The same device graph seen through another context (not owning it)."""
    v = api.Graph.__new__(api.Graph)
    v.ctx, v._h, v.n, v.nnz = other, g.handle, g.n, g.nnz
    return v


@pytest.mark.parametrize("free_with", ["maker", "other"])
def test_graph_used_and_freed_across_contexts(ctx, free_with):
    off, col = CASES[1][1]
    k = 8
    w1 = O.phase1(off, col, k, 10000)
    w2 = O.phase2(off, col, k, w1.labels, 5, 0.03)
    other = api.Context(0)
    try:
        g = ctx.upload(off, col)
        view = view_in(g, other)
        p1 = view.phase1(k, 10000)
        p2 = view.phase2(k, 5, 0.03)
        assert np.array_equal(p1.labels, w1.labels)
        assert np.array_equal(p2.labels, w2.labels)
        if free_with == "maker":
            view._h = None
            g.free()
        else:
            g._h = None
            view.free()
    finally:
        other.close()
    # Both pools keep working afterwards.
    run_and_check(ctx, off, col, k)
    with api.Context(0) as third:
        run_and_check(third, off, col, k)
