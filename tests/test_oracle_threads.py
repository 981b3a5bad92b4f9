"""The oracle's OpenMP sweeps give the same results on any thread count:
phase 1, phase 2 (with its trace), the P-rank mode and the pieces, at one
thread and at several, on graphs where ties and claims collide."""
from __future__ import annotations

import numpy as np
import pytest

import graphs
import oracle_py as O
from dmolp_amd import api


@pytest.fixture
def restore_threads():
    t = O.get_threads()
    yield
    O.set_threads(t)


@pytest.mark.parametrize("spec", ["rmat14", "plaw", "grid"])
def test_results_do_not_depend_on_threads(restore_threads, spec):
    if spec == "rmat14":
        off, col = api.gen_host(api.rmat_spec(14, 16, seed=17))
        k, theta = 16, 300
    elif spec == "plaw":
        off, col = graphs.powerlaw(3000, 30000, 4)
        k, theta = 8, 50
    else:
        off, col = graphs.grid(40, 50)
        k, theta = 5, 7
    runs = []
    for threads in (1, 4):
        O.set_threads(threads)
        p1 = O.phase1(off, col, k, theta)
        p2 = O.phase2(off, col, k, p1.labels, 15, 0.01)
        pr = O.phase2_ranks(off, col, k, p1.labels, 3, 0, 8, 0.01)
        runs.append((p1, p2, pr))
    (a1, a2, ar), (b1, b2, br) = runs
    assert np.array_equal(a1.labels, b1.labels) and np.array_equal(a1.seeds, b1.seeds)
    assert a1.m.rounds == b1.m.rounds and a1.m.initial_edge_cut == b1.m.initial_edge_cut
    assert np.array_equal(a2.labels, b2.labels) and a2.m.edge_cut == b2.m.edge_cut
    for f in a2.trace:
        assert np.array_equal(a2.trace[f], b2.trace[f]), f
    assert np.array_equal(ar.labels, br.labels)
    assert np.array_equal(ar.trace["ghost_updates"], br.trace["ghost_updates"])
