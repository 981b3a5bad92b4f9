"""Phase 1 orders only a prefix of the hub list on the host before its seed
search (dmolp_ctx_set_hub_prefix), and repeats a search that runs out of
listed hubs with the whole list.  Seeds, labels and metrics must equal the
oracle's whatever the prefix: prefix 1 makes almost every search run out
(the repeat path), larger ones let most finish inside the prefix, 0 orders
every hub.  Grids make every vertex a hub (near-regular degrees)."""
from __future__ import annotations

import numpy as np
import pytest

import graphs
import oracle_py as O
from dmolp_amd import api

pytestmark = pytest.mark.gpu

GRAPHS = {
    "grid30x30": graphs.grid(30, 30),
    "star12": graphs.star(12),
    "cliques6": graphs.two_cliques(6),
    "rand300": graphs.random_graph(300, 900, 7),
    "plaw500": graphs.powerlaw(500, 3000, 11),
}


@pytest.fixture
def prefixed(ctx):
    yield ctx
    ctx.set_hub_prefix(1 << 16)


@pytest.mark.parametrize("name", list(GRAPHS))
@pytest.mark.parametrize("prefix", [1, 2, 8, 64, 0])
def test_phase1_equals_oracle_for_any_hub_prefix(prefixed, name, prefix):
    off, col = GRAPHS[name]
    n = len(off) - 1
    prefixed.set_hub_prefix(prefix)
    g = prefixed.upload(off, col)
    try:
        for k in (2, 8, 16):
            if k > n:
                continue
            want = O.phase1(off, col, k, 10000)
            got = g.phase1(k, 10000)
            assert np.array_equal(want.seeds, got.seeds), k
            assert np.array_equal(want.labels, got.labels), k
            assert want.m.n_hubs == got.metrics.n_hubs
            assert want.m.n_seeds_found == got.metrics.n_seeds_found
            assert want.m.rounds == got.metrics.rounds
    finally:
        g.free()


def test_batch_with_short_hub_prefix_equals_oracle(prefixed):
    prefixed.set_hub_prefix(2)
    jobs, want = [], []
    for name, k in (("grid30x30", 8), ("rand300", 16), ("plaw500", 4), ("cliques6", 2)):
        off, col = GRAPHS[name]
        g = prefixed.upload(off, col)
        jobs.append(api.BatchJob(g, k, 10000, 3, 0.03))
        o1 = O.phase1(off, col, k, 10000)
        want.append(O.phase2(off, col, k, o1.labels, 3, 0.03))
    try:
        res, _ = api.batch_run(prefixed, jobs, want_labels=True)
        for r, w in zip(res, want):
            assert r.status == 0
            assert np.array_equal(r.labels, w.labels)
            assert r.edge_cut == w.m.edge_cut
    finally:
        for j in jobs:
            j.graph.free()
