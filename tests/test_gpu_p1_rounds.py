"""Phase-1 rounds that need no selection run for each job in one workgroup
(k_p1_rounds_wg, block barriers only) and hand over to the per-kernel launches
at the first round that does.  Every run here must equal the oracle (labels,
seeds, rounds, fill round): small theta makes runs switch between the two ways
round by round (and leave the workgroup rounds after odd and even round
counts), graphs with many components make the fill run inside them, and a
batch mixes jobs that need selections with jobs that never do."""
from __future__ import annotations

import numpy as np
import pytest

import graphs
import oracle_py as O
from dmolp_amd import api

pytestmark = pytest.mark.gpu


def islands(n_main, m_main, extra, seed):
    off, col = graphs.random_graph(n_main, m_main, seed)
    rng = np.random.default_rng(seed)
    edges = [(v, int(u)) for v in range(n_main) for u in col[off[v]:off[v + 1]] if v < u]
    v = n_main
    while v < n_main + extra:
        ln = int(rng.integers(1, 5))
        for i in range(ln - 1):
            if v + i + 1 < n_main + extra:
                edges.append((v + i, v + i + 1))
        v += ln
    return graphs.from_edges(n_main + extra, edges)


CASES = {
    "grid40x50": lambda: graphs.grid(40, 50),
    "path300": lambda: graphs.path(300),
    "islands": lambda: islands(400, 700, 120, 3),
    "sparse": lambda: graphs.random_graph(900, 1000, 8),
    "plaw": lambda: graphs.powerlaw(1500, 6000, 2),
    "rmat12": lambda: api.gen_host(api.rmat_spec(12, 8, seed=21)),
}


def check(got, want):
    assert np.array_equal(got.labels, want.labels)
    assert np.array_equal(np.asarray(got.seeds), np.asarray(want.seeds))
    assert got.metrics.rounds == want.m.rounds
    assert got.metrics.fill_round == want.m.fill_round


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("theta", [1, 2, 3, 7, 40, 10 ** 6])
def test_phase1_rounds_match_oracle(ctx, name, theta):
    off, col = CASES[name]()
    g = ctx.upload(off, col)
    try:
        for k in (2, 5):
            print(f"[p1rounds] {name} theta={theta} k={k}", flush=True)
            want = O.phase1(off, col, k, theta)
            got = g.phase1(k, theta)
            check(got, want)
    finally:
        g.free()


def test_batch_mixes_selection_and_persistent_rounds(ctx):
    specs = [CASES[n]() for n in ("grid40x50", "islands", "rmat12", "plaw", "path300", "sparse")]
    specs += [graphs.grid(r, c) for r, c in ((30, 90), (64, 64), (7, 300))]
    ks = [2, 4, 8, 3, 2, 6, 5, 16, 2]
    thetas = [10000, 10000, 50, 10000, 10000, 10000, 10000, 10000, 10000]
    gs = [ctx.upload(o, c) for o, c in specs]
    try:
        jobs = [api.BatchJob(g, k, th, 8, 0.02) for g, k, th in zip(gs, ks, thetas)]
        res, _ = api.batch_run(ctx, jobs, want_labels=True)
        for (o, c), k, th, r in zip(specs, ks, thetas, res):
            w1 = O.phase1(o, c, k, th)
            w2 = O.phase2(o, c, k, w1.labels, 8, 0.02)
            assert np.array_equal(r.labels, w2.labels)
            assert r.edge_cut == w2.m.edge_cut
    finally:
        for g in gs:
            g.free()
