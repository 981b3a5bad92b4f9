"""Batched mode: many graphs of uneven size partitioned in one job set.
Every job's labels and metrics must equal its plain (one graph at a time)
run, whatever the batch composition or order."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_py as O
from dmolp_amd import api

pytestmark = pytest.mark.gpu


def mixed_specs(seed=7, count=12):
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        k = int(rng.choice([2, 4, 8, 16, 64]))
        if i % 2 == 0:
            spec = api.rmat_spec(int(rng.integers(8, 13)), 16, seed=int(rng.integers(1, 1000)))
        else:
            r, c = (int(x) for x in np.exp(rng.uniform(np.log(12), np.log(60), 2)))
            spec = api.grid_spec(r, c)
        specs.append((spec, k))
    return specs


def run_plain(ctx, specs, theta, rounds, eps):
    out = []
    for spec, k in specs:
        g = ctx.generate(spec)
        p1 = g.phase1(k, theta)
        p2 = g.phase2(k, rounds, eps)
        out.append((p1, p2))
        g.free()
    return out


def run_batch(ctx, specs, theta, rounds, eps, order=None):
    order = list(range(len(specs))) if order is None else order
    graphs = [ctx.generate(specs[i][0]) for i in order]
    jobs = [api.BatchJob(graphs[j], specs[i][1], theta, rounds, eps) for j, i in enumerate(order)]
    res, ms = api.batch_run(ctx, jobs, want_labels=True)
    for g in graphs:
        g.free()
    back = [None] * len(specs)
    for j, i in enumerate(order):
        back[i] = res[j]
    return back, ms


@pytest.mark.parametrize("theta", [100, 10000])
def test_batch_equals_plain(ctx, theta):
    specs = mixed_specs()
    print(f"[batch] {len(specs)} jobs theta={theta}", flush=True)
    plain = run_plain(ctx, specs, theta, 10, 0.03)
    got, ms = run_batch(ctx, specs, theta, 10, 0.03)
    assert ms > 0
    for (p1, p2), r in zip(plain, got):
        assert r.status == 0
        assert np.array_equal(r.labels, p2.labels)
        assert r.edge_cut == p2.metrics.edge_cut
        assert r.phase1_cut == p1.metrics.initial_edge_cut == p2.metrics.initial_cut
        assert r.rounds == p2.metrics.rounds
        assert r.exit_reason == p2.metrics.exit_reason
        assert r.lp_edges == p2.metrics.lp_edges
        assert r.vertex_balance == p2.metrics.vertex_balance
        assert r.edge_balance == p2.metrics.edge_balance
        assert r.n_seeds_fallback == p1.metrics.n_seeds_fallback


def test_batch_is_order_and_composition_independent(ctx):
    specs = mixed_specs(seed=11, count=10)
    a, _ = run_batch(ctx, specs, 500, 12, 0.01)
    rev = list(reversed(range(len(specs))))
    b, _ = run_batch(ctx, specs, 500, 12, 0.01, order=rev)
    c, _ = run_batch(ctx, specs[:3], 500, 12, 0.01)
    for x, y in zip(a, b):
        assert np.array_equal(x.labels, y.labels) and x.edge_cut == y.edge_cut
    for x, y in zip(a[:3], c):
        assert np.array_equal(x.labels, y.labels) and x.edge_cut == y.edge_cut


def test_batch_matches_oracle(ctx):
    specs = mixed_specs(seed=3, count=6)
    got, _ = run_batch(ctx, specs, 10000, 10, 0.03)
    for (spec, k), r in zip(specs, got):
        off, col = api.gen_host(spec)
        w1 = O.phase1(off, col, k, 10000)
        w2 = O.phase2(off, col, k, w1.labels, 10, 0.03)
        assert np.array_equal(r.labels, w2.labels)
        assert r.edge_cut == w2.m.edge_cut


def test_batch_tuning_does_not_change_results(ctx):
    specs = mixed_specs(seed=5, count=8)
    base, _ = run_batch(ctx, specs, 1000, 10, 0.03)
    old = ctx.tuning()
    try:
        ctx.set_tuning(heavy_degree=32, chunk_edges=64, batch_work_per_block=1024)
        tuned, _ = run_batch(ctx, specs, 1000, 10, 0.03)
    finally:
        ctx.set_tuning(heavy_degree=old.heavy_degree, chunk_edges=old.chunk_edges,
                       batch_work_per_block=old.batch_work_per_block)
    for x, y in zip(base, tuned):
        assert np.array_equal(x.labels, y.labels) and x.edge_cut == y.edge_cut


def test_batch_rejects_bad_jobs(ctx):
    g = ctx.generate(api.grid_spec(8, 8))
    try:
        with pytest.raises(api.DmolpError):
            api.batch_run(ctx, [api.BatchJob(g, 4, 0, 10, 0.03)])
        with pytest.raises(api.DmolpError):
            api.batch_run(ctx, [api.BatchJob(g, 4, 10, 10, 0.03), api.BatchJob(g, 4, 10, 10, 0.03)])
    finally:
        g.free()


def test_autotune_keeps_results(ctx):
    specs = mixed_specs(seed=9, count=8)
    base, _ = run_batch(ctx, specs, 1000, 10, 0.03)
    old = ctx.tuning()
    graphs = [ctx.generate(sp) for sp, _ in specs]
    try:
        jobs = [api.BatchJob(g, k, 1000, 10, 0.03) for g, (_, k) in zip(graphs, specs)]
        t = api.batch_autotune(ctx, jobs, max_trials=4)
        assert t.heavy_degree in (256, 1024, 4096)
        cur = ctx.tuning()
        assert cur.heavy_degree == t.heavy_degree and cur.batch_work_per_block == t.batch_work_per_block
        res, _ = api.batch_run(ctx, jobs, want_labels=True)
        for x, y in zip(base, res):
            assert np.array_equal(x.labels, y.labels) and x.edge_cut == y.edge_cut
    finally:
        for g in graphs:
            g.free()
        ctx.set_tuning(heavy_degree=old.heavy_degree, chunk_edges=old.chunk_edges,
                       batch_work_per_block=old.batch_work_per_block)


def test_tuning_rejects_out_of_range_values(ctx):
    # chunk_edges is bounded so a lane's 16-bit chunk counters cannot wrap;
    # heavy_degree so light rows' 16-bit kept counts cannot.
    old = ctx.tuning()
    for bad in (dict(chunk_edges=(1 << 21) + 1), dict(heavy_degree=65536), dict(light_hist=3)):
        with pytest.raises(api.DmolpError):
            ctx.set_tuning(**bad)
    now = ctx.tuning()
    assert (now.chunk_edges, now.heavy_degree) == (old.chunk_edges, old.heavy_degree)
    ctx.set_tuning(chunk_edges=1 << 21)
    assert ctx.tuning().chunk_edges == 1 << 21
    ctx.set_tuning(chunk_edges=old.chunk_edges)
