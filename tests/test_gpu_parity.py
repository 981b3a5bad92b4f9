"""Parity of the HIP path with the CPU oracle (bit-exact), through the C ABI."""
from __future__ import annotations

import numpy as np
import pytest

import graphs
import oracle_py as O
from dmolp_amd import api

pytestmark = pytest.mark.gpu

TINY = {
    "path9": graphs.path(9),
    "star12": graphs.star(12),
    "grid5x7": graphs.grid(5, 7),
    "cliques6": graphs.two_cliques(6),
    "rand300": graphs.random_graph(300, 900, 7),
    "plaw500": graphs.powerlaw(500, 3000, 11),
}


def c1():
    return api.gen_host(api.rmat_spec(16, 16, seed=1))


DEFAULT_TUNING = dict(heavy_degree=1024, chunk_edges=4096, pull_per_mille=50,
                      hist_recount_per_mille=200, hist_push_per_mille=20, light_hist=1)
TUNINGS = {
    "default": DEFAULT_TUNING,
    # Tiny heavy threshold: the chunked heavy-row kernels run on small graphs.
    "heavy": dict(DEFAULT_TUNING, heavy_degree=5, chunk_edges=64),
    # Candidates always pushed from B_{t-1} / always pulled from the boundary.
    "push": dict(DEFAULT_TUNING, pull_per_mille=1000000),
    "pull": dict(DEFAULT_TUNING, heavy_degree=5, chunk_edges=64, pull_per_mille=1),
    # Heavy-row histograms always followed by one mode: bitmap scan, push of
    # the changes, or recount (1001 per mille: never).
    "hist_scan": dict(DEFAULT_TUNING, heavy_degree=5, chunk_edges=64,
                      hist_recount_per_mille=1001, hist_push_per_mille=1),
    "hist_push": dict(DEFAULT_TUNING, heavy_degree=5, chunk_edges=64,
                      hist_recount_per_mille=1001, hist_push_per_mille=1000),
    "hist_recount": dict(DEFAULT_TUNING, heavy_degree=5, chunk_edges=64,
                         hist_recount_per_mille=1, hist_push_per_mille=1),
    # Light rows gather their neighbours' labels every round instead of
    # keeping histograms (alone, and with heavy histograms pushed).
    "gather": dict(DEFAULT_TUNING, light_hist=2),
    "gather_push": dict(DEFAULT_TUNING, heavy_degree=5, chunk_edges=64, light_hist=2,
                        hist_recount_per_mille=1001, hist_push_per_mille=1000),
}


@pytest.fixture(params=list(TUNINGS))
def tuned(ctx, request):
    ctx.set_tuning(**TUNINGS[request.param])
    yield ctx
    ctx.set_tuning(**DEFAULT_TUNING)


@pytest.mark.parametrize("spec", [
    api.rmat_spec(12, 16, seed=5),
    api.rmat_spec(10, 8, a=0.45, b=0.22, c=0.22, seed=9),
    api.chung_lu_spec(5000, 40000, gamma=2.0, dmax=800.0, seed=3),
    api.grid_spec(13, 17),
], ids=["rmat12", "rmat10", "chunglu", "grid"])
def test_device_generator_matches_host(ctx, spec):
    off_h, col_h = api.gen_host(spec)
    g = ctx.generate(spec)
    off_d, col_d = g.download()
    assert np.array_equal(off_h, off_d)
    assert np.array_equal(col_h, col_d)


@pytest.mark.parametrize("name", list(TINY))
def test_stats_parity(tuned, name):
    off, col = TINY[name]
    g = tuned.upload(off, col)
    rng = np.random.default_rng(1)
    # k <= 16 and 17..32 take the packed label mirror (4 and 5 bits), 40-254
    # the bytes (labels from 32 on have no place in the kept rows), 255 and
    # 300 the 32-bit labels.
    for k in (1, 3, 8, 16, 20, 32, 40, 64, 128, 254, 255, 300):
        labels = rng.integers(0, k, len(off) - 1).astype(np.int32)
        rc, vc, ic, cut = O.stats(off, col, k, labels)
        assert rc == 0
        dvc, dic, dcut = g.stats(k, labels)
        assert np.array_equal(vc, dvc) and np.array_equal(ic, dic) and cut == dcut


@pytest.mark.parametrize("name", ["grid5x7", "plaw500"])
def test_stats_invalid_label_status(ctx, name):
    # A -1 label that reaches the cut fails the sweep with the oracle's
    # status; the packed mirror is skipped when any label is invalid.
    off, col = TINY[name]
    g = ctx.upload(off, col)
    rng = np.random.default_rng(3)
    for k in (4, 20, 40):
        labels = rng.integers(0, k, len(off) - 1).astype(np.int32)
        labels[int(col[0])] = -1
        rc = O.stats(off, col, k, labels)[0]
        assert rc == -3
        with pytest.raises(api.DmolpError) as e:
            g.stats(k, labels)
        assert e.value.status == rc


def test_penalties_bit_exact(ctx):
    rng = np.random.default_rng(3)
    for trial in range(200):
        k = int(rng.integers(1, 130))
        vc = rng.integers(0, 10 ** int(rng.integers(1, 9)), k)
        ic = rng.integers(0, 10 ** int(rng.integers(1, 9)), k)
        if trial % 7 == 0:
            ic[:] = 0
        t = int(rng.integers(1, 40))
        pen, re = O.penalties(k, vc, ic, t)
        dpen, dre = ctx.penalties(k, vc, ic, t)
        assert np.array_equal(pen.view(np.uint64), dpen.view(np.uint64)), trial
        assert np.array_equal(re.view(np.uint64), dre.view(np.uint64)), trial


@pytest.mark.parametrize("name", list(TINY))
def test_lp_round_parity(tuned, name):
    off, col = TINY[name]
    n = len(off) - 1
    g = tuned.upload(off, col)
    rng = np.random.default_rng(5)
    for k in (2, 5, 32, 37):
        for trial in range(3):
            labels = rng.integers(0, k, n).astype(np.int32)
            b = np.sort(rng.choice(n, size=max(1, n // 2), replace=False)).astype(np.int32)
            pen = rng.normal(0, 0.5, k)
            if trial == 2:
                pen[:] = 0.0
            want = O.lp_round(off, col, labels, b, pen, k)
            got = g.lp_round(labels, b, pen, k)
            assert np.array_equal(want, got), (k, trial)


@pytest.mark.parametrize("name", list(TINY))
def test_select_boundary_parity(tuned, name):
    off, col = TINY[name]
    n = len(off) - 1
    g = tuned.upload(off, col)
    rng = np.random.default_rng(9)
    for k in (2, 4, 9):
        for t in (1, 2, 5, 13):
            labels = rng.integers(0, k, n).astype(np.int32)
            prev = np.sort(rng.choice(n, size=max(1, n // 3), replace=False)).astype(np.int32)
            re = rng.uniform(0.5, 1.5, k)
            want = O.select_boundary(off, col, labels, prev, re, k, t)
            got = g.select_boundary(labels, prev, re, k, t)
            assert np.array_equal(want, got), (k, t)


def _check_phase2(g, off, col, k, labels0, rounds, eps):
    want = O.phase2(off, col, k, labels0, rounds, eps)
    g.set_labels(labels0)
    got = g.phase2(k, rounds, eps)
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
    # Candidates are counted only in rounds that push from B_{t-1} (-1 when
    # they are pulled from the boundary side).
    pushed = got.trace["cand_count"] >= 0
    assert np.array_equal(want.trace["cand_count"][pushed], got.trace["cand_count"][pushed])
    return want


@pytest.mark.parametrize("name", list(TINY))
def test_phase2_parity_tiny(tuned, name):
    off, col = TINY[name]
    g = tuned.upload(off, col)
    rng = np.random.default_rng(13)
    for k in (2, 3, 8):
        labels0 = rng.integers(0, k, len(off) - 1).astype(np.int32)
        _check_phase2(g, off, col, k, labels0, 12, 0.03)


@pytest.mark.parametrize("name", list(TINY))
def test_phase1_parity_tiny(tuned, name):
    off, col = TINY[name]
    g = tuned.upload(off, col)
    n = len(off) - 1
    for k in (1, 2, 4):
        for theta in (1, 2, 10000):
            if k > n:
                continue
            want = O.phase1(off, col, k, theta)
            got = g.phase1(k, theta)
            assert np.array_equal(want.seeds, got.seeds), (k, theta)
            assert np.array_equal(want.labels, got.labels), (k, theta)
            assert want.m.rounds == got.metrics.rounds
            assert want.m.fill_round == got.metrics.fill_round
            assert want.m.initial_edge_cut == got.metrics.initial_edge_cut
            assert np.array_equal(want.vc, got.vc) and np.array_equal(want.ec, got.ec)
            assert want.m.n_seeds_found == got.metrics.n_seeds_found
            assert want.m.n_hubs == got.metrics.n_hubs


def test_config1_end_to_end(tuned):
    """R-MAT scale 16, edge factor 16, seed 1; k=8, theta=10000, 10 rounds."""
    off, col = c1()
    g = tuned.upload(off, col)
    want1 = O.phase1(off, col, 8, 10000)
    got1 = g.phase1(8, 10000)
    assert np.array_equal(want1.labels, got1.labels)
    assert np.array_equal(want1.seeds, got1.seeds)
    assert want1.m.initial_edge_cut == got1.metrics.initial_edge_cut
    assert want1.m.initial_vertex_balance == got1.metrics.initial_vertex_balance
    assert want1.m.initial_edge_balance == got1.metrics.initial_edge_balance
    _check_phase2(g, off, col, 8, want1.labels, 10, 0.03)


def test_graph_replays_match_eager(ctx):
    """Phase 2 normally replays one captured round as a HIP graph; with
    profiling on it launches the kernels directly.  Several graph runs, with
    eager runs between them, must all equal the eager result and the oracle."""
    off, col = api.gen_host(api.rmat_spec(14, 16, seed=4))
    g = ctx.upload(off, col)
    p1 = g.phase1(8, 100)
    g.checkpoint_labels()
    want = O.phase2(off, col, 8, p1.labels, 15, 0.01)
    runs = []
    for profiled in (True, False, False, True, False):
        ctx.set_profiling(profiled)
        g.checkpoint_labels(restore=True)
        r = g.phase2(8, 15, 0.01)
        runs.append(r)
    ctx.set_profiling(False)
    for r in runs:
        assert np.array_equal(r.labels, want.labels)
        assert r.metrics.edge_cut == want.m.edge_cut and r.metrics.rounds == want.m.rounds
        assert np.array_equal(r.trace["b_count"], want.trace["b_count"])


@pytest.mark.parametrize("k", [33, 40, 254, 255, 300, 1024])
def test_phase2_parity_large_k(tuned, k):
    """k above the LP's 32-label cap (labels >= 32 always move) and above
    the byte mirror's exact range (k > 254: sweeps use the int32 labels)."""
    off, col = api.gen_host(api.rmat_spec(12, 8, seed=21))
    g = tuned.upload(off, col)
    rng = np.random.default_rng(k)
    labels0 = rng.integers(0, k, len(off) - 1).astype(np.int32)
    _check_phase2(g, off, col, k, labels0, 8, 0.01)


@pytest.mark.parametrize("k", [40, 300])
def test_phase1_then_phase2_large_k(tuned, k):
    off, col = api.gen_host(api.rmat_spec(11, 16, seed=22))
    g = tuned.upload(off, col)
    want1 = O.phase1(off, col, k, 500)
    got1 = g.phase1(k, 500)
    assert np.array_equal(want1.labels, got1.labels)
    _check_phase2(g, off, col, k, want1.labels, 10, 0.03)


def test_trace_light_parts_and_profile_groups(ctx):
    """The light parts of Delta_t and B_t in the trace, checked against the
    labels themselves, and the shape of the per-round kernel-group times."""
    ctx.set_tuning(**dict(DEFAULT_TUNING, heavy_degree=40))
    try:
        off, col = api.gen_host(api.rmat_spec(13, 16, seed=7))
        deg = np.diff(off)
        g = ctx.upload(off, col)
        labels0 = g.phase1(8, 1000).labels.copy()
        g.checkpoint_labels()
        r = g.phase2(8, 1, 0.01)
        changed = np.flatnonzero(r.labels != labels0)
        light = changed[deg[changed] <= 40]
        assert r.trace["chg_count"][0] == len(changed)
        assert r.trace["chg_light_count"][0] == len(light)
        assert r.trace["chg_light_deg"][0] == int(deg[light].sum())
        ctx.set_profiling(True)
        g.checkpoint_labels(restore=True)
        r = g.phase2(8, 6, 0.01)
        ctx.set_profiling(False)
        rounds = r.metrics.rounds
        for group in range(8):
            t = ctx.round_times(group)
            assert len(t) == rounds and np.all(t >= 0), group
        assert len(ctx.round_times(8)) == 1  # the initial light sweep, once per run
        # The light Delta kernel is part of the Delta + apply group.
        assert np.all(ctx.round_times(7) <= ctx.round_times(4) + 1e-3)
        assert np.all(r.trace["b_light_count"] <= r.trace["b_count"])
        assert np.all(r.trace["chg_light_count"] <= r.trace["chg_count"])
        assert np.all(r.trace["chg_light_deg"] <= r.trace["chg_deg"])
    finally:
        ctx.set_profiling(False)
        ctx.set_tuning(**DEFAULT_TUNING)


# Graphs whose rows reach the selection's last level-1 bin (degree 255 and
# up), so label cuts fall there too and take the digit passes, next to cuts
# inside one exact degree (ranked by id chunks); k = 70 adds labels beyond
# the first 64, which always take the digit passes.
HIGH_DEGREE = {
    "dense700": graphs.random_graph(700, 700 * 150, 5),
    "plaw3000": graphs.powerlaw(3000, 60000, 7),
}


@pytest.mark.parametrize("name", list(HIGH_DEGREE))
def test_select_boundary_parity_high_degree(ctx, name):
    off, col = HIGH_DEGREE[name]
    n = len(off) - 1
    assert int(np.diff(off).max()) >= 255
    g = ctx.upload(off, col)
    rng = np.random.default_rng(17)
    for k in (2, 4, 9, 70):
        for t in (1, 2, 5, 13):
            labels = rng.integers(0, k, n).astype(np.int32)
            prev = np.sort(rng.choice(n, size=max(1, n // 2), replace=False)).astype(np.int32)
            re = rng.uniform(0.5, 1.5, k)
            want = O.select_boundary(off, col, labels, prev, re, k, t)
            got = g.select_boundary(labels, prev, re, k, t)
            assert np.array_equal(want, got), (k, t)


@pytest.mark.parametrize("name", list(HIGH_DEGREE))
def test_phase2_parity_high_degree(ctx, name):
    off, col = HIGH_DEGREE[name]
    g = ctx.upload(off, col)
    rng = np.random.default_rng(19)
    for k in (4, 9, 70):
        labels0 = rng.integers(0, k, len(off) - 1).astype(np.int32)
        _check_phase2(g, off, col, k, labels0, 15, 0.0)


def test_phase2_byte_then_word_labels_same_graph(tuned):
    # Runs on one graph keep their work arrays: a run with byte labels
    # (k <= 254) followed by runs with word labels (k > 254) and back.
    off, col = api.gen_host(api.rmat_spec(12, 8, seed=23))
    g = tuned.upload(off, col)
    rng = np.random.default_rng(29)
    for k in (16, 300, 32, 1024):
        labels0 = rng.integers(0, k, len(off) - 1).astype(np.int32)
        _check_phase2(g, off, col, k, labels0, 8, 0.01)


def test_probe_atomics_counts_and_rejects(ctx):
    """The atomic-rate probe reports its adds per launch (blocks x 256 lanes x
    reps) and a positive time, and rejects shares that do not divide 64."""
    ms, adds = ctx.probe_atomics(1 << 16, 1, reps=4)
    assert ms > 0 and adds > 0 and adds % (256 * 4) == 0
    ms2, adds2 = ctx.probe_atomics(1 << 16, 8, reps=4)
    assert ms2 > 0 and adds2 == adds
    with pytest.raises(api.DmolpError):
        ctx.probe_atomics(1 << 16, 3, reps=4)
