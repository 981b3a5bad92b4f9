"""dmolp_stats_noalloc (the statistics entry point that allocates no
device memory) against dmolp_stats, the entry point it is a variant of:
equal counts, cut and statuses on two calls in a row, the second on a
smaller graph (the shared scratch keeps the first call's contents), and a
clear error past its size limit."""
from __future__ import annotations

import numpy as np
import pytest

import graphs
from dmolp_amd import api

pytestmark = pytest.mark.gpu

NOALLOC_MAX_N = 1 << 24


def test_noalloc_stats_equal_stats_twice_second_smaller(ctx):
    cases = [(api.gen_host(api.rmat_spec(14, 16, seed=3)), 12),
             (graphs.grid(23, 31), 5)]
    rng = np.random.default_rng(21)
    for (off, col), k in cases:
        print(f"[noalloc] n={len(off) - 1} k={k}", flush=True)
        g = ctx.upload(off, col)
        try:
            labels = rng.integers(0, k, len(off) - 1).astype(np.int32)
            want = g.stats(k, labels)
            got = g.stats_noalloc(k, labels)
            assert np.array_equal(want[0], got[0]) and np.array_equal(want[1], got[1])
            assert want[2] == got[2]
            labels[3] = -1
            for fn in (g.stats, g.stats_noalloc):
                with pytest.raises(api.DmolpError) as e:
                    fn(k, labels)
                assert e.value.status == -3
        finally:
            g.free()


def test_noalloc_rejects_graphs_past_its_limit(ctx):
    g = ctx.generate(api.grid_spec(4097, 4096))
    try:
        assert g.n > NOALLOC_MAX_N
        with pytest.raises(api.DmolpError) as e:
            g.stats_noalloc(4, np.zeros(g.n, np.int32))
        assert e.value.status == -7
    finally:
        g.free()
