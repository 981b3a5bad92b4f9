"""The hub search with its partition steps on the device (dmolp_hubs_device,
and dmolp_phase1 for graphs of at least the device threshold) against the
oracle's std::sort order (ref_hubs, init.cpp:139-177) and find_first_seed
(init.cpp:179-187): the same hubs in the same order, ties included, the
same hub count and first seed.  host_max = 17 keeps every range longer than
16 entries on the device (so the device does nearly every partition step,
many ranges per launch when every vertex is a hub); larger values hand
the short ranges to the host as dmolp_phase1 does.  Then phase 1, the
batched mode and the served mode with the device search forced on small
graphs, against the oracle and the host search."""
from __future__ import annotations

import numpy as np
import pytest

import graphs
import oracle_py as O
from dmolp_amd import api
from test_hubs_cpu import first_seed_sequences, long_sequences, sequences

pytestmark = pytest.mark.gpu


def oracle(deg):
    off = np.zeros(len(deg) + 1, np.int64)
    off[1:] = np.cumsum(deg)
    return O.hubs(off), O.first_seed(off)


@pytest.fixture
def dev(ctx):
    yield ctx
    ctx.set_hub_device(1 << 18, 1 << 15)
    ctx.set_hub_prefix(1 << 16)


SEQ = [(nm, d) for nm, d in sequences() if len(d) >= 2]


@pytest.mark.parametrize("host_max", [17, 1000, 1 << 15])
@pytest.mark.parametrize("name,deg", SEQ, ids=[f"{nm}-{len(d)}" for nm, d in SEQ])
def test_device_hubs_equal_oracle(ctx, name, deg, host_max):
    deg = np.asarray(deg, np.int32)
    want, fs = oracle(deg)
    hubs, nh, seed, d = ctx.hubs_device(deg, -1, host_max)
    assert nh == len(want), name
    assert np.array_equal(hubs, want), name
    assert (seed, d) == tuple(fs), name


@pytest.mark.parametrize("name,deg", long_sequences(), ids=[nm for nm, _ in long_sequences()])
def test_device_hubs_long_sequences(ctx, name, deg):
    deg = np.asarray(deg, np.int32)
    want = api.hubs_host(deg)  # equal to the oracle's (test_hubs_cpu)
    for host_max in (17, 1 << 15):
        hubs, nh, seed, d = ctx.hubs_device(deg, -1, host_max)
        assert nh == len(want)
        assert np.array_equal(hubs, want), (name, host_max)
        assert (seed, d) == api.first_seed_host(deg)


@pytest.mark.parametrize("limit", [0, 1, 7, 100, 4000])
def test_device_hub_prefix(ctx, limit):
    """Only the first min(limit, M) hubs: the top of the full order."""
    rng = np.random.default_rng(limit + 3)
    for deg in (np.full(5000, 7), rng.integers(0, 4, 3000), np.minimum(rng.zipf(1.7, 200_000), 10**5)):
        deg = np.asarray(deg, np.int32)
        want, _ = oracle(deg)
        for host_max in (17, 1 << 15):
            hubs, nh, _, _ = ctx.hubs_device(deg, limit, host_max)
            assert nh == len(want)
            assert np.array_equal(hubs, want[:limit]), (limit, host_max)


def test_device_first_seed(ctx):
    for name, deg in first_seed_sequences():
        deg = np.asarray(deg, np.int32)
        if len(deg) < 2:
            continue
        _, _, seed, d = ctx.hubs_device(deg, 0, 1 << 15)
        assert (seed, d) == api.first_seed_host(deg), name


def test_device_hubs_random_sweep(ctx):
    rng = np.random.default_rng(99)
    for trial in range(60):
        n = int(rng.integers(2, 20000))
        kind = trial % 3
        if kind == 0:
            deg = rng.integers(0, int(rng.integers(1, 20)), n)
        elif kind == 1:
            deg = np.minimum(rng.zipf(1.5 + rng.random(), n), 10**6)
        else:
            deg = np.sort(rng.integers(0, 50, n))[::int(rng.choice([1, -1]))]
        deg = np.ascontiguousarray(deg, np.int32)
        want, fs = oracle(deg)
        hubs, nh, seed, d = ctx.hubs_device(deg, -1, int(rng.choice([17, 64, 1000])))
        assert np.array_equal(hubs, want), trial
        assert nh == len(want) and (seed, d) == tuple(fs), trial


GRAPHS = {
    "grid30x30": graphs.grid(30, 30),
    "star12": graphs.star(12),
    "rand300": graphs.random_graph(300, 900, 7),
    "plaw500": graphs.powerlaw(500, 3000, 11),
}


@pytest.mark.parametrize("name", list(GRAPHS))
@pytest.mark.parametrize("prefix", [1, 0, 1 << 16])
def test_phase1_with_device_hubs_equals_oracle(dev, name, prefix):
    off, col = GRAPHS[name]
    n = len(off) - 1
    dev.set_hub_device(0, 17)
    dev.set_hub_prefix(prefix)
    g = dev.upload(off, col)
    try:
        for k in (2, 8, 16):
            if k > n:
                continue
            want = O.phase1(off, col, k, 10000)
            got = g.phase1(k, 10000)
            assert np.array_equal(want.seeds, got.seeds), k
            assert np.array_equal(want.labels, got.labels), k
            assert want.m.n_hubs == got.metrics.n_hubs
            assert want.m.first_seed == got.metrics.first_seed
            assert want.m.n_seeds_found == got.metrics.n_seeds_found
            assert want.m.n_seeds_fallback == got.metrics.n_seeds_fallback
    finally:
        g.free()


def test_batch_with_device_hubs_equals_oracle(dev):
    dev.set_hub_device(0, 17)
    jobs, want = [], []
    for name, k in (("grid30x30", 8), ("rand300", 16), ("plaw500", 4), ("star12", 2)):
        off, col = GRAPHS[name]
        g = dev.upload(off, col)
        jobs.append(api.BatchJob(g, k, 10000, 3, 0.03))
        o1 = O.phase1(off, col, k, 10000)
        want.append(O.phase2(off, col, k, o1.labels, 3, 0.03))
    try:
        res, _ = api.batch_run(dev, jobs, want_labels=True)
        for r, w in zip(res, want):
            assert r.status == 0
            assert np.array_equal(r.labels, w.labels)
            assert r.edge_cut == w.m.edge_cut
    finally:
        for j in jobs:
            j.graph.free()


def test_served_with_device_hubs_equals_host_hubs(dev):
    rng = np.random.default_rng(21)
    specs = [(api.rmat_spec(int(rng.integers(8, 13)), int(rng.integers(8, 33)),
                            seed=int(rng.integers(1, 10**6))),
              int(rng.choice([4, 16, 128]))) for _ in range(6)]
    want = []
    for sp, k in specs:  # host searches (default threshold: these graphs are short)
        g = dev.generate(sp)
        try:
            g.phase1(k, 10000, want_labels=False)
            want.append(g.phase2(k, 20, 0.03))
        finally:
            g.free()
    dev.set_hub_device(0, 17)
    srv = api.Server(dev, 8 << 30, 4)
    try:
        tickets = [srv.submit(k, 10000, 20, 0.03, spec=sp, want_labels=True) for sp, k in specs]
        got = {}
        for _ in range(20000):
            srv.step()
            for t in tickets:
                if t not in got:
                    r = srv.take(t)
                    if r is not None:
                        got[t] = r
            if len(got) == len(tickets):
                break
    finally:
        srv.close()
    assert len(got) == len(tickets)
    for w, t in zip(want, tickets):
        served, labels = got[t]
        assert served.result.status == 0
        assert np.array_equal(labels, w.labels)
        assert served.result.edge_cut == w.metrics.edge_cut


@pytest.mark.parametrize("host_max", [17, 1000, 1 << 15])
def test_device_top_by_degree_equals_sorted_order(ctx, host_max):
    """The seed fallback's candidates: the first m vertices by (degree
    descending, id ascending), against numpy's lexicographic sort, on
    sequences with many ties at the cut."""
    rng = np.random.default_rng(host_max)
    seqs = [np.full(5000, 7), rng.integers(0, 4, 30000), np.minimum(rng.zipf(1.7, 300_000), 10**5),
            rng.integers(0, 3, 1_200_000), (np.arange(100_000) % 97)[::-1].copy()]
    for deg in seqs:
        deg = np.asarray(deg, np.int32)
        order = np.lexsort((np.arange(len(deg)), -deg.astype(np.int64)))
        for m in (1, 2, 33, 256, 2048):
            got = ctx.top_degree_device(deg, m, host_max)
            assert np.array_equal(got, order[:m]), (len(deg), m)
