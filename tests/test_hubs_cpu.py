"""The library's host hub search (find_hub_nodes, init.cpp:139-177), which
sorts only the top of the (id, degree) order, against the oracle's
ref_hubs, which calls std::sort on the whole sequence: the same hubs in the
same order (ties included) on degree sequences built to stress the
introsort: many equal degrees, sorted and reversed runs, organ pipes,
saw-tooth, power laws, a constant sequence (every vertex a hub) and sizes
around the insertion-sort threshold.  No GPU."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_py as O
from dmolp_amd import api


def oracle_hubs(deg):
    off = np.zeros(len(deg) + 1, np.int64)
    off[1:] = np.cumsum(deg)
    return O.hubs(off)


def sequences():
    rng = np.random.default_rng(123)
    out = []
    for n in (2, 3, 15, 16, 17, 31, 33, 100, 1000, 5000):
        out.append(("uniform-small", rng.integers(0, 4, n)))
        out.append(("uniform-wide", rng.integers(0, 1000, n)))
    for n in (10_000, 100_000, 400_000):
        pl = np.floor(2000.0 / (np.arange(n) + 1.0) ** 0.8).astype(np.int64)
        out.append(("power-law", pl))
        out.append(("power-law-shuffled", rng.permutation(pl)))
        out.append(("zipf", np.minimum(rng.zipf(1.7, n), 100_000)))
        out.append(("few-values", rng.choice([1, 2, 3, 50, 50, 50, 400], n)))
        out.append(("ascending", np.arange(n) // 7))
        out.append(("descending", (n - np.arange(n)) // 7))
        half = np.arange(n // 2)
        out.append(("organ-pipe", np.concatenate([half, half[::-1]]) % 997))
        out.append(("saw-tooth", np.arange(n) % 64))
        top = rng.integers(0, 10, n)
        top[rng.choice(n, 40, replace=False)] = 5000  # a tie among the hubs
        out.append(("tied-hubs", top))
    out.append(("constant", np.full(5000, 7)))
    out.append(("zeros", np.zeros(100, np.int64)))
    return out


@pytest.mark.parametrize("name,deg", sequences(), ids=lambda x: x if isinstance(x, str) else "")
def test_host_hubs_match_full_sort(name, deg):
    deg = np.asarray(deg, np.int32)
    want = oracle_hubs(deg)
    got = api.hubs_host(deg)
    assert np.array_equal(want, got), (name, len(want), len(got))


def long_sequences():
    """Above 2^20 entries the long partition steps run on several threads
    (the test process's host threads): the result must not change."""
    rng = np.random.default_rng(77)
    n = 1_300_000
    pl = np.floor(3000.0 / (np.arange(n) + 1.0) ** 0.7).astype(np.int64)
    return [
        ("power-law-ids", pl),  # hubs at low ids: the descending-id order ends with them
        ("power-law-shuffled", rng.permutation(pl)),
        ("zipf", np.minimum(rng.zipf(1.6, 2_400_000), 1_000_000)),
        ("few-values", rng.choice([1, 2, 2, 3, 40, 40, 900], 1_700_000)),
        ("descending", (1_100_000 - np.arange(1_100_000)) // 3),
        ("organ-pipe", np.concatenate([np.arange(700_000), np.arange(700_000)[::-1]]) % 5003),
    ]


@pytest.mark.parametrize("name,deg", long_sequences(), ids=lambda x: x if isinstance(x, str) else "")
def test_host_hubs_long_sequences(name, deg):
    deg = np.asarray(deg, np.int32)
    want = oracle_hubs(deg)
    got = api.hubs_host(deg)
    assert len(want) > 0
    assert np.array_equal(want, got), (name, len(want), len(got))


def test_host_hubs_random_sweep():
    rng = np.random.default_rng(2024)
    for trial in range(300):
        n = int(rng.integers(2, 3000))
        kind = trial % 4
        if kind == 0:
            deg = rng.integers(0, int(rng.integers(1, 20)), n)
        elif kind == 1:
            deg = np.minimum(rng.zipf(1.5 + rng.random(), n), 10 ** 6)
        elif kind == 2:
            deg = np.sort(rng.integers(0, 50, n))[::int(rng.choice([1, -1]))]
        else:
            deg = rng.integers(0, 5, n)
            deg[rng.integers(0, n, 5)] = int(rng.integers(10, 100))
        deg = np.ascontiguousarray(deg, np.int32)
        assert np.array_equal(oracle_hubs(deg), api.hubs_host(deg)), trial


def oracle_first_seed(deg):
    off = np.zeros(len(deg) + 1, np.int64)
    off[1:] = np.cumsum(deg)
    return O.first_seed(off)


def first_seed_sequences():
    """Degree sequences for find_first_seed (init.cpp:179-187): one maximum;
    several maxima with ids below and above the maximum degree (the tie test
    compares an id with the running best degree); maxima only above it;
    long sequences that the library splits over threads."""
    rng = np.random.default_rng(7)
    out = [("one", np.array([3])), ("pair", np.array([2, 2])), ("zeros", np.zeros(9, np.int64))]
    for n in (5, 40, 1000):
        d = rng.integers(0, 10, n)
        d[n // 2] = 50
        out.append(("single-max", d))
        d = rng.integers(0, 30, n)
        d[[1, n // 3, n - 2]] = 30  # ids below 30 and above
        out.append(("ties-below-and-above", d))
        d = rng.integers(0, 5, n)
        d[[n - 1, n - 3]] = 5  # all maxima at ids >= 5 (n >= 40 case)
        out.append(("ties-above", d))
    for n in (1 << 20, (1 << 21) + 17):
        d = rng.integers(0, 100, n)
        out.append(("many-ties", d))
        d = np.minimum(rng.zipf(1.9, n), 5000)
        d[rng.integers(0, n, 7)] = 5000
        out.append(("zipf-capped", d))
        d = rng.integers(0, 100, n)
        d[[n - 1, n - 5]] = 100000  # maxima only at large ids
        out.append(("big-max-high-ids", d))
    return out


@pytest.mark.parametrize("name,deg", first_seed_sequences(),
                         ids=[f"{nm}-{len(d)}" for nm, d in first_seed_sequences()])
def test_first_seed_equals_oracle_scan(name, deg):
    assert api.first_seed_host(deg) == oracle_first_seed(deg)
