"""Committed end-to-end fixtures (tests/golden/partitions.json, written by
tests/golden/make_golden.py from the oracle): the oracle and the host
generator must keep reproducing them on the CPU, and the device path must
reproduce them bit for bit on the GPU."""
from __future__ import annotations

import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

import oracle_py as O
from golden.make_golden import CASES, spec_of

GOLD = {c["name"]: c for c in json.loads((Path(__file__).parent / "golden" / "partitions.json").read_text())}
PARAMS = [pytest.param(c, id=c[0]) for c in CASES]


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_fixture_covers_every_case():
    assert set(GOLD) == {c[0] for c in CASES}


@pytest.mark.parametrize("case", PARAMS)
def test_oracle_reproduces_fixture(case):
    from dmolp_amd import api
    name, gen, k, theta, rounds, eps = case
    want = GOLD[name]
    off, col = api.gen_host(spec_of(gen))
    assert sha(off.astype(np.int64), col.astype(np.int32)) == want["csr_sha256"]
    p1 = O.phase1(off, col, k, theta)
    assert sha(p1.labels) == want["phase1"]["labels_sha256"]
    assert p1.m.initial_edge_cut == want["phase1"]["cut"]
    p2 = O.phase2(off, col, k, p1.labels, rounds, eps)
    assert sha(p2.labels) == want["phase2"]["labels_sha256"]
    assert p2.m.edge_cut == want["phase2"]["cut"]
    assert p2.vc.tolist() == want["phase2"]["vc"] and p2.ic.tolist() == want["phase2"]["ic"]


@pytest.mark.gpu
@pytest.mark.parametrize("case", PARAMS)
def test_device_reproduces_fixture(ctx, case):
    from dmolp_amd import api
    name, gen, k, theta, rounds, eps = case
    want = GOLD[name]
    print(f"[golden] {name}", flush=True)
    g = ctx.generate(spec_of(gen))
    try:
        off, col = g.download()
        assert sha(off.astype(np.int64), col.astype(np.int32)) == want["csr_sha256"]
        p1 = g.phase1(k, theta)
        w1 = want["phase1"]
        assert sha(p1.labels) == w1["labels_sha256"]
        assert p1.metrics.initial_edge_cut == w1["cut"]
        assert p1.metrics.rounds == w1["rounds"]
        assert p1.metrics.n_seeds_fallback == w1["n_seeds_fallback"]
        assert p1.seeds.tolist() == w1["seeds"]
        p2 = g.phase2(k, rounds, eps, want_trace=True)
        w2 = want["phase2"]
        assert sha(p2.labels) == w2["labels_sha256"]
        assert p2.metrics.edge_cut == w2["cut"]
        assert p2.metrics.rounds == w2["rounds"]
        assert p2.metrics.exit_reason == w2["exit_reason"]
        assert p2.metrics.lp_edges == w2["lp_edges"]
        assert p2.vc.tolist() == w2["vc"] and p2.ic.tolist() == w2["ic"]
        for f, v in w2["trace"].items():
            assert p2.trace[f][:len(v)].tolist() == v, f
    finally:
        g.free()
