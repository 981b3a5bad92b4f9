"""Bit-exact parity at the measured workload sizes, against the oracle's
committed fixtures (tests/golden/large_*.json, written on the CPU by
tests/golden/make_large_golden.py).

  R-MAT scale 22 (k=16, 20 rounds) and the 41.6M-vertex Chung-Lu graph
  (k=32, 30 rounds): phase 1 and phase 2 in full, every label (SHA-256),
  the cut, part sizes, round counts, exit reason and per-round trace.
  The 256-graph batch and the served request mix: every job's labels and
  metrics, batched / served next to the oracle.
"""
from __future__ import annotations

import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from dmolp_amd import api, workloads as W

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).parent / "golden"


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def fixture(name):
    p = GOLD / f"large_{name}.json"
    if not p.exists():
        pytest.skip(f"{p.name} not written")
    return json.loads(p.read_text())


def check_plain(ctx, name):
    want = fixture(name)
    cfg = W.CONFIGS[name]
    k, R, eps = cfg["k"], cfg["rounds"], cfg["eps"]
    assert (k, R, eps) == (want["k"], want["max_rounds"], want["epsilon"])
    print(f"[large] {name}: generate", flush=True)
    g = ctx.generate(W.spec_for(cfg))
    try:
        assert (g.n, g.nnz) == (want["n"], want["nnz"])
        if "csr_sha256" in want:
            off, col = g.download()
            assert sha(off.astype(np.int64), col.astype(np.int32)) == want["csr_sha256"]
            del off, col
        print(f"[large] {name}: phase 1", flush=True)
        p1 = g.phase1(k, W.THETA)
        w1 = want["phase1"]
        assert p1.seeds.tolist() == w1["seeds"]
        assert p1.metrics.n_hubs == w1["n_hubs"]
        assert p1.metrics.rounds == w1["rounds"]
        assert p1.metrics.fill_round == w1["fill_round"]
        assert p1.metrics.n_seeds_fallback == w1["n_seeds_fallback"]
        assert p1.metrics.initial_edge_cut == w1["cut"]
        assert p1.vc.tolist() == w1["vc"] and p1.ec.tolist() == w1["ec"]
        assert p1.metrics.initial_vertex_balance == w1["vertex_balance"]
        assert p1.metrics.initial_edge_balance == w1["edge_balance"]
        assert sha(p1.labels) == w1["labels_sha256"]
        print(f"[large] {name}: phase 2", flush=True)
        p2 = g.phase2(k, R, eps, want_trace=True)
        w2 = want["phase2"]
        assert p2.metrics.initial_cut == w2["initial_cut"]
        assert p2.metrics.rounds == w2["rounds"]
        assert p2.metrics.exit_reason == w2["exit_reason"]
        assert p2.metrics.lp_edges == w2["lp_edges"]
        assert p2.metrics.edge_cut == w2["cut"]
        assert p2.vc.tolist() == w2["vc"] and p2.ic.tolist() == w2["ic"]
        assert p2.metrics.vertex_balance == w2["vertex_balance"]
        assert p2.metrics.edge_balance == w2["edge_balance"]
        for f, v in w2["trace"].items():
            got = p2.trace[f].tolist()
            if f == "cand_count":  # -1 in rounds that pull candidates from the boundary side
                got = [w if x < 0 else x for x, w in zip(got, v)]
            assert got == v, f
        assert sha(p2.labels) == w2["labels_sha256"]
        # The incremental counts equal a from-scratch sweep of the final labels.
        vc, ic, cut = g.stats(k, p2.labels)
        assert cut == w2["cut"] and vc.tolist() == w2["vc"] and ic.tolist() == w2["ic"]
    finally:
        g.free()


def test_rmat_s22_full_run_matches_oracle(ctx):
    check_plain(ctx, "c2")


def test_chunglu_41m_full_run_matches_oracle(ctx):
    check_plain(ctx, "c3")


def _check_job(res_labels, res, w, tag):
    got = dict(status=res.status, phase1_cut=res.phase1_cut, n_seeds_fallback=res.n_seeds_fallback,
               cut=res.edge_cut, rounds=res.rounds, exit_reason=res.exit_reason,
               lp_edges=res.lp_edges, vertex_balance=res.vertex_balance,
               edge_balance=res.edge_balance, labels_sha256=sha(res_labels))
    exp = dict(status=0, **{f: w[f] for f in got if f != "status"})
    assert got == exp, tag


def test_batch256_every_job_matches_oracle(ctx):
    want = fixture("c4")
    specs = W.c4_specs()
    assert len(specs) == len(want["jobs"])
    graphs = [ctx.generate(sp) for sp, _ in specs]
    try:
        for (sp, k), w, g in zip(specs, want["jobs"], graphs):
            assert W.spec_key(sp) == w["generator"] and k == w["k"]
            assert (g.n, g.nnz) == (w["n"], w["nnz"])
        jobs = [api.BatchJob(g, k, want["theta"], want["max_rounds"], want["epsilon"])
                for g, (_, k) in zip(graphs, specs)]
        print("[large] c4: batch of 256", flush=True)
        res, _ = api.batch_run(ctx, jobs, want_labels=True)
        for i, (r, w) in enumerate(zip(res, want["jobs"])):
            _check_job(r.labels, r, w, i)
    finally:
        for g in graphs:
            g.free()


def test_served_mix_every_job_matches_oracle(ctx):
    """The served request mix (vertices 1e3-1e7, edge factor 8-32, k up to
    128), arriving while others run: every result equals the oracle's."""
    want = fixture("c5")
    jobs = W.c5_jobs(len(want["jobs"]))
    srv = api.Server(ctx, 160 << 30, 32)
    try:
        rng = np.random.default_rng(55)
        pending = list(range(len(jobs)))
        tickets = {}
        done = {}
        print(f"[large] c5: {len(jobs)} served jobs", flush=True)
        while len(done) < len(jobs):
            # Admit a few requests per step, so later ones arrive while
            # earlier ones are in every stage.
            for _ in range(int(rng.integers(0, 4))):
                if pending:
                    i = pending.pop(0)
                    sp, k = jobs[i]
                    tickets[i] = srv.submit(k, want["theta"], want["max_rounds"], want["epsilon"],
                                            spec=sp, want_labels=True)
            srv.step()
            for i, t in tickets.items():
                if i not in done:
                    r = srv.take(t)
                    if r is not None:
                        done[i] = r
        for i, w in enumerate(want["jobs"]):
            served, labels = done[i]
            assert (served.n, served.nnz) == (w["n"], w["nnz"]), i
            _check_job(labels, served.result, w, i)
    finally:
        srv.close()
