"""Write the oracle's results at the measured workload sizes as fixtures:

  tests/golden/large_c2.json  R-MAT scale 22 (k=16, 20 rounds): phase 1 and
                              phase 2 in full, with the per-round trace
  tests/golden/large_c3.json  the 41.6M-vertex Chung-Lu graph (k=32, 30
                              rounds): phase 1 and phase 2 in full
  tests/golden/large_c4.json  every job of the 256-graph batch (phase 1 +
                              10 rounds each)
  tests/golden/large_c5.json  the first 96 served requests (phase 1 + 20
                              rounds each)

Each label array is stored as its SHA-256 (int32, native order), next to
the cut, part sizes, round counts and exit reasons, so the GPU tests and
bench.py can check a full-size run without running the oracle on the GPU
box.  The oracle runs on the host's OpenMP threads; its results do not
depend on the thread count.

Run from the repository root:
    python tests/golden/make_large_golden.py [--csr-cache DIR] c2 [c3 c4 c5]
--csr-cache keeps each generated plain graph as DIR/<workload>.{off,col}.npy
and reuses it (the host generator takes minutes on the 2.4e9-entry graph).
"""
from __future__ import annotations

import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, str(HERE.parent))

C5_FIXTURE_JOBS = 96


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def plain_graph(cfg, cache):
    from dmolp_amd import api, workloads as W
    if cache:
        po = Path(cache) / f"{cfg['workload']}.off.npy"
        pc = Path(cache) / f"{cfg['workload']}.col.npy"
        if po.exists() and pc.exists():
            return np.load(po, mmap_mode="r"), np.load(pc, mmap_mode="r")
    off, col = api.gen_host(W.spec_for(cfg))
    if cache:
        Path(cache).mkdir(parents=True, exist_ok=True)
        np.save(po, off)
        np.save(pc, col)
    return off, col


def plain_case(name, cache=None):
    import oracle_py as O
    from dmolp_amd import workloads as W
    cfg = W.CONFIGS[name]
    k, R, eps = cfg["k"], cfg["rounds"], cfg["eps"]
    t0 = time.time()
    off, col = plain_graph(cfg, cache)
    print(f"[{name}] generated n={len(off) - 1} nnz={len(col)} in {time.time() - t0:.1f} s",
          flush=True)
    rec = {"workload": cfg["workload"], "generator": W.spec_key(W.spec_for(cfg)), "k": k,
           "theta": W.THETA, "max_rounds": R, "epsilon": eps, "n": int(len(off) - 1),
           "nnz": int(len(col)), "off_sha256": sha(off.astype(np.int64))}
    if len(col) < (1 << 30):
        rec["csr_sha256"] = sha(off.astype(np.int64), col.astype(np.int32))
    t0 = time.time()
    p1 = O.phase1(off, col, k, W.THETA)
    t1 = time.time()
    print(f"[{name}] oracle phase 1: {t1 - t0:.1f} s, cut {p1.m.initial_edge_cut}, "
          f"rounds {p1.m.rounds}", flush=True)
    rec["phase1"] = {
        "labels_sha256": sha(p1.labels.astype(np.int32)), "seeds": p1.seeds.tolist(),
        "vc": p1.vc.tolist(), "ec": p1.ec.tolist(), "cut": int(p1.m.initial_edge_cut),
        "vertex_balance": float(p1.m.initial_vertex_balance),
        "edge_balance": float(p1.m.initial_edge_balance),
        "rounds": int(p1.m.rounds), "fill_round": int(p1.m.fill_round),
        "n_hubs": int(p1.m.n_hubs), "first_seed": int(p1.m.first_seed),
        "n_seeds_found": int(p1.m.n_seeds_found),
        "n_seeds_fallback": int(p1.m.n_seeds_fallback),
    }
    p2 = O.phase2(off, col, k, p1.labels, R, eps)
    print(f"[{name}] oracle phase 2: {time.time() - t1:.1f} s, cut {p2.m.initial_cut} -> "
          f"{p2.m.edge_cut}, rounds {p2.m.rounds}", flush=True)
    rec["phase2"] = {
        "labels_sha256": sha(p2.labels.astype(np.int32)), "vc": p2.vc.tolist(),
        "ic": p2.ic.tolist(), "cut": int(p2.m.edge_cut), "initial_cut": int(p2.m.initial_cut),
        "rounds": int(p2.m.rounds), "exit_reason": int(p2.m.exit_reason),
        "lp_edges": int(p2.m.lp_edges), "vertex_balance": float(p2.m.vertex_balance),
        "edge_balance": float(p2.m.edge_balance),
        "trace": {f: p2.trace[f].tolist() for f in ("prev_count", "prev_deg", "cand_count",
                                                    "bd_count", "b_count", "b_deg",
                                                    "chg_count", "chg_deg", "cut")},
    }
    return rec


def job_record(spec, k, theta, rounds, eps):
    import oracle_py as O
    from dmolp_amd import api, workloads as W
    off, col = api.gen_host(spec)
    p1 = O.phase1(off, col, k, theta)
    p2 = O.phase2(off, col, k, p1.labels, rounds, eps)
    return {"generator": W.spec_key(spec), "k": k, "n": int(len(off) - 1), "nnz": int(len(col)),
            "phase1_cut": int(p1.m.initial_edge_cut),
            "n_seeds_fallback": int(p1.m.n_seeds_fallback),
            "labels_sha256": sha(p2.labels.astype(np.int32)), "cut": int(p2.m.edge_cut),
            "rounds": int(p2.m.rounds), "exit_reason": int(p2.m.exit_reason),
            "lp_edges": int(p2.m.lp_edges), "vertex_balance": float(p2.m.vertex_balance),
            "edge_balance": float(p2.m.edge_balance)}


def jobs_case(name):
    from dmolp_amd import workloads as W
    if name == "c4":
        jobs, rounds, eps, wl = W.c4_specs(), W.C4_ROUNDS, W.C4_EPS, W.C4_WORKLOAD
    else:
        jobs, rounds, eps, wl = (W.c5_jobs(C5_FIXTURE_JOBS), W.C5_ROUNDS, W.C5_EPS,
                                 W.C5_WORKLOAD)
    out = []
    t0 = time.time()
    for i, (spec, k) in enumerate(jobs):
        out.append(job_record(spec, k, W.THETA, rounds, eps))
        if i % 16 == 15:
            print(f"[{name}] {i + 1}/{len(jobs)} jobs in {time.time() - t0:.1f} s", flush=True)
    return {"workload": wl, "theta": W.THETA, "max_rounds": rounds, "epsilon": eps, "jobs": out}


def main(argv):
    cache = None
    if argv[:1] == ["--csr-cache"]:
        cache, argv = argv[1], argv[2:]
    for name in argv or ["c2"]:
        rec = plain_case(name, cache) if name in ("c1", "c2", "c3") else jobs_case(name)
        (HERE / f"large_{name}.json").write_text(json.dumps(rec, indent=1) + "\n")
        print(f"[{name}] written", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
