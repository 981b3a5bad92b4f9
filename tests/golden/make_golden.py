"""Write tests/golden/partitions.json: end-to-end results of the CPU oracle
(phase 1 + phase 2, one rank) on seeded generator graphs.

The fixtures freeze the oracle's results so that a later change to the
oracle, the generators or the device path that alters any label shows up as
a diff.  Each case stores the generator spec, a SHA-256 of the CSR, and for
each phase the SHA-256 of the int32 label array with the cut, part sizes,
round count, exit reason and per-round trace.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, str(HERE.parent))

CASES = [
    # name, generator, k, theta, max_rounds, epsilon
    ("rmat-s16-ef16-seed1-k8", ("rmat", 16, 16, 1), 8, 10000, 10, 0.03),
    ("grid-64x64-k4", ("grid", 64, 64), 4, 100, 10, 0.03),
    ("rmat-s10-ef16-seed11-k64", ("rmat", 10, 16, 11), 64, 1000, 10, 0.03),
    ("chunglu-n20000-m200000-k32", ("chunglu", 20000, 200000, 2000.0, 3), 32, 500, 15, 0.05),
    ("rmat-s12-ef8-seed5-k16", ("rmat", 12, 8, 5), 16, 50, 20, 0.01),
]


def spec_of(gen):
    from dmolp_amd import api
    if gen[0] == "rmat":
        return api.rmat_spec(gen[1], gen[2], seed=gen[3])
    if gen[0] == "grid":
        return api.grid_spec(gen[1], gen[2])
    return api.chung_lu_spec(gen[1], gen[2], gamma=2.0, dmax=gen[3], seed=gen[4])


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def run_case(name, gen, k, theta, rounds, eps):
    import oracle_py as O
    from dmolp_amd import api
    off, col = api.gen_host(spec_of(gen))
    p1 = O.phase1(off, col, k, theta)
    p2 = O.phase2(off, col, k, p1.labels, rounds, eps)
    return {
        "name": name, "generator": list(gen), "k": k, "theta": theta, "max_rounds": rounds,
        "epsilon": eps, "n": int(len(off) - 1), "nnz": int(len(col)),
        "csr_sha256": sha(off.astype(np.int64), col.astype(np.int32)),
        "phase1": {
            "labels_sha256": sha(p1.labels.astype(np.int32)),
            "seeds": p1.seeds.tolist(), "vc": p1.vc.tolist(), "ec": p1.ec.tolist(),
            "cut": int(p1.m.initial_edge_cut), "rounds": int(p1.m.rounds),
            "n_hubs": int(p1.m.n_hubs), "n_seeds_found": int(p1.m.n_seeds_found),
            "n_seeds_fallback": int(p1.m.n_seeds_fallback),
        },
        "phase2": {
            "labels_sha256": sha(p2.labels.astype(np.int32)),
            "vc": p2.vc.tolist(), "ic": p2.ic.tolist(), "cut": int(p2.m.edge_cut),
            "initial_cut": int(p2.m.initial_cut), "rounds": int(p2.m.rounds),
            "exit_reason": int(p2.m.exit_reason), "lp_edges": int(p2.m.lp_edges),
            "trace": {f: p2.trace[f].tolist() for f in ("b_count", "b_deg", "chg_count", "cut")},
        },
    }


def main():
    out = [run_case(*c) for c in CASES]
    (HERE / "partitions.json").write_text(json.dumps(out, indent=1) + "\n")
    for c in out:
        print(c["name"], "n", c["n"], "nnz", c["nnz"], "p1 cut", c["phase1"]["cut"],
              "p2 cut", c["phase2"]["cut"], "rounds", c["phase2"]["rounds"],
              "fallback seeds", c["phase1"]["n_seeds_fallback"])


if __name__ == "__main__":
    main()
