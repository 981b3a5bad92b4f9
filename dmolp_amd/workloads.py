"""The measured workloads, defined once for bench.py, the tests and the
fixture scripts.

Plain graphs (one graph per run):
  c1  R-MAT scale 16, edge factor 16, a/b/c = 0.57/0.19/0.19, seed 1;
      k=8, 10 rounds, epsilon 0.03
  c2  R-MAT scale 22, edge factor 16, same a/b/c, seed 2; k=16, 20 rounds,
      epsilon 0.03
  c3  Chung-Lu power-law graph, 41,652,230 vertices, 1,468,365,182 sampled
      directed edges, degree exponent 2.0, expected maximum degree 3e6,
      seed 3; k=32, 30 rounds, epsilon 0.05
Batched (c4): 256 graphs, half R-MAT (scale uniform in 10..18, edge factor
16), half 2-D 5-point grids (sides log-uniform in 64..1024); k from
{2,4,8,16,64}; 10 rounds.
Served (c5): R-MAT requests, vertices log-uniform in 1e3..1e7 (scale = the
nearest power of two), edge factor uniform in 8..32, k from {4,...,128};
20 rounds.

theta = 10000 everywhere (the frontier throttle of phase 1).
"""
from __future__ import annotations

import numpy as np

THETA = 10000

CONFIGS = {
    "c1": dict(workload="rmat-s16-ef16-seed1", gen=("rmat", 16, 16, 1), k=8, rounds=10, eps=0.03),
    "c2": dict(workload="rmat-s22-ef16-seed2", gen=("rmat", 22, 16, 2), k=16, rounds=20, eps=0.03),
    "c3": dict(workload="chunglu-n41652230-m1468365182-g2.0-dmax3e6-seed3",
               gen=("chunglu", 41652230, 1468365182, 3), k=32, rounds=30, eps=0.05),
}

C4_WORKLOAD = "batch256-rmat-s10to18-grid64to1024-seed4"
C4_ROUNDS, C4_EPS = 10, 0.03
C5_WORKLOAD = "served-rmat-n1e3to1e7-ef8to32-seed5"
C5_ROUNDS, C5_EPS = 20, 0.03


def spec_for(cfg):
    from . import api
    g = cfg["gen"]
    if g[0] == "rmat":
        return api.rmat_spec(g[1], g[2], seed=g[3])
    return api.chung_lu_spec(g[1], g[2], gamma=2.0, dmax=3.0e6, seed=g[3])


def c4_specs(seed=4, count=256):
    """(spec, k) of the 256 batched jobs, in submission order."""
    from . import api
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(count):
        k = int(rng.choice([2, 4, 8, 16, 64]))
        if i % 2 == 0:
            jobs.append((api.rmat_spec(int(rng.integers(10, 19)), 16, seed=1000 + i), k))
        else:
            r, c = (int(round(x)) for x in np.exp(rng.uniform(np.log(64), np.log(1024), 2)))
            jobs.append((api.grid_spec(r, c), k))
    return jobs


def c5_jobs(count, seed=5):
    """(spec, k) of the first `count` served requests, in arrival order."""
    from . import api
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = float(np.exp(rng.uniform(np.log(1e3), np.log(1e7))))
        scale = int(min(23, max(10, round(np.log2(n)))))
        out.append((api.rmat_spec(scale, int(rng.integers(8, 33)), seed=5000 + i),
                    int(rng.choice([4, 8, 16, 32, 64, 128]))))
    return out


def spec_key(spec) -> list:
    """A JSON-able description of a generator spec (fixtures record it)."""
    if spec.kind == 1:
        return ["rmat", int(spec.scale), int(spec.edge_factor), float(spec.a), float(spec.b),
                float(spec.c), int(spec.seed)]
    if spec.kind == 3:
        return ["grid", int(spec.rows), int(spec.cols)]
    return ["chunglu", int(spec.n), int(spec.m), float(spec.gamma), float(spec.dmax),
            int(spec.seed)]
