"""DMOLP_NO_GRAPHS=1 launches the phase-1 and phase-2 rounds kernel by kernel
instead of replaying captured graphs: same kernels, same results.  The
switch is read once per process, so the run without graphs is a child
process; its labels and statistics must equal this process's (graphs on)
and the oracle's."""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_py as O
from dmolp_amd import api

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import hashlib, json, sys
sys.path.insert(0, {root!r})
from dmolp_amd import api
with api.Context(0) as ctx:
    g = ctx.upload(*api.gen_host(api.rmat_spec(15, 16, seed=77)))
    p1 = g.phase1(16, 300)
    p2 = g.phase2(16, 12, 0.01)
    print(json.dumps({{"p1": hashlib.sha256(p1.labels.tobytes()).hexdigest(),
                      "p1_rounds": int(p1.metrics.rounds),
                      "p2": hashlib.sha256(p2.labels.tobytes()).hexdigest(),
                      "cut": int(p2.metrics.edge_cut), "rounds": int(p2.metrics.rounds)}}))
    g.free()
"""


def run_child(no_graphs: bool) -> dict:
    env = dict(os.environ)
    if no_graphs:
        env["DMOLP_NO_GRAPHS"] = "1"
    else:
        env.pop("DMOLP_NO_GRAPHS", None)
    print(f"[nographs] child DMOLP_NO_GRAPHS={'1' if no_graphs else 'unset'}", flush=True)
    out = subprocess.run([sys.executable, "-c", CHILD.format(root=ROOT)], env=env, capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    return json.loads(out.stdout.strip().splitlines()[-1])


def test_rounds_without_graphs_match(ctx):
    off, col = api.gen_host(api.rmat_spec(15, 16, seed=77))
    want1 = O.phase1(off, col, 16, 300)
    want2 = O.phase2(off, col, 16, want1.labels, 12, 0.01)
    a = run_child(False)
    b = run_child(True)
    assert a == b
    assert a["p1"] == hashlib.sha256(np.ascontiguousarray(want1.labels, np.int32).tobytes()).hexdigest()
    assert a["p1_rounds"] == want1.m.rounds
    assert a["p2"] == hashlib.sha256(np.ascontiguousarray(want2.labels, np.int32).tobytes()).hexdigest()
    assert a["cut"] == want2.m.edge_cut and a["rounds"] == want2.m.rounds
