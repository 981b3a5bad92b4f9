"""The drop-in path end to end: the reference's command line
(hpc_partitioning -m bfs|random <graph> <k> <theta>, main.cpp:123-283) on a
file in the reference's text adjacency format, through run_phase1 /
run_phase2 on both graph layouts (the reference's int32 Graph and
GraphWide), against the oracle: final labels, the phase-1 metrics lines and
the comparison report's cut lines."""
from __future__ import annotations

import re
import subprocess

import numpy as np
import pytest

import oracle_py as O
from conftest import ROOT
from dmolp_amd import api

pytestmark = pytest.mark.gpu

CLI = ROOT / "dmolp_amd" / "hpc_partitioning"


def run_cli(tmp_path, graph, mode, k, theta, rounds, eps, layout, verbose=False):
    out = tmp_path / f"labels-{mode}-{layout}.bin"
    cmd = [str(CLI), "-m", mode, str(graph), str(k), str(theta), "--rounds", str(rounds),
           "--epsilon", str(eps), "--labels-out", str(out), "--layout", layout]
    if verbose:
        cmd.insert(1, "--verbose")
    print("[dropin]", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout, np.fromfile(out, np.int32)


def field(text, pattern):
    m = re.search(pattern, text, flags=re.M)
    assert m, pattern
    return m.group(1)


@pytest.fixture(scope="module")
def graph_file(tmp_path_factory):
    off, col = api.gen_host(api.rmat_spec(12, 16, seed=31))
    p = tmp_path_factory.mktemp("dropin") / "rmat12.txt"
    api.write_adjacency_text(p, off, col)
    return p, off, col


@pytest.mark.parametrize("layout", ["graph", "wide"])
def test_cli_bfs_matches_oracle(tmp_path, graph_file, layout):
    path, off, col = graph_file
    k, theta, rounds, eps = 8, 200, 10, 0.03
    text, labels = run_cli(tmp_path, path, "bfs", k, theta, rounds, eps, layout)
    w1 = O.phase1(off, col, k, theta)
    w2 = O.phase2(off, col, k, w1.labels, rounds, eps)
    assert np.array_equal(labels, w2.labels)
    assert int(field(text, r"^Initial edge cut: (\d+)")) == w1.m.initial_edge_cut
    assert field(text, r"^Partition vertex counts: ([\d ]+)").split() == [str(x) for x in w1.vc]
    assert field(text, r"^Partition edge counts: ([\d ]+)").split() == [str(x) for x in w1.ec]
    assert int(field(text, r"Phase 1 \(초기\):\s+(\d+)")) == w1.m.initial_edge_cut
    assert int(field(text, r"Phase 2 \(최종\):\s+(\d+)")) == w2.m.edge_cut
    assert float(field(text, r"^Initial vertex balance: ([\d.]+)")) == \
        pytest.approx(w1.m.initial_vertex_balance, abs=5e-5)


@pytest.mark.parametrize("layout", ["graph", "wide"])
def test_cli_random_mode_matches_oracle(tmp_path, graph_file, layout):
    """Random mode (random.cpp:71-74): every label is the vertex id mod k;
    phase 2 then runs from those labels."""
    path, off, col = graph_file
    n = len(off) - 1
    k, rounds, eps = 5, 12, 0.01
    text, labels = run_cli(tmp_path, path, "random", k, 100, rounds, eps, layout)
    start = (np.arange(n) % k).astype(np.int32)
    vc, ec, m = O.phase1_metrics(off, col, k, start)
    w2 = O.phase2(off, col, k, start, rounds, eps)
    assert np.array_equal(labels, w2.labels)
    assert int(field(text, r"^Initial edge cut: (\d+)")) == m.initial_edge_cut
    assert field(text, r"^Partition vertex counts: ([\d ]+)").split() == [str(x) for x in vc]
    assert int(field(text, r"Phase 2 \(최종\):\s+(\d+)")) == w2.m.edge_cut


def test_cli_second_run_reads_the_binary_cache(tmp_path):
    """The first run parses the text file and leaves <file>.csr; the next
    run loads that cache and gives the same partition."""
    off, col = api.gen_host(api.rmat_spec(11, 8, seed=33))
    p = tmp_path / "rmat11.txt"
    api.write_adjacency_text(p, off, col)
    t1, l1 = run_cli(tmp_path, p, "bfs", 4, 100, 8, 0.03, "wide", verbose=True)
    assert "(text)" in t1 and (tmp_path / "rmat11.txt.csr").exists()
    t2, l2 = run_cli(tmp_path, p, "bfs", 4, 100, 8, 0.03, "wide", verbose=True)
    assert "(binary cache)" in t2
    assert np.array_equal(l1, l2)
    w1 = O.phase1(off, col, 4, 100)
    assert np.array_equal(l2, O.phase2(off, col, 4, w1.labels, 8, 0.03).labels)


def test_phase1_random_abi_labels(ctx, graph_file):
    """dmolp_phase1_random directly: labels id mod k and the oracle's
    metrics for them."""
    _, off, col = graph_file
    n = len(off) - 1
    g = ctx.upload(off, col)
    try:
        for k in (1, 3, 16, 40):
            r = g.phase1_random(k)
            start = (np.arange(n) % k).astype(np.int32)
            assert np.array_equal(r.labels, start), k
            assert np.array_equal(g.get_labels(), start), k
            rc, wvc, wic, wcut = O.stats(off, col, k, start)
            assert rc == 0 and np.array_equal(r.vc, wvc) and r.metrics.initial_edge_cut == wcut, k
            assert np.array_equal(r.ec, wic), k
    finally:
        g.free()
