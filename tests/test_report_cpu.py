"""The end-of-run report keeps the reference's text format
(printComparisonReport, utils.cpp:337-394, after Phase1Metrics::print,
graph_types.h:287-308): a small C++ caller prints both through the library
and the output must equal the format restated here from those lines.  No
GPU is touched."""
from __future__ import annotations

import subprocess

import pytest

from conftest import ROOT

CALLER = r"""
#include <cstdlib>
#include "graph_types.h"
#include "phase2/phase2.h"
int main(int argc, char **argv) {
    Phase1Metrics m;
    m.initial_edge_cut = std::atoi(argv[1]);
    m.initial_vertex_balance = std::atof(argv[2]);
    m.initial_edge_balance = std::atof(argv[3]);
    m.loading_time_ms = std::atol(argv[4]);
    m.distribution_time_ms = std::atol(argv[5]);
    m.total_vertices = std::atoi(argv[6]);
    m.total_edges = std::strtoull(argv[7], nullptr, 10);
    m.partition_vertex_counts = {3, 1, 4};
    m.partition_edge_counts = {1, 5, 9};
    PartitioningMetrics p1(m, 3), p2;
    p2.edge_cut = std::atoi(argv[8]);
    p2.vertex_balance = std::atof(argv[9]);
    p2.edge_balance = std::atof(argv[10]);
    p2.loading_time_ms = std::atol(argv[11]);
    p2.num_partitions = 3;
    m.print();
    printComparisonReport(p1, p2);
    return 0;
}
"""


def expected(a):
    """The reference's lines for arguments a (same order as the caller)."""
    cut1, vb1, eb1, load, dist, nv, ne, cut2, vb2, eb2, ms2 = a
    cut1, load, dist, nv, ne, cut2, ms2 = map(int, (cut1, load, dist, nv, ne, cut2, ms2))
    vb1, eb1, vb2, eb2 = map(float, (vb1, eb1, vb2, eb2))
    out = ["=== Phase 1 Metrics ===",
           f"Initial edge cut: {cut1}",
           f"Initial vertex balance: {vb1:.4f}",
           f"Initial edge balance: {eb1:.4f}",
           f"Loading time: {load} ms",
           f"Distribution time: {dist} ms",
           f"Total vertices: {nv}",
           f"Total edges: {ne}",
           "Partition vertex counts: 3 1 4 ",
           "Partition edge counts: 1 5 9 ",
           ""]
    rule = "─" * 61
    cut_pct = (cut1 - cut2) / cut1 * 100.0 if cut1 > 0 else 0.0
    vb_pct = (vb1 - vb2) / vb1 * 100.0 if vb1 > 0 else 0.0
    eb_pct = (eb1 - eb2) / eb1 * 100.0 if eb1 > 0 else 0.0
    total = load + dist + ms2
    it, ft, pt, mt = (" " * 30 + "│", " " * 29 + "│", "%" + " " * 29 + "│", " ms" + " " * 26 + "│")
    blank = "│" + " " * 61 + "│"
    out += ["",
            f"┌{rule}┐",
            "│                    메트릭 비교 결과                         │",
            f"├{rule}┤",
            "│ 그래프 정보:                                                │",
            f"│   총 노드 수:     {nv:>10d}{it}",
            f"│   총 간선 수:     {ne:>10d}{it}",
            f"│   파티션 수:      {3:>10d}{it}",
            blank,
            "│ Edge-cut:                                                   │",
            f"│   Phase 1 (초기): {cut1:>10d}{it}",
            f"│   Phase 2 (최종): {cut2:>10d}{it}",
            f"│   개선율:         {cut_pct:>8.2f}{pt}",
            blank,
            "│ Vertex Balance:                                             │",
            f"│   Phase 1 (초기): {vb1:>8.4f}{ft}",
            f"│   Phase 2 (최종): {vb2:>8.4f}{ft}",
            f"│   개선율:         {vb_pct:>8.2f}{pt}",
            blank,
            "│ Edge Balance:                                               │",
            f"│   Phase 1 (초기): {eb1:>8.4f}{ft}",
            f"│   Phase 2 (최종): {eb2:>8.4f}{ft}",
            f"│   개선율:         {eb_pct:>8.2f}{pt}",
            blank,
            "│ 실행 시간:                                                  │",
            f"│   Phase 1 (로딩): {load:>6d}{mt}",
            f"│   Phase 1 (분산): {dist:>6d}{mt}",
            f"│   Phase 2 (7단계): {ms2:>5d}{mt}",
            f"│   총 소요시간:     {total:>5d}{mt}",
            f"└{rule}┘",
            "",
            "=== 알고리즘 성능 요약 ==="]

    def verdict(what, pct, b, c, fmt):
        word = "개선" if pct > 0 else "악화"
        mark = "✓" if pct > 0 else "⚠"
        shown = pct if pct > 0 else -pct  # the reference negates: 0.0 prints as -0.0
        return f"{mark} {what} {shown:.1f}% {word} ({fmt(b)} → {fmt(c)})"
    out.append(verdict("Edge-cut", cut_pct, cut1, cut2, str))
    out.append(verdict("Vertex Balance", vb_pct, vb1, vb2, lambda x: f"{x:.1f}"))
    out.append(verdict("Edge Balance", eb_pct, eb1, eb2, lambda x: f"{x:.1f}"))
    out += [f"총 소요시간: {total} ms", "", "=== 7단계 알고리즘 완료 ==="]
    return "\n".join(out) + "\n"


@pytest.fixture(scope="module")
def caller(tmp_path_factory):
    d = tmp_path_factory.mktemp("report")
    src = d / "caller.cpp"
    src.write_text(CALLER)
    exe = d / "caller"
    lib = ROOT / "dmolp_amd"
    if not (lib / "libdmolp_hip.so").exists():
        pytest.skip("library not built")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", str(ROOT / "include"), str(src), "-o",
                    str(exe), "-L", str(lib), "-ldmolp_hip", f"-Wl,-rpath,{lib}"], check=True)
    return exe


@pytest.mark.parametrize("args", [
    ("12295", "1.8731", "2.2504", "41", "250", "4096", "48347", "12268", "1.0402", "1.1984", "17"),
    # worse cut and balances, and a zero phase-1 cut
    ("100", "1.0", "1.0", "1", "2", "10", "20", "120", "1.25", "3.5", "123456"),
    ("0", "0.0", "0.0", "0", "0", "1", "0", "0", "0.0", "0.0", "0"),
])
def test_report_matches_reference_format(caller, args):
    out = subprocess.run([str(caller), *args], check=True, capture_output=True, text=True).stdout
    assert out == expected(args)
