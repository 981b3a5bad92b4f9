// report.cpp -- printComparisonReport (utils.cpp:337-394): the reference's
// end-of-run table comparing the phase-1 partition with the phase-2 one.
// The text is an output format that downstream log parsers read, so every
// line, label, field width and precision is the reference's; the numbers
// go through std::cout with its stream state as the reference leaves it
// (std::fixed stays set after the first fixed field).
#include <iomanip>
#include <iostream>
#include <string>

#include "../../../include/graph_types.h"
#include "../../../include/phase2/phase2.h"

namespace {

const char *const kRule = "─────────────────────────────────────────────────────────────";
const char *const kBlank = "│                                                             │\n";
const char *const kIntTail = "                              │\n";     // after a width-10 integer
const char *const kFixTail = "                             │\n";      // after a width-8 fixed value
const char *const kPctTail = "%                             │\n";
const char *const kMsTail = " ms                          │\n";

double improvement(double before, double after) {
    return before > 0 ? (before - after) / before * 100.0 : 0.0;
}

void section(const char *title) { std::cout << title; }

// "│   Phase 1 (초기): " and its phase-2 twin, then the width-10 integers.
void int_pair(long long a, long long b) {
    std::cout << "│   Phase 1 (초기): " << std::setw(10) << a << kIntTail;
    std::cout << "│   Phase 2 (최종): " << std::setw(10) << b << kIntTail;
}

void fix_pair(double a, double b) {
    std::cout << "│   Phase 1 (초기): " << std::setw(8) << std::fixed << std::setprecision(4) << a
              << kFixTail;
    std::cout << "│   Phase 2 (최종): " << std::setw(8) << std::fixed << std::setprecision(4) << b
              << kFixTail;
}

void pct_line(double pct) {
    std::cout << "│   개선율:         " << std::setw(8) << std::fixed << std::setprecision(2) << pct
              << kPctTail;
}

// The summary lines: improved (✓) or worse (⚠), one decimal, with the two
// values in the stream's current format.
template <typename T>
void verdict(const char *what, double pct, T before, T after) {
    if (pct > 0)
        std::cout << "✓ " << what << " " << std::fixed << std::setprecision(1) << pct << "% 개선 (";
    else
        std::cout << "⚠ " << what << " " << std::fixed << std::setprecision(1) << -pct << "% 악화 (";
    std::cout << before << " → " << after << ")\n";
}

}  // namespace

void printComparisonReport(const PartitioningMetrics &m1, const PartitioningMetrics &m2) {
    const double cut_pct =
        m1.edge_cut > 0 ? static_cast<double>(m1.edge_cut - m2.edge_cut) / m1.edge_cut * 100.0 : 0.0;
    const double vb_pct = improvement(m1.vertex_balance, m2.vertex_balance);
    const double eb_pct = improvement(m1.edge_balance, m2.edge_balance);
    const long total_ms = m1.loading_time_ms + m1.distribution_time_ms + m2.loading_time_ms;

    std::cout << "\n┌" << kRule << "┐\n";
    std::cout << "│                    메트릭 비교 결과                         │\n";
    std::cout << "├" << kRule << "┤\n";
    section("│ 그래프 정보:                                                │\n");
    std::cout << "│   총 노드 수:     " << std::setw(10) << m1.total_vertices << kIntTail;
    std::cout << "│   총 간선 수:     " << std::setw(10) << m1.total_edges << kIntTail;
    std::cout << "│   파티션 수:      " << std::setw(10) << m1.num_partitions << kIntTail;
    std::cout << kBlank;
    section("│ Edge-cut:                                                   │\n");
    int_pair(m1.edge_cut, m2.edge_cut);
    pct_line(cut_pct);
    std::cout << kBlank;
    section("│ Vertex Balance:                                             │\n");
    fix_pair(m1.vertex_balance, m2.vertex_balance);
    pct_line(vb_pct);
    std::cout << kBlank;
    section("│ Edge Balance:                                               │\n");
    fix_pair(m1.edge_balance, m2.edge_balance);
    pct_line(eb_pct);
    std::cout << kBlank;
    section("│ 실행 시간:                                                  │\n");
    std::cout << "│   Phase 1 (로딩): " << std::setw(6) << m1.loading_time_ms << kMsTail;
    std::cout << "│   Phase 1 (분산): " << std::setw(6) << m1.distribution_time_ms << kMsTail;
    std::cout << "│   Phase 2 (7단계): " << std::setw(5) << m2.loading_time_ms << kMsTail;
    std::cout << "│   총 소요시간:     " << std::setw(5) << total_ms << kMsTail;
    std::cout << "└" << kRule << "┘\n";
    std::cout << "\n=== 알고리즘 성능 요약 ===\n";
    verdict("Edge-cut", cut_pct, m1.edge_cut, m2.edge_cut);
    verdict("Vertex Balance", vb_pct, m1.vertex_balance, m2.vertex_balance);
    verdict("Edge Balance", eb_pct, m1.edge_balance, m2.edge_balance);
    std::cout << "총 소요시간: " << total_ms << " ms\n";
    std::cout << "\n=== 7단계 알고리즘 완료 ===\n";
}
