// gen_host.hpp -- host generators and CSR builder (gen_host.cpp).
#pragma once
#include <cstdint>
#include <functional>
#include <vector>

#include "../../include/dmolp_hip.h"
#include "gen_common.hpp"

namespace dmolp {

int host_threads();
// body(a, b) over ranges of [0, n) on the host threads; grain > 0 fixes the
// range length (default: about n / (64 threads), at least 1024).
void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)> &body, int64_t grain = 0);
RmatThresholds rmat_thresholds(double a, double b, double c);
std::vector<uint64_t> chung_lu_cdf(int64_t n, int64_t m, double gamma, double dmax);
void build_csr(int64_t n, int64_t m, const uint32_t *src, const uint32_t *dst, dmolp_csr *out);
void gen_host(const dmolp_gen_spec *spec, dmolp_csr *out);

}  // namespace dmolp
