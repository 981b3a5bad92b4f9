// opbatch.hip -- OpBatch: fills, uploads and downloads of a job set in a
// few launches (see opbatch.hpp).
#include <algorithm>
#include <cstring>
#include <memory>

#include "opbatch.hpp"

namespace dmolp {
namespace {

constexpr uint32_t kPieceBytes = 1u << 18;  // one block per piece
constexpr uint64_t kAlign = 16;

// Pieces start 4-byte aligned; each block copies or fills whole words, the
// tail byte by byte.
__global__ void k_pieces(const OpBatch::Piece *P, int n, uint64_t src_base) {
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const OpBatch::Piece p = P[i];
        uint8_t *dst = reinterpret_cast<uint8_t *>(p.dst);
        const uint8_t *src = p.src ? reinterpret_cast<const uint8_t *>(p.src + src_base) : nullptr;
        const uint32_t words = p.bytes / 4;
        uint32_t *dw = reinterpret_cast<uint32_t *>(dst);
        if (src) {
            const uint32_t *sw = reinterpret_cast<const uint32_t *>(src);
            for (uint32_t w = threadIdx.x; w < words; w += blockDim.x) dw[w] = sw[w];
            for (uint32_t b = words * 4 + threadIdx.x; b < p.bytes; b += blockDim.x) dst[b] = src[b];
        } else {
            for (uint32_t w = threadIdx.x; w < words; w += blockDim.x) dw[w] = p.value;
            for (uint32_t b = words * 4 + threadIdx.x; b < p.bytes; b += blockDim.x)
                dst[b] = (uint8_t)(p.value & 0xffu);
        }
    }
}

void split(std::vector<OpBatch::Piece> &out, uint64_t dst, uint64_t src, uint64_t bytes, uint32_t value) {
    for (uint64_t o = 0; o < bytes; o += kPieceBytes) {
        const uint32_t b = (uint32_t)std::min<uint64_t>(kPieceBytes, bytes - o);
        out.push_back(OpBatch::Piece{dst + o, src ? src + o : 0, b, value});
    }
}

uint64_t round_up(uint64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

}  // namespace

void OpBatch::fill(void *dst, uint8_t byte, size_t bytes) {
    if (bytes == 0) return;
    DMOLP_REQUIRE(((uintptr_t)dst & 3) == 0, DMOLP_E_INTERNAL, "unaligned device region");
    split(fills_, (uint64_t)(uintptr_t)dst, 0, bytes, 0x01010101u * byte);
}

void OpBatch::upload(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return;
    DMOLP_REQUIRE(((uintptr_t)dst & 3) == 0, DMOLP_E_INTERNAL, "unaligned device region");
    if (bytes >= kDirectBytes) {
        direct_up_.push_back(Direct{dst, src, bytes});
        return;
    }
    const uint64_t off = up_data_.size();
    up_data_.resize(round_up(off + bytes));
    std::memcpy(up_data_.data() + off, src, bytes);
    // src holds the offset + 1 in the staging area (0 marks a fill); the
    // kernel adds the staging base minus one.
    split(ups_, (uint64_t)(uintptr_t)dst, off + 1, bytes, 0);
}

void OpBatch::download(void *host, const void *src, size_t bytes) {
    if (bytes == 0) return;
    DMOLP_REQUIRE(((uintptr_t)src & 3) == 0, DMOLP_E_INTERNAL, "unaligned device region");
    if (bytes >= kDirectBytes) {
        direct_down_.push_back(Direct{host, src, bytes});
        return;
    }
    // dst holds the offset in the staging area; the host side adds the base.
    split(downs_, down_bytes_, (uint64_t)(uintptr_t)src, bytes, 0);
    down_host_.push_back(host);
    down_off_.push_back(down_bytes_);
    down_len_.push_back(bytes);
    down_bytes_ = round_up(down_bytes_ + bytes);
}

void OpBatch::run(dmolp_ctx *ctx) {
    hipStream_t s = ctx->stream;
    const int grid = ctx->cu_count * 4;
    // One upload: piece lists, then the data of the uploads.
    const uint64_t nf = fills_.size(), nu = ups_.size(), nd = downs_.size();
    const uint64_t desc = round_up(sizeof(Piece) * (nf + nu + nd));
    const uint64_t total_up = desc + up_data_.size();
    if (nf + nu + nd + direct_up_.size() + direct_down_.size() == 0) return;
    DevMem &dev = ctx->ops_dev;
    dev.grow(total_up + down_bytes_);
    uint8_t *base = dev.as<uint8_t>();
    uint8_t *down_area = base + total_up;
    for (Piece &p : downs_) p.dst += (uint64_t)(uintptr_t)down_area;
    std::unique_ptr<uint8_t[]> host(new uint8_t[total_up]);  // left uninitialised
    Piece *hp = reinterpret_cast<Piece *>(host.get());
    std::copy(fills_.begin(), fills_.end(), hp);
    std::copy(ups_.begin(), ups_.end(), hp + nf);
    std::copy(downs_.begin(), downs_.end(), hp + nf + nu);
    if (!up_data_.empty()) std::memcpy(host.get() + desc, up_data_.data(), up_data_.size());
    DMOLP_HIP_CHECK(hipMemcpyAsync(base, host.get(), total_up, hipMemcpyHostToDevice, s));
    const Piece *dp = reinterpret_cast<const Piece *>(base);
    if (nf) hipLaunchKernelGGL(k_pieces, dim3(std::min<uint64_t>(nf, grid)), dim3(256), 0, s, dp, (int)nf, 0ull);
    for (const Direct &d : direct_up_)  // after the fills, like the packed uploads
        DMOLP_HIP_CHECK(hipMemcpyAsync(d.dst, d.src, d.bytes, hipMemcpyHostToDevice, s));
    if (nu)
        hipLaunchKernelGGL(k_pieces, dim3(std::min<uint64_t>(nu, grid)), dim3(256), 0, s, dp + nf, (int)nu,
                           (uint64_t)(uintptr_t)(base + desc) - 1ull);
    if (nd) hipLaunchKernelGGL(k_pieces, dim3(std::min<uint64_t>(nd, grid)), dim3(256), 0, s, dp + nf + nu, (int)nd, 0ull);
    DMOLP_HIP_CHECK(hipGetLastError());
    for (const Direct &d : direct_down_)
        DMOLP_HIP_CHECK(hipMemcpyAsync(d.dst, d.src, d.bytes, hipMemcpyDeviceToHost, s));
    std::unique_ptr<uint8_t[]> back(new uint8_t[down_bytes_ > 0 ? down_bytes_ : 1]);
    if (down_bytes_) DMOLP_HIP_CHECK(hipMemcpyAsync(back.get(), down_area, down_bytes_, hipMemcpyDeviceToHost, s));
    DMOLP_HIP_CHECK(hipStreamSynchronize(s));
    for (size_t i = 0; i < down_host_.size(); ++i)
        std::memcpy(down_host_[i], back.get() + down_off_[i], down_len_[i]);
    fills_.clear();
    direct_up_.clear();
    direct_down_.clear();
    ups_.clear();
    downs_.clear();
    up_data_.clear();
    down_host_.clear();
    down_off_.clear();
    down_len_.clear();
    down_bytes_ = 0;
}

}  // namespace dmolp
