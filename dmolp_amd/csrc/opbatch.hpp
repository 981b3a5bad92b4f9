// opbatch.hpp -- many small device fills, uploads and downloads of a job set
// in a few launches.
#pragma once
#include <cstdint>
#include <vector>

#include "internal.hpp"

namespace dmolp {

// Each hipMemsetAsync or hipMemcpyAsync is a runtime call with tens of
// microseconds of stream time, and a set of a few hundred jobs issues
// thousands of them.  An OpBatch collects them and run() performs all fills,
// then all uploads (one packed host-to-device copy, then one kernel), then
// all downloads (one gather kernel, one device-to-host copy), and returns when
// everything has completed.  Transfers of kDirectBytes or more are copied
// directly (packing them would only add host copies).  Device regions must be
// 4-byte aligned (every allocation of this library is); sizes may be anything.
class OpBatch {
public:
    void fill(void *dst, uint8_t byte, size_t bytes);
    void upload(void *dst, const void *src, size_t bytes);
    void download(void *host, const void *src, size_t bytes);
    void run(dmolp_ctx *ctx);

    struct Piece {
        uint64_t dst, src;  // src == 0: fill with `value`
        uint32_t bytes, value;
    };

private:
    static constexpr size_t kDirectBytes = size_t(1) << 16;
    struct Direct {
        void *dst;
        const void *src;
        size_t bytes;
    };
    std::vector<Piece> fills_, ups_, downs_;  // ups_/downs_ src/dst hold staging offsets
    std::vector<Direct> direct_up_, direct_down_;
    std::vector<uint8_t> up_data_;
    std::vector<void *> down_host_;
    std::vector<uint64_t> down_off_, down_len_;
    uint64_t down_bytes_ = 0;
};

}  // namespace dmolp
