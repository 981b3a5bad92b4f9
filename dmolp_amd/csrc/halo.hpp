// halo.hpp -- rank shares of one graph with on-device ghost labels
// (halo.hip).
#pragma once
#include <memory>
#include <vector>

#include "internal.hpp"

namespace dmolp {

// Device view of one rank's share: owned rows (local id i is global id
// r + i * P), columns in unified ids (owned: local id; ghost: nl + slot),
// unified labels (owned, then ghosts), and the ghost lookup (global ids
// ascending with their slots).
struct ShareDev {
    int64_t nl, ng, nnz;
    const int64_t *off;
    const int32_t *col;
    int32_t *lab;
    int32_t *nxt;               // this round's decisions (Jacobi)
    const int32_t *ghost_gid;   // slot -> global id
    const int32_t *sgid;        // ghost global ids ascending
    const int32_t *sslot;       // their slots
    uint8_t *inb, *nb;          // previous / new boundary flags of owned vertices
    uint8_t *ghost_inb;         // ghosts in some rank's previous boundary (scope 1)
    unsigned long long *key;    // boundary record keys of owned vertices
};

struct Share {
    int64_t nl = 0, ng = 0, nnz = 0;
    DevMem off, col, lab, nxt, ghost_gid, sgid, sslot, inb, nb, ghost_inb, key;
    DevMem d_gid, d_lab;  // this share's deltas of the round
    int64_t d_cap = 0;
    ShareDev view() const {
        return ShareDev{nl,
                        ng,
                        nnz,
                        off.as<int64_t>(),
                        col.as<int32_t>(),
                        lab.as<int32_t>(),
                        nxt.as<int32_t>(),
                        ghost_gid.as<int32_t>(),
                        sgid.as<int32_t>(),
                        sslot.as<int32_t>(),
                        inb.as<uint8_t>(),
                        nb.as<uint8_t>(),
                        ghost_inb.as<uint8_t>(),
                        key.as<unsigned long long>()};
    }
};

}  // namespace dmolp

struct dmolp_shares {
    int32_t P = 0;
    int64_t n = 0;
    dmolp_graph *g = nullptr;
    std::vector<dmolp::Share> sh;
    std::vector<int64_t> deg_h;   // global degrees (host), for the trace
    std::vector<int64_t> dcount;  // deltas per share this round
    dmolp::DevMem dev;            // ShareDev[P]
    dmolp::DevMem temp;           // hipCUB temporary storage
    dmolp::DevMem count, upd, parts;
    void upload_views(dmolp_ctx *ctx);
    int64_t global_deg(int64_t v) const { return deg_h[(size_t)v]; }
    int64_t sh_deg(int32_t r, int64_t i) const { return deg_h[(size_t)(r + i * P)]; }
    void deltas_of(int32_t r, int64_t m) {
        dmolp::Share &s = sh[(size_t)r];
        if (s.d_cap < m || !s.d_gid.p) {
            const int64_t cap = m > 0 ? m : 1;
            s.d_gid.alloc(sizeof(int32_t) * (size_t)cap);
            s.d_lab.alloc(sizeof(int32_t) * (size_t)cap);
            s.d_cap = cap;
        }
    }
};

namespace dmolp {

dmolp_shares *shares_create(dmolp_ctx *ctx, dmolp_graph *g, int32_t P);
// Apply m deltas (device arrays of global ids and labels) to share r's
// ghost labels; returns how many ghost labels were rewritten.
int64_t halo_refresh(dmolp_ctx *ctx, dmolp_shares *S, int32_t r, const int32_t *d_gid,
                     const int32_t *d_lab, int64_t m);
void shares_phase2(dmolp_ctx *ctx, dmolp_shares *S, int32_t k, int32_t max_rounds, double epsilon,
                   int32_t scope, int32_t *labels_out, int64_t *vc_out, int64_t *ic_out,
                   dmolp_phase2_metrics *m, dmolp_trace *tr, int64_t *ghost_updates);

}  // namespace dmolp
