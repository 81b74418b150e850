// nyl_ctx.h -- the model context shared by the translation units of
// libnyles_hip.so (not part of the public ABI).
#pragma once
#include "nyl_internal.h"
#include "../../include/nyles_hip.h"

namespace nyl {

constexpr int kMaxLevels = 32;

struct Level {
    Geo g;
    uint64_t x, b, y;   // byte offsets in the workspace
};

}  // namespace nyl

struct nyl_ctx {
    nyl_desc d;
    nyl::Geo g;
    int mixed;               // NYL_PREC_MIXED: FP32 model fields and inner V-cycles
    int64_t ncell;
    int nlev;
    nyl::Level lev[nyl::kMaxLevels];
    int exact;               // the coarsest level has an odd size: solved directly
    uint64_t cs_tab;         // its eigenvector tables (byte offset)
    nyl::MGTabs tabs;
    // byte offsets of the fields
    uint64_t state[2][4];
    uint64_t ds[3][4];
    uint64_t sb[4];          // LFAM3 'before' state
    // passive tracers t0.. (state double-buffered with b, u, v, w)
    int ntr;
    uint64_t tst[2][NYL_MAX_TRACERS], tds[3][NYL_MAX_TRACERS], tsb[NYL_MAX_TRACERS];
    int lf_first;            // next LFAM3 step is the first one
    uint64_t ke, vor[3], uc[3], p;
    uint64_t mgx, mgb;       // FP64 solution and right-hand side of the projection
    uint64_t partials, result;
    uint64_t ws_bytes;
    int cur;                 // which state buffer is current
    char *ws;
    hipStream_t stream;
    double *host_out;        // pinned host memory, coherent: 8 result doubles + a sequence flag
    double *host_dev;        // host_out as the device addresses it
    uint32_t pub_seq;        // last published result sequence number
    int total_cycles;
    // the V-cycle below level 1, replayed from a hipGraph (captured on first use)
    int use_graphs;
    hipGraphExec_t g_coarse;
    // HIP-event timing of one kernel class (nyl_prof_start / nyl_prof_stop)
    int prof_cls;
    int prof_used;
    hipEvent_t prof_ev[2 * 4096];
    int prof_nev;            // events created
    double prof_ms, prof_bytes;
    int64_t prof_launches;
};

namespace nyl {

template <typename T>
inline T *F(const nyl_ctx *c, uint64_t off) { return reinterpret_cast<T *>(c->ws + off); }

// record an error message for nyl_last_error() and return code
int set_error(int code, const char *fmt, ...);

// Level l holds FP64 arrays in FP64 mode; in mixed mode the inner V-cycle
// levels are FP32 and the coarsest level stays FP64.
// Mixed precision: levels 0 .. NYL_MIXED_F32_LEVELS-1 in FP32 (never the
// coarsest), the others in FP64.
#ifndef NYL_MIXED_F32_LEVELS
#define NYL_MIXED_F32_LEVELS 64
#endif
inline bool lev64(const nyl_ctx *c, int l) {
    return !c->mixed || l == c->nlev - 1 || l >= NYL_MIXED_F32_LEVELS;
}

}  // namespace nyl
