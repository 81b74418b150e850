// nyl_api.hip -- context, workspace layout and the extern "C" entry points of
// libnyles_hip.so (declared in include/nyles_hip.h).
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <type_traits>
#include <utility>
#include <vector>

#include "nyl_ctx.h"
#include <atomic>
#include <chrono>

using namespace nyl;

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

}  // namespace

namespace nyl {
int set_error(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}
}  // namespace nyl

namespace {

#define HIPCHK(...)                                                                    \
    do {                                                                               \
        hipError_t e_ = (__VA_ARGS__);                                                 \
        if (e_ != hipSuccess)                                                          \
            return fail(NYL_E_HIP, "%s: %s (%s:%d)", #__VA_ARGS__,                    \
                        hipGetErrorString(e_), __FILE__, __LINE__);                    \
    } while (0)

}  // namespace

namespace {

// prolongation fused into the first post-smoothing sweep of tiled levels
#ifndef NYL_FUSE_PROLONG
#define NYL_FUSE_PROLONG 0
#endif


uint64_t align256(uint64_t v) { return (v + 255) & ~uint64_t(255); }

Geo make_geo(int nx, int ny, int nz, int px, int py, int pz) {
    Geo g;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.px = px; g.py = py; g.pz = pz;
    g.sy = nx;
    g.sz = (int64_t)nx * ny;
    return g;
}

// core/mgfor/mg_setup.f90:225-307 on one rank: halve x and y until one of
// them equals 2; z (interior) halves alongside.  A level is restricted only
// when all its sizes are even, so that its cells pair up exactly
// (core/mgfor/basicoperators.f90:32-60).  The reference has no rule for a
// level with an odd size (its restriction drops a cell, and on one rank a
// size 3 never reaches 2): here the hierarchy stops at that level, which is
// then solved directly (nyl_coarse.hip).  Grids the reference halves down to
// x or y = 2 get exactly the reference's hierarchy.
int build_hierarchy(nyl_ctx *c) {
    int x = c->d.nx, y = c->d.ny, z = c->d.nz;
    int n = 0;
    c->exact = 0;
    c->lev[n++].g = make_geo(x, y, z, c->g.px, c->g.py, c->g.pz);
    while (!(x == 2 || y == 2)) {
        if ((x & 1) || (y & 1) || (z & 1)) {
            if (n == 1)
                return fail(NYL_E_SIZE,
                            "grid %dx%dx%d: the multigrid needs at least one halving "
                            "(nx, ny and nz even)", c->d.nx, c->d.ny, c->d.nz);
            const Geo &g = c->lev[n - 1].g;
            if (!coarse_fits(g))
                return fail(NYL_E_SIZE,
                            "grid %dx%dx%d: the coarsest multigrid level %dx%dx%d has an odd "
                            "size and more than %d cells along x or %d along y or z (direct "
                            "coarse solve)",
                            c->d.nx, c->d.ny, c->d.nz, x, y, z, CS_MAXX, CS_MAX);
            c->exact = 1;
            break;
        }
        x /= 2; y /= 2; z /= 2;
        if (n >= kMaxLevels) return fail(NYL_E_SIZE, "too many multigrid levels");
        c->lev[n++].g = make_geo(x, y, z, c->g.px, c->g.py, c->g.pz);
    }
    c->nlev = n;
    return NYL_OK;
}

void build_tabs(MGTabs &t) {
    for (int d = 0; d <= 6; d++) {
        t.diag[d] = (double)d;
        t.idiag[d] = d > 0 ? 1.0 / (double)d : 0.0;
    }
    for (int m = 0; m < 8; m++) {
        const int mx = m & 1, my = (m >> 1) & 1, mz = (m >> 2) & 1;
        t.pcoef[m] = 1.0 / (double)((3 + mx) * (3 + my) * (3 + mz));
    }
}


void build_layout(nyl_ctx *c) {
    const uint64_t es = c->mixed ? sizeof(float) : sizeof(double);
    const uint64_t fb = align256((uint64_t)c->ncell * es);
    const uint64_t fb64 = align256((uint64_t)c->ncell * sizeof(double));
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) { uint64_t o = off; off += align256(bytes); return o; };
    for (int s = 0; s < 2; s++)
        for (int f = 0; f < 4; f++) c->state[s][f] = take(fb);
    for (int s = 0; s < 3; s++)
        for (int f = 0; f < 4; f++) c->ds[s][f] = take(fb);
    for (int f = 0; f < 4; f++) c->sb[f] = take(fb);
    c->ke = take(fb);
    for (int f = 0; f < 3; f++) c->vor[f] = take(fb);
    for (int f = 0; f < 3; f++) c->uc[f] = take(fb);
    c->p = take(fb64);
    for (int l = 0; l < c->nlev; l++) {
        const uint64_t lb = (uint64_t)c->lev[l].g.nx * c->lev[l].g.ny * c->lev[l].g.nz *
                            (lev64(c, l) ? sizeof(double) : sizeof(float));
        c->lev[l].x = take(lb);
        c->lev[l].b = take(lb);
        c->lev[l].y = take(lb);
    }
    if (c->mixed) {
        c->mgx = take(fb64);
        c->mgb = take(fb64);
    } else {
        c->mgx = c->lev[0].x;
        c->mgb = c->lev[0].b;
    }
    c->partials = take(NYL_RED_BLOCKS * sizeof(double));
    c->result = take(8 * sizeof(double));
    c->cs_tab = c->exact ? take(coarse_table_doubles() * sizeof(double)) : 0;
    for (int q = 0; q < c->ntr; q++) {
        c->tst[0][q] = take(fb);
        c->tst[1][q] = take(fb);
        for (int k = 0; k < 3; k++) c->tds[k][q] = take(fb);
        c->tsb[q] = take(fb);
    }
    c->ws_bytes = off;
}

int check_ready(const nyl_ctx *c) {
    if (!c) return fail(NYL_E_ARG, "null context");
    if (!c->ws) return fail(NYL_E_STATE, "workspace not bound (call nyl_bind_workspace)");
    return NYL_OK;
}

int prof_drain(nyl_ctx *c) {
    if (c->prof_used == 0) return NYL_OK;
    HIPCHK(hipEventSynchronize(c->prof_ev[2 * c->prof_used - 1]));
    for (int q = 0; q < c->prof_used; q++) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->prof_ev[2 * q], c->prof_ev[2 * q + 1]));
        c->prof_ms += ms;
    }
    c->prof_used = 0;
    return NYL_OK;
}

// Launch through `fn`; when the kernel class is being profiled, the first
// kernel `fn` launches records its start and end on a pair of events (the
// kernel's own duration) and its algorithmic bytes are counted.
constexpr int kFineLevel = 0x100;   // launch works on the finest multigrid level

template <typename Fn>
int timed_launch(nyl_ctx *c, int cls, double bytes, Fn fn) {
    // a level-1 sweep (two-sweep pass) also belongs to the NYL_K_MG_SWEEP_FINE
    // (NYL_K_MG_SWEEP2_FINE) class
    const int base = cls & 0xff;
    const bool fine = (cls & kFineLevel) != 0;
    const bool on = c->prof_cls == base ||
                    (c->prof_cls == NYL_K_MG_SWEEP_FINE && fine && base == NYL_K_MG_SWEEP) ||
                    (c->prof_cls == NYL_K_MG_SWEEP2_FINE && fine && base == NYL_K_MG_SWEEP2);
    if (on) {
        if (c->prof_used * 2 + 2 > c->prof_nev) {
            const int cap = (int)(sizeof(c->prof_ev) / sizeof(c->prof_ev[0]));
            if (c->prof_nev + 2 <= cap) {
                HIPCHK(hipEventCreate(&c->prof_ev[c->prof_nev]));
                HIPCHK(hipEventCreate(&c->prof_ev[c->prof_nev + 1]));
                c->prof_nev += 2;
            } else {
                int r = prof_drain(c);
                if (r) return r;
            }
        }
        g_launch_events = {c->prof_ev[2 * c->prof_used], c->prof_ev[2 * c->prof_used + 1]};
    }
    const hipError_t e = fn();
    const bool timed = on && g_launch_events.start == nullptr;
    g_launch_events = {nullptr, nullptr};
    HIPCHK(e);
    if (timed) {
        c->prof_used++;
        c->prof_launches++;
        c->prof_bytes += bytes;
    }
    return NYL_OK;
}

#define LAUNCH(cls, bytes, ...)                                                        \
    do {                                                                               \
        int r_ = timed_launch(c, (cls), (double)(bytes), [&]() { return (__VA_ARGS__); }); \
        if (r_) return r_;                                                             \
    } while (0)

double cells(const Geo &g) { return (double)g.nx * g.ny * g.nz; }

// The results reach the host through coherent pinned memory: a one-thread
// kernel stores them there and then a sequence number, and the host polls
// that number instead of a stream synchronisation plus a copy (each host
// round trip of the multigrid loop is on the critical path).  After a
// bounded poll the host falls back to a blocking synchronisation, which also
// reports a failed kernel.
// The final sum that writes the last of the next n results publishes them
// itself (launch_final_sum), one launch fewer per read; read_results then
// only polls.
void arm_publish(nyl_ctx *c, int n) {
    Publish &r = g_publish;
    r.src = reinterpret_cast<const double *>(c->ws + c->result);
    r.n = n;
    r.dst = c->host_dev;
    r.flag = c->host_dev ? reinterpret_cast<uint32_t *>(c->host_dev + 8) : nullptr;
    r.seq = ++c->pub_seq;
    r.consumed = 0;
}

int read_results(nyl_ctx *c, double *out, int n) {
    const double *src = reinterpret_cast<const double *>(c->ws + c->result);
    Publish &r = g_publish;
    const bool armed = r.consumed && r.dst && r.dst == c->host_dev && r.src == src && r.n == n;
    const uint32_t seq = armed ? r.seq : ++c->pub_seq;
    r = Publish{};
    volatile uint32_t *flag = reinterpret_cast<volatile uint32_t *>(c->host_out + 8);
    if (!armed)
        HIPCHK(launch_publish(src, n, c->host_dev, reinterpret_cast<uint32_t *>(c->host_dev + 8),
                              seq, c->stream));
    const auto t0 = std::chrono::steady_clock::now();
    for (uint64_t it = 1; *flag != seq; it++) {
        if ((it & 1023) == 0 &&
            std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) {
            HIPCHK(hipStreamSynchronize(c->stream));
            if (*flag != seq) return fail(NYL_E_HIP, "result of the device reduction not published");
            break;
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    const volatile double *v = c->host_out;
    for (int q = 0; q < n; q++) out[q] = v[q];
    return NYL_OK;
}

int read_result(nyl_ctx *c, double *out) { return read_results(c, out, 1); }

template <typename T>
StageArgs<T> stage_args(nyl_ctx *c) {
    StageArgs<T> a;
    const int s = c->cur, t = 1 - c->cur;
    a.b = F<T>(c, c->state[s][0]); a.u = F<T>(c, c->state[s][1]);
    a.v = F<T>(c, c->state[s][2]); a.w = F<T>(c, c->state[s][3]);
    a.bn = F<T>(c, c->state[t][0]); a.un = F<T>(c, c->state[t][1]);
    a.vn = F<T>(c, c->state[t][2]); a.wn = F<T>(c, c->state[t][3]);
    a.ox = F<T>(c, c->vor[0]); a.oy = F<T>(c, c->vor[1]); a.oz = F<T>(c, c->vor[2]);
    a.ke = F<T>(c, c->ke);
    a.idx2 = c->d.idx2; a.idy2 = c->d.idy2; a.idz2 = c->d.idz2;
    a.d0b = F<T>(c, c->ds[0][0]); a.d0u = F<T>(c, c->ds[0][1]);
    a.d0v = F<T>(c, c->ds[0][2]); a.d0w = F<T>(c, c->ds[0][3]);
    a.d1b = F<T>(c, c->ds[1][0]); a.d1u = F<T>(c, c->ds[1][1]);
    a.d1v = F<T>(c, c->ds[1][2]); a.d1w = F<T>(c, c->ds[1][3]);
    a.tb = a.tu = a.tv = a.tw = nullptr;
    a.sbb = F<T>(c, c->sb[0]); a.sbu = F<T>(c, c->sb[1]);
    a.sbv = F<T>(c, c->sb[2]); a.sbw = F<T>(c, c->sb[3]);
    a.dz = c->d.dz;
    a.c = 0.0;
    a.c2 = 0.0;
    a.vcx = a.vcy = a.vcz = 0.0;
    a.visc = 0;
    a.flags = 0;
    a.upw_a = c->d.upwind_a;
    a.upw_vf = c->d.upwind_vf;
    return a;
}

template <typename T>
void set_tendency_out(nyl_ctx *c, StageArgs<T> &a, int slot) {
    a.tb = F<T>(c, c->ds[slot][0]); a.tu = F<T>(c, c->ds[slot][1]);
    a.tv = F<T>(c, c->ds[slot][2]); a.tw = F<T>(c, c->ds[slot][3]);
}

// One smooth() of level l: the two fsmoother3d sweeps x -> y -> x.
// xc_level >= 0 first adds the prolongation from that coarse level.
// pre_done: the first sweep already ran into y (speculatively, with the
// convergence test of the previous V-cycle, or from e = 0 in the outer
// residual of the mixed solve).
// addx (level 0 of the mixed solve): the last sweep adds its result to the
// FP64 solution instead of storing it (tiled levels), or an add follows.
// Post-smoothing of the tiled levels as one two-sweep pass after the
// prolongation into y: src and b are read once for two sweeps (256^3:
// 37.8 us FP32 and 93.4 us FP64 against 2 x 31.6 and 2 x 80 us).  The pass
// can also add the second sweep's result to an FP64 solution.
#ifndef NYL_SWEEP2
#define NYL_SWEEP2 1
#endif

// Levels 1 .. nlev-2 run the fused level kernels (down: second pre-sweep +
// restriction; up: prolongation + both post-sweeps), whose up-leg leaves the
// level's correction in y.
#ifndef NYL_FUSED_LEVELS
#define NYL_FUSED_LEVELS 1
#endif
#ifndef NYL_FUSED_MAX_CELLS
#define NYL_FUSED_MAX_CELLS 32768
#endif
bool fused_level(const nyl_ctx *c, int l) {
    return NYL_FUSED_LEVELS && l >= 1 && l <= c->nlev - 2 &&
           cells(c->lev[l].g) <= NYL_FUSED_MAX_CELLS &&
           fused_fits(c->lev[l].g, c->lev[l + 1].g);
}

// offset of the buffer that holds level l's correction after its up-leg
uint64_t sol_off(const nyl_ctx *c, int l) {
    return fused_level(c, l) ? c->lev[l].y : c->lev[l].x;
}

template <typename Tf>
int smooth_level_t(nyl_ctx *c, int l, int xc_level, bool pre_done, double *addx,
                   bool sol_in_y) {
    Level &f = c->lev[l];
    const double om = c->d.mg_omega;
    Tf *x = F<Tf>(c, f.x), *b = F<Tf>(c, f.b), *y = F<Tf>(c, f.y);
    const int SWC = NYL_K_MG_SWEEP | (l == 0 ? kFineLevel : 0);
    const double by = 3 * cells(f.g) * sizeof(Tf);
    const bool tiled = sweep25_fits(f.g);
    if (xc_level >= 0) {
        Level &q = c->lev[xc_level];
        const uint64_t qx = sol_off(c, xc_level);
        const bool same = lev64(c, xc_level) == std::is_same<Tf, double>::value;
        if (sol_in_y) {
            // the level's solution is in y (both pre-sweeps in one pass):
            // y += P(xc) in place, then both post-sweeps y -> x
            if (!(NYL_SWEEP2 && same && restrict25_fits(f.g) && tiled))
                return fail(NYL_E_STATE, "level solution in y without the two-sweep pass");
            Tf *yv = F<Tf>(c, f.y);
            LAUNCH(NYL_K_MG_PROLONG, 2.125 * cells(f.g) * sizeof(Tf),
                   launch_mg_prolong25<Tf, Tf>(f.g, q.g, yv, F<Tf>(c, qx), c->tabs, c->stream));
            const double b2 = addx ? (2 * sizeof(Tf) + 2 * sizeof(double)) * cells(f.g) : by;
            LAUNCH(NYL_K_MG_SWEEP2 | (l == 0 ? kFineLevel : 0), b2,
                   launch_mg_sweep2x25<Tf>(f.g, yv, x, b, om, c->tabs, c->stream, addx));
            return NYL_OK;
        }
        if (NYL_FUSE_PROLONG && tiled && same) {
            // prolongation fused into the first sweep: x + P(xc) -> y, then y -> x
            const double bp = (3 + 0.125) * cells(f.g) * sizeof(Tf);
            LAUNCH(SWC, bp,
                   launch_mg_sweep25<Tf>(f.g, x, y, b, om, c->tabs, nullptr, nullptr, c->stream,
                                         &q.g, F<Tf>(c, qx)));
            LAUNCH(SWC, by, launch_mg_sweep25<Tf>(f.g, y, x, b, om, c->tabs, nullptr,
                                                              nullptr, c->stream));
            return NYL_OK;
        }
        const double bp = 2.125 * cells(f.g) * sizeof(Tf);
        if (NYL_SWEEP2 && same && restrict25_fits(f.g) && tiled && !pre_done) {
            // y = x + P(xc), then both post-smoothing sweeps y -> x in one
            // pass (with addx: the second sweep's result added to the FP64 x)
            Tf *yv = F<Tf>(c, f.y);
            LAUNCH(NYL_K_MG_PROLONG, bp,
                   launch_mg_prolong25<Tf, Tf>(f.g, q.g, x, F<Tf>(c, qx), c->tabs, c->stream, yv));
            const double b2 = addx ? (2 * sizeof(Tf) + 2 * sizeof(double)) * cells(f.g) : by;
            LAUNCH(NYL_K_MG_SWEEP2 | (l == 0 ? kFineLevel : 0), b2,
                   launch_mg_sweep2x25<Tf>(f.g, yv, x, b, om, c->tabs, c->stream, addx));
            return NYL_OK;
        }
        if (same && restrict25_fits(f.g)) {
            LAUNCH(NYL_K_MG_PROLONG, bp,
                   launch_mg_prolong25<Tf, Tf>(f.g, q.g, x, F<Tf>(c, qx), c->tabs, c->stream));
        } else if (lev64(c, xc_level)) {
            LAUNCH(NYL_K_MG_PROLONG, bp,
                   launch_mg_prolong<Tf, double>(f.g, q.g, x, F<double>(c, qx), c->tabs,
                                                 c->stream));
        } else if constexpr (std::is_same<Tf, float>::value) {
            LAUNCH(NYL_K_MG_PROLONG, bp,
                   launch_mg_prolong<float, float>(f.g, q.g, x, F<float>(c, qx), c->tabs,
                                                   c->stream));
        } else {
            return fail(NYL_E_STATE, "FP64 level above an FP32 level");
        }
    }
    if (tiled) {
        if (!pre_done)
            LAUNCH(SWC, by, launch_mg_sweep25<Tf>(f.g, x, y, b, om, c->tabs, nullptr,
                                                              nullptr, c->stream));
        if (addx) {
            const double bx = (2 * sizeof(Tf) + 2 * sizeof(double)) * cells(f.g);
            LAUNCH(SWC, bx, launch_mg_sweep25<Tf>(f.g, y, x, b, om, c->tabs, nullptr,
                                                              nullptr, c->stream, nullptr,
                                                              nullptr, addx));
        } else {
            LAUNCH(SWC, by, launch_mg_sweep25<Tf>(f.g, y, x, b, om, c->tabs, nullptr,
                                                              nullptr, c->stream));
        }
        return NYL_OK;
    }
    if (!pre_done)
        LAUNCH(SWC, by, launch_mg_sweep<Tf>(f.g, x, y, b, om, c->tabs, c->stream));
    LAUNCH(SWC, by, launch_mg_sweep<Tf>(f.g, y, x, b, om, c->tabs, c->stream));
    if (addx) {
        if constexpr (std::is_same<Tf, float>::value) {
            LAUNCH(NYL_K_MG_UPDATE, (2 * sizeof(double) + sizeof(float)) * cells(f.g),
                   launch_add_correction<float>(f.g, addx, x, c->stream));
        } else {
            return fail(NYL_E_STATE, "FP64 inner level");
        }
    }
    return NYL_OK;
}

int smooth_level(nyl_ctx *c, int l, int xc_level, bool pre_done = false,
                 double *addx = nullptr, bool sol_in_y = false) {
    return lev64(c, l) ? smooth_level_t<double>(c, l, xc_level, pre_done, addx, sol_in_y)
                       : smooth_level_t<float>(c, l, xc_level, pre_done, addx, sol_in_y);
}

// residual of level l restricted into level l + 1, with level l+1's first
// smoothing sweep from x = 0 written to its y (from_y: level l's solution is
// in its y)
int restrict_level(nyl_ctx *c, int l, bool from_y = false) {
    Level &f = c->lev[l], &q = c->lev[l + 1];
    const uint64_t fx = from_y ? f.y : f.x;   // the level's solution
    const bool f64 = lev64(c, l), c64 = lev64(c, l + 1);
    const double om = c->d.mg_omega;
    const double bytes = 2.25 * cells(f.g) * (f64 ? sizeof(double) : sizeof(float));
    if (f64 == c64 && restrict25_fits(f.g)) {
        if (f64)
            LAUNCH(NYL_K_MG_RESTRICT, bytes,
                   launch_mg_restrict25<double, double>(f.g, q.g, F<double>(c, fx),
                                                        F<double>(c, f.b), F<double>(c, q.b),
                                                        F<double>(c, q.x), c->tabs, c->stream,
                                                        F<double>(c, q.y), om));
        else
            LAUNCH(NYL_K_MG_RESTRICT, bytes,
                   launch_mg_restrict25<float, float>(f.g, q.g, F<float>(c, fx),
                                                      F<float>(c, f.b), F<float>(c, q.b),
                                                      F<float>(c, q.x), c->tabs, c->stream,
                                                      F<float>(c, q.y), om));
    } else if (f64 && c64) {
        LAUNCH(NYL_K_MG_RESTRICT, bytes,
               launch_mg_resid_restrict<double, double>(f.g, q.g, F<double>(c, fx),
                                                        F<double>(c, f.b), F<double>(c, q.b),
                                                        F<double>(c, q.x), c->tabs, c->stream,
                                                        F<double>(c, q.y), om));
    } else if (!f64 && c64) {
        LAUNCH(NYL_K_MG_RESTRICT, bytes,
               launch_mg_resid_restrict<float, double>(f.g, q.g, F<float>(c, fx),
                                                       F<float>(c, f.b), F<double>(c, q.b),
                                                       F<double>(c, q.x), c->tabs, c->stream,
                                                       F<double>(c, q.y), om));
    } else if (!f64 && !c64) {
        LAUNCH(NYL_K_MG_RESTRICT, bytes,
               launch_mg_resid_restrict<float, float>(f.g, q.g, F<float>(c, fx),
                                                      F<float>(c, f.b), F<float>(c, q.b),
                                                      F<float>(c, q.x), c->tabs, c->stream,
                                                      F<float>(c, q.y), om));
    } else {
        return fail(NYL_E_STATE, "FP64 level above an FP32 level");
    }
    return NYL_OK;
}

// level l's second pre-sweep (y -> x) and the restriction into level l + 1
// (b and the first sweep from zero into y) in one launch
int fused_down(nyl_ctx *c, int l) {
    Level &f = c->lev[l], &q = c->lev[l + 1];
    const bool f64 = lev64(c, l), c64 = lev64(c, l + 1);
    const double om = c->d.mg_omega;
    const double es = f64 ? sizeof(double) : sizeof(float);
    const double bytes = 3 * cells(f.g) * es + 2 * cells(q.g) * (c64 ? 8.0 : 4.0);
    if (f64 && c64)
        LAUNCH(NYL_K_MG_SWEEP, bytes,
               launch_mg_down<double, double>(f.g, q.g, F<double>(c, f.y), F<double>(c, f.b),
                                              F<double>(c, f.x), F<double>(c, q.b),
                                              F<double>(c, q.y), om, c->tabs, c->stream));
    else if (!f64 && c64)
        LAUNCH(NYL_K_MG_SWEEP, bytes,
               launch_mg_down<float, double>(f.g, q.g, F<float>(c, f.y), F<float>(c, f.b),
                                             F<float>(c, f.x), F<double>(c, q.b),
                                             F<double>(c, q.y), om, c->tabs, c->stream));
    else if (!f64 && !c64)
        LAUNCH(NYL_K_MG_SWEEP, bytes,
               launch_mg_down<float, float>(f.g, q.g, F<float>(c, f.y), F<float>(c, f.b),
                                            F<float>(c, f.x), F<float>(c, q.b),
                                            F<float>(c, q.y), om, c->tabs, c->stream));
    else
        return fail(NYL_E_STATE, "FP64 level above an FP32 level");
    return NYL_OK;
}

// level l's prolongation from level l + 1 and both post-sweeps, x -> y
int fused_up(nyl_ctx *c, int l) {
    Level &f = c->lev[l], &q = c->lev[l + 1];
    const bool f64 = lev64(c, l), c64 = lev64(c, l + 1);
    const double om = c->d.mg_omega;
    const uint64_t qx = sol_off(c, l + 1);
    const double es = f64 ? sizeof(double) : sizeof(float);
    const double bytes = 3 * cells(f.g) * es + cells(q.g) * (c64 ? 8.0 : 4.0);
    if (f64 && c64)
        LAUNCH(NYL_K_MG_PROLONG, bytes,
               launch_mg_up<double, double>(f.g, q.g, F<double>(c, f.x), F<double>(c, qx),
                                            F<double>(c, f.b), F<double>(c, f.y), om, c->tabs,
                                            c->stream));
    else if (!f64 && c64)
        LAUNCH(NYL_K_MG_PROLONG, bytes,
               launch_mg_up<float, double>(f.g, q.g, F<float>(c, f.x), F<double>(c, qx),
                                           F<float>(c, f.b), F<float>(c, f.y), om, c->tabs,
                                           c->stream));
    else if (!f64 && !c64)
        LAUNCH(NYL_K_MG_PROLONG, bytes,
               launch_mg_up<float, float>(f.g, q.g, F<float>(c, f.x), F<float>(c, qx),
                                          F<float>(c, f.b), F<float>(c, f.y), om, c->tabs,
                                          c->stream));
    else
        return fail(NYL_E_STATE, "FP64 level above an FP32 level");
    return NYL_OK;
}

// the coarsest level's step of a V-cycle: mgfor's smooth() (two sweeps, the
// first already written by the restriction), or the direct solve of a level
// with an odd size
int coarsest(nyl_ctx *c) {
    const int L = c->nlev - 1;
    if (!c->exact) return smooth_level(c, L, -1, true);
    Level &f = c->lev[L];
    CoarseArgs a;
    a.g = f.g;
    a.b = F<double>(c, f.b);
    a.s = F<double>(c, f.y);
    a.x = F<double>(c, f.x);
    a.tab = F<double>(c, c->cs_tab);
    LAUNCH(NYL_K_MG_COARSE, 4 * cells(f.g) * sizeof(double), launch_coarse_solve(a, c->stream));
    return NYL_OK;
}

// The smallest levels (fused, one type, at most NYL_TAIL_MAX_CELLS cells,
// all of them with the coarsest level within the tail kernel's LDS) and the
// coarsest level run in one single-workgroup launch (k_mg_tail): below a
// few hundred cells a level's launches cost more than its work, while one
// workgroup's issue rate makes a larger level cheaper as separate launches.
#ifndef NYL_TAIL
#define NYL_TAIL 1
#endif
#ifndef NYL_TAIL_MAX_CELLS
#define NYL_TAIL_MAX_CELLS TAIL_T
#endif
// first level of the tail, -1 when there is none
int tail_level(const nyl_ctx *c) {
    const int L = c->nlev;
    if (!NYL_TAIL || c->exact || L < 3 || L > NYL_TAIL_MAXLEV) return -1;
    auto lds = [&](int l) {
        const int64_t n = (int64_t)cells(c->lev[l].g) * (lev64(c, l) ? 8 : 4);
        return 3.0 * (double)((n + 15) / 16 * 16);
    };
    double used = lds(L - 1);
    int t0 = -1;
    for (int l = L - 2; l >= 1; l--) {
        used += lds(l);
        if (!fused_level(c, l) || used > NYL_TAIL_LDS || cells(c->lev[l].g) > NYL_TAIL_MAX_CELLS ||
            lev64(c, l) != lev64(c, L - 2))
            break;
        t0 = l;
    }
    return t0;
}

int tail(nyl_ctx *c, int t0) {
    const int L = c->nlev;
    TailArgs a;
    memset(&a, 0, sizeof(a));
    a.ws = c->ws;
    double bytes = 0.0;
    for (int l = 0; l < L; l++) {
        const Level &v = c->lev[l];
        a.lev[l].g = v.g;
        a.lev[l].x = v.x;
        a.lev[l].b = v.b;
        a.lev[l].y = v.y;
        if (l >= t0) bytes += 3.0 * cells(v.g) * (lev64(c, l) ? 8.0 : 4.0);
    }
    a.l0 = t0;
    a.nlev = L;
    a.omega = c->d.mg_omega;
    a.tab = c->tabs;
    const bool t64 = lev64(c, t0), c64 = lev64(c, L - 1);
    if (t64 && c64) LAUNCH(NYL_K_MG_COARSE, bytes, (launch_mg_tail<double, double>(a, c->stream)));
    else if (!t64 && c64) LAUNCH(NYL_K_MG_COARSE, bytes, (launch_mg_tail<float, double>(a, c->stream)));
    else if (!t64 && !c64) LAUNCH(NYL_K_MG_COARSE, bytes, (launch_mg_tail<float, float>(a, c->stream)));
    else return fail(NYL_E_STATE, "FP64 level above an FP32 level");
    return NYL_OK;
}

// the part of a V-cycle below level 1: from level 2's smoothing (its first
// sweep written by the restriction from level 1) down to the coarsest level
// and back up to level 2's post-smoothing
int coarse_cycle(nyl_ctx *c) {
    const int L = c->nlev;
    const int t0 = tail_level(c);
    const int down_end = t0 >= 0 ? t0 : L - 1;   // levels [1, down_end) run here
    int rc;
    for (int l = 1; l < down_end; l++) {
        if (fused_level(c, l)) {
            if ((rc = fused_down(c, l))) return rc;
            continue;
        }
        if ((rc = smooth_level(c, l, -1, true))) return rc;
        if ((rc = restrict_level(c, l))) return rc;
    }
    if (t0 >= 0) {
        if ((rc = tail(c, t0))) return rc;
    } else if ((rc = coarsest(c))) {
        return rc;
    }
    for (int l = down_end - 1; l >= 1; l--) {
        if (fused_level(c, l)) {
            if ((rc = fused_up(c, l))) return rc;
            continue;
        }
        if ((rc = smooth_level(c, l, l + 1))) return rc;
    }
    return NYL_OK;
}

void drop_graphs(nyl_ctx *c) {
    if (c->g_coarse) {
        (void)hipGraphExecDestroy(c->g_coarse);
        c->g_coarse = nullptr;
    }
}

// The coarse part is the same launch sequence every V-cycle: it is captured
// once into a hipGraph and replayed (the per-launch overhead dominates these
// small levels).  Direct launches are used on the null stream, and while a
// kernel class that includes coarse-level launches is being timed.
int coarse_cycle_run(nyl_ctx *c) {
    const int pc = c->prof_cls;
    const bool timed = pc == NYL_K_MG_SWEEP || pc == NYL_K_MG_RESTRICT || pc == NYL_K_MG_PROLONG ||
                       pc == NYL_K_MG_SWEEP2 || pc == NYL_K_MG_COARSE ||
                       pc == NYL_K_MG_SWEEP2_FINE;
    if (!c->use_graphs || !c->stream || timed) return coarse_cycle(c);
    if (!c->g_coarse) {
        hipGraph_t graph;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int prof = c->prof_cls;
        c->prof_cls = 0;                 // no event records inside the graph
        int rc = coarse_cycle(c);
        c->prof_cls = prof;
        hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (rc) return rc;
        HIPCHK(e);
        HIPCHK(hipGraphInstantiate(&c->g_coarse, graph, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(graph));
    }
    HIPCHK(hipGraphLaunch(c->g_coarse, c->stream));
    return NYL_OK;
}

// core/mgfor/solvers.f90:35-55.  pre_done: the level-1 pre-smooth already
// ran into the spare buffer.  addx: the mixed solve's x, which receives the
// level-1 result (x += e).  Below level 1 the first sweep of each level is
// written by the restriction.
// pre: the level-1 pre-smoothing done before the call, 0 none, 1 the first
// sweep (in y), 2 both (in y, by the two-sweep pass of the convergence
// test): then the restriction reads y, the prolongation adds to y in place
// and the post-smoothing pass writes x.
int vcycle(nyl_ctx *c, int pre, double *addx = nullptr) {
    int rc;
    if (pre < 2 && (rc = smooth_level(c, 0, -1, pre == 1))) return rc;
    if ((rc = restrict_level(c, 0, pre == 2))) return rc;
    if (c->nlev > 2) {
        if ((rc = coarse_cycle_run(c))) return rc;
    } else {
        if ((rc = coarsest(c))) return rc;
    }
    return smooth_level(c, 0, 1, false, addx, pre == 2);
}

// The convergence test's residual norm can come with both pre-smoothing
// sweeps of the next V-cycle (k_mg_sweep2t, solution into y): level 1 tiled
// for the two-sweep pass and level 2 of the same type (the post-smoothing
// that follows uses the pass too)
bool spec2_fits(const nyl_ctx *c) {
    const Level &f = c->lev[0];
    return NYL_SWEEP2 && c->nlev >= 2 && sweep25_fits(f.g) && restrict25_fits(f.g) &&
           lev64(c, 0) == lev64(c, 1) &&
           sweep2x_blocks(f.g, lev64(c, 0) ? 8 : 4) <= NYL_RED_BLOCKS;
}

void fill_stats(nyl_ctx *c, nyl_mg_stats *st, int nite, double res, double normb) {
    c->total_cycles += nite;
    if (st) {
        st->nite = nite;
        st->res = res;
        st->normb = normb;
        st->total_cycles = c->total_cycles;
    }
}

// core/mgfor/solvers.f90:8-33 with normresidual (core/mgfor/operators.f90:81-96).
// When level 1 uses the tiled sweeps, the residual norm of x is computed by
// the pre-smoothing of the next V-cycle, launched before the test (both
// sweeps in the two-sweep pass, or the first one): if the test stops the
// solve, that output (in y) is unused and x is untouched.
int mg_solve_fp64(nyl_ctx *c, nyl_mg_stats *st) {
    Level &f = c->lev[0];
    double *part = F<double>(c, c->partials), *res_dev = F<double>(c, c->result);
    const bool spec = sweep25_fits(f.g) && sweep25_blocks(f.g) <= NYL_RED_BLOCKS;
    const bool spec2 = spec2_fits(c);
    int rc;
    double normb, rn, res;
    arm_publish(c, 1);
    HIPCHK(launch_norm2<double>(f.g, F<double>(c, f.b), part, res_dev, c->stream));
    if ((rc = read_result(c, &normb))) return rc;
    int pending = 0;   // pre-smoothing sweeps of the next V-cycle already done
    auto normresidual = [&](double *out) -> int {
        pending = 0;
        if (normb > 0.0) {
            arm_publish(c, 1);
            if (spec2) {
                LAUNCH(NYL_K_MG_SWEEP2 | kFineLevel, 3 * cells(f.g) * sizeof(double),
                       launch_mg_sweep2x25<double>(f.g, F<double>(c, f.x), F<double>(c, f.y),
                                                   F<double>(c, f.b), c->d.mg_omega, c->tabs,
                                                   c->stream, nullptr, part, res_dev));
                pending = 2;
            } else if (spec) {
                LAUNCH(NYL_K_MG_SWEEP | kFineLevel, 3 * cells(f.g) * sizeof(double),
                       launch_mg_sweep25<double>(f.g, F<double>(c, f.x), F<double>(c, f.y),
                                                 F<double>(c, f.b), c->d.mg_omega, c->tabs,
                                                 part, res_dev, c->stream));
                pending = 1;
            } else {
                LAUNCH(NYL_K_RESNORM, 2 * cells(f.g) * sizeof(double),
                       launch_resnorm2<double>(f.g, F<double>(c, f.x), F<double>(c, f.b), c->tabs,
                                               part, res_dev, c->stream));
            }
            int r2 = read_result(c, &rn);
            if (r2) return r2;
            *out = rn / normb;
        } else {
            *out = 0.0;
        }
        return NYL_OK;
    };
    if ((rc = normresidual(&res))) return rc;
    int nite = 0;
    for (;;) {
        if (res < c->d.mg_tol) break;
        if ((rc = vcycle(c, pending))) return rc;
        nite = nite + 1;
        if (nite >= c->d.mg_maxite) break;
        if ((rc = normresidual(&res))) return rc;
    }
    fill_stats(c, st, nite, res, normb);
    return NYL_OK;
}

// Mixed precision: iterative refinement with the stop rule of mgfor's solve.
// The compatibility of the singular Neumann/periodic problem is enforced in
// FP64 first: one pass sums b and b**2, the mean is removed from b inside
// every outer residual and |b - mean|**2 is the norm of the stop test (the
// discrete divergence sums to zero up to rounding, so this changes nothing
// measurable, but a right-hand side with a mean could not converge).  An
// outer step computes r = (b - mean) - A x in
// FP64 and solves A e = r from e = 0 with FP32 V-cycles under mgfor's loop:
// before each further V-cycle the squared residual r - A e = b - A(x + e),
// evaluated in FP32 and summed in FP64, is tested against tol*|b|^2 (fused
// into that V-cycle's pre-smoothing, as in the FP64 path).  Then x += e in FP64
// and the FP64 residual confirms the stop (or starts another outer step).
// The V-cycle count follows mgfor's rule (nite, maxite).
// One inner solve resolves a limited reduction of r: A e of a smooth e
// cancels to ~eps32*(n/pi)^2 of its terms, so below that the FP32 residual
// norm is rounding noise and the FP32 e stops improving.  Hence an outer step
// whose tolerance asks for more than kInnerRed2*|r|^2 tests the inner
// residual against kInnerRed2*|r|^2 instead and runs at most kInnerMax
// V-cycles; the FP64 residual decides.  At mgfor's tolerance (1e-6) the
// tolerance is the target and the V-cycle count is that of the FP64 solve.
constexpr double kInnerRed2 = 1e-6;   // squared reduction of one inner solve
constexpr int kInnerMax = 8;
int mg_solve_mixed(nyl_ctx *c, nyl_mg_stats *st) {
    Level &f = c->lev[0];
    double *part = F<double>(c, c->partials), *res_dev = F<double>(c, c->result);
    double *xo = F<double>(c, c->mgx), *bo = F<double>(c, c->mgb);
    float *e = F<float>(c, f.x), *r = F<float>(c, f.b), *y = F<float>(c, f.y);
    const bool spec = sweep25_fits(f.g) && sweep25_blocks(f.g) <= NYL_RED_BLOCKS;
    const bool spec2 = spec2_fits(c);
    int rc;
    double rn, res;
    // sum of b and of b**2 in FP64: the mean to remove and the norm of b - mean
    double s2[2];
    arm_publish(c, 2);
    HIPCHK(launch_sum_sumsq(f.g, bo, part, res_dev, c->stream));
    if ((rc = read_results(c, s2, 2))) return rc;
    const double ncell = cells(f.g);
    const double shift = s2[0] / ncell;
    double normb = s2[1] - ncell * shift * shift;
    if (normb < 0.0) normb = 0.0;
    // FP64 residual of x into r (and the first sweep from e = 0 into y)
    auto outer = [&](double *out) -> int {
        if (normb > 0.0) {
            arm_publish(c, 1);
            LAUNCH(NYL_K_RESNORM, (2 * sizeof(double) + 2 * sizeof(float)) * cells(f.g),
                   launch_mg_outer_resid<float>(f.g, xo, bo, shift, r, spec2 ? nullptr : y,
                                                c->d.mg_omega, c->tabs, part, res_dev,
                                                c->stream, false));
            int r2 = read_result(c, &rn);
            if (r2) return r2;
            *out = rn / normb;
        } else {
            *out = 0.0;
        }
        return NYL_OK;
    };
    // FP32 residual r - A e (speculative pre-smoothing of the next V-cycle:
    // both sweeps into y, or the first)
    auto inner = [&](double *out) -> int {
        arm_publish(c, 1);
        if (spec2) {
            LAUNCH(NYL_K_MG_SWEEP2 | kFineLevel, 3 * cells(f.g) * sizeof(float),
                   launch_mg_sweep2x25<float>(f.g, e, y, r, c->d.mg_omega, c->tabs, c->stream,
                                              nullptr, part, res_dev));
        } else if (spec) {
            LAUNCH(NYL_K_MG_SWEEP | kFineLevel, 3 * cells(f.g) * sizeof(float),
                   launch_mg_sweep25<float>(f.g, e, y, r, c->d.mg_omega, c->tabs, part, res_dev,
                                            c->stream));
        } else {
            LAUNCH(NYL_K_RESNORM, 2 * cells(f.g) * sizeof(float),
                   launch_resnorm2<float>(f.g, e, r, c->tabs, part, res_dev, c->stream));
        }
        int r2 = read_result(c, &rn);
        if (r2) return r2;
        *out = rn / normb;
        return NYL_OK;
    };
    if ((rc = outer(&res))) return rc;
    // mgfor's loop shape (core/mgfor/solvers.f90:22-29): test, V-cycle,
    // count, stop at maxite (so maxite = 0 still runs one V-cycle)
    int nite = 0;
    bool at_max = false;
    while (!(res < c->d.mg_tol) && !at_max) {
        // both pre-smoothing sweeps from e = 0 in one pass into y, or the first
        // one (written by the outer residual)
        int pending = 1;
        if (spec2) {
            // (not in the fine-level class of the roofline: it reads one field)
            LAUNCH(NYL_K_MG_SWEEP2, 2 * cells(f.g) * sizeof(float),
                   launch_mg_sweep2x25<float>(f.g, nullptr, y, r, c->d.mg_omega, c->tabs,
                                              c->stream));
            pending = 2;
        }
        const bool rel = res * kInnerRed2 > c->d.mg_tol;
        const double target = rel ? res * kInnerRed2 : c->d.mg_tol;
        for (int k = 1;; k++) {
            if ((rc = vcycle(c, pending))) return rc;
            nite = nite + 1;
            at_max = nite >= c->d.mg_maxite;
            if (at_max || (rel && k >= kInnerMax)) break;
            double rin;
            if ((rc = inner(&rin))) return rc;
            pending = spec2 ? 2 : spec ? 1 : 0;
            if (rin < target) break;
        }
        LAUNCH(NYL_K_MG_UPDATE, (2 * sizeof(double) + sizeof(float)) * cells(f.g),
               launch_add_correction<float>(f.g, xo, e, c->stream));
        if ((rc = outer(&res))) return rc;
    }
    fill_stats(c, st, nite, res, normb);
    return NYL_OK;
}

int mg_solve(nyl_ctx *c, nyl_mg_stats *st) {
    return c->mixed ? mg_solve_mixed(c, st) : mg_solve_fp64(c, st);
}

template <typename T>
DiagArgs<T> diag_args(nyl_ctx *c) {
    DiagArgs<T> a;
    const int s = c->cur;
    a.u = F<T>(c, c->state[s][1]); a.v = F<T>(c, c->state[s][2]);
    a.w = F<T>(c, c->state[s][3]);
    a.Ux = F<T>(c, c->uc[0]); a.Uy = F<T>(c, c->uc[1]); a.Uz = F<T>(c, c->uc[2]);
    a.ox = F<T>(c, c->vor[0]); a.oy = F<T>(c, c->vor[1]); a.oz = F<T>(c, c->vor[2]);
    a.ke = F<T>(c, c->ke);
    a.idx2 = c->d.idx2; a.idy2 = c->d.idy2; a.idz2 = c->d.idz2;
    a.fpar = c->d.fparameter > 0.0 ? c->d.fparameter : 0.0;
    return a;
}

template <typename T>
int apply_op_t(nyl_ctx *c, int op) {
    const int s = c->cur;
    T *u = F<T>(c, c->state[s][1]), *v = F<T>(c, c->state[s][2]), *w = F<T>(c, c->state[s][3]);
    switch (op) {
    case NYL_OP_U_FROM_U:
        HIPCHK(launch_diag<T>(c->g, diag_args<T>(c), 1, c->stream));
        break;
    case NYL_OP_DIV:
        LAUNCH(NYL_K_DIV, (3 * sizeof(T) + sizeof(double)) * cells(c->g),
               launch_div<T, double>(c->g, u, v, w, c->d.idx2, c->d.idy2, c->d.idz2,
                                     F<double>(c, c->mgb), c->stream));
        break;
    case NYL_OP_CORRECT:
        LAUNCH(NYL_K_CORRECT, (6 * sizeof(T) + sizeof(double)) * cells(c->g),
               launch_correct<T, double>(c->g, u, v, w, F<double>(c, c->mgx), c->d.dx2,
                                         c->stream));
        break;
    case NYL_OP_VORTICITY:
        HIPCHK(launch_diag<T>(c->g, diag_args<T>(c), 2, c->stream));
        break;
    case NYL_OP_KINENERGY:
        HIPCHK(launch_diag<T>(c->g, diag_args<T>(c), 4, c->stream));
        break;
    case NYL_OP_PRESSURE:
        HIPCHK(launch_pressure<double, double>(c->g, F<double>(c, c->mgx), c->d.dx2,
                                               F<double>(c, c->p), c->stream));
        break;
    default:
        return fail(NYL_E_ARG, "unknown op %d", op);
    }
    return NYL_OK;
}

int apply_op(nyl_ctx *c, int op) {
    return c->mixed ? apply_op_t<float>(c, op) : apply_op_t<double>(c, op);
}

// core/model_les.py:97-127 (halo fills are implicit in the wrapped reads)
template <typename T>
int diagnose_t(nyl_ctx *c, nyl_mg_stats *st) {
    int rc;
    if ((rc = apply_op_t<T>(c, NYL_OP_DIV))) return rc;
    if ((rc = mg_solve(c, st))) return rc;
    if ((rc = apply_op_t<T>(c, NYL_OP_CORRECT))) return rc;
    // vorticity and ke (not in the linear model, core/model_les.py:122-127);
    // U is not stored (the stage kernel recomputes it, NYL_OP_U_FROM_U fills
    // it on demand)
    if (c->d.model != NYL_MODEL_LINEAR)
        LAUNCH(NYL_K_DIAG, 7 * cells(c->g) * sizeof(T),
               launch_diag<T>(c->g, diag_args<T>(c), 6, c->stream));
    return NYL_OK;
}

int diagnose(nyl_ctx *c, nyl_mg_stats *st) {
    return c->mixed ? diagnose_t<float>(c, st) : diagnose_t<double>(c, st);
}

// words of HBM traffic per cell of each stage-kernel mode: reads b,u,v,w,
// vor(3), ke plus the mode's extra inputs; writes the mode's outputs
double stage_words(int mode) {
    switch (mode) {
    case 0: return 8 + 4;            // tendencies only
    case 1: return 8 + 8;            // + next state + ds0
    case 2: return 12 + 8;           // + ds0 read
    case 3: return 16 + 4;           // + ds0, ds1 read
    case 4: return 12 + 8;           // + sb read; next state + sb written
    case 5: return 8 + 4 + 4;        // next state read and written
    case 6: return 8 + 8;            // next state + sb written
    default: return 0;
    }
}

// mode st of the stage kernel; slot >= 0 also writes the tendencies there
template <typename T>
int stage_t(nyl_ctx *c, int st, int slot, double coef, double coef2, bool last) {
    StageArgs<T> a = stage_args<T>(c);
    if (slot >= 0) set_tendency_out<T>(c, a, slot);
    a.c = coef;
    a.c2 = coef2;
    if (last && c->d.viscosity != 0.0) {   // core/model_les.py:140-141
        a.visc = 1;
        a.vcx = c->d.viscosity * c->d.idx2;
        a.vcy = c->d.viscosity * c->d.idy2;
        a.vcz = c->d.viscosity * c->d.idz2;
    }
    if (c->d.model == NYL_MODEL_LINEAR) a.flags |= NYL_SF_NO_VF;
    if (c->d.model == NYL_MODEL_EULER) a.flags |= NYL_SF_NO_B;
    const double words = stage_words(st);
    LAUNCH(NYL_K_STAGE, words * cells(c->g) * sizeof(T), launch_stage<T>(c->g, a, st, c->stream));
    // passive tracers: the same right-hand side and update, one launch each;
    // Euler3d advects none of them (zero tendencies)
    const int s = c->cur, t = 1 - c->cur;
    for (int q = 0; q < c->ntr; q++) {
        if (c->d.model == NYL_MODEL_EULER) {
            HIPCHK(launch_zero<T>(F<T>(c, c->tds[slot][q]), c->ncell, c->stream));
            continue;
        }
        TracerArgs<T> ta;
        ta.q = F<T>(c, c->tst[s][q]);
        ta.qn = F<T>(c, c->tst[t][q]);
        ta.u = a.u; ta.v = a.v; ta.w = a.w;
        ta.d0 = F<T>(c, c->tds[0][q]);
        ta.d1 = F<T>(c, c->tds[1][q]);
        ta.tout = slot >= 0 ? F<T>(c, c->tds[slot][q]) : nullptr;
        ta.sb = F<T>(c, c->tsb[q]);
        ta.idx2 = c->d.idx2; ta.idy2 = c->d.idy2; ta.idz2 = c->d.idz2;
        ta.c = coef;
        ta.c2 = coef2;
        ta.upw = c->d.upwind_a;
        LAUNCH(NYL_K_STAGE, (words / 4.0) * cells(c->g) * sizeof(T),
               launch_tracer_stage<T>(c->g, ta, st, c->stream));
    }
    return NYL_OK;
}

// the stage combination of mode st from the stored tendencies of every
// prognostic field (b, u, v, w, tracers): slot 0 for modes 1 and 4..6, slots
// 0..1 for mode 2 and 0..2 for mode 3
template <typename T>
int update_t(nyl_ctx *c, int st, double coef, double coef2) {
    const int s = c->cur, t = 1 - c->cur;
    auto one = [&](uint64_t cur, uint64_t nxt, uint64_t d0, uint64_t d1, uint64_t d2,
                   uint64_t sb) -> int {
        UpdateArgs<T> a;
        a.s = F<T>(c, cur);
        a.sn = F<T>(c, nxt);
        a.ds0 = F<T>(c, d0);
        a.ds1 = F<T>(c, d1);
        a.ds2 = F<T>(c, d2);
        a.sb = F<T>(c, sb);
        a.c = coef;
        a.c2 = coef2;
        HIPCHK(launch_update<T>(c->ncell, a, st, c->stream));
        return NYL_OK;
    };
    int rc;
    for (int f = 0; f < 4; f++)
        if ((rc = one(c->state[s][f], c->state[t][f], c->ds[0][f], c->ds[1][f], c->ds[2][f],
                      c->sb[f])))
            return rc;
    for (int q = 0; q < c->ntr; q++)
        if ((rc = one(c->tst[s][q], c->tst[t][q], c->tds[0][q], c->tds[1][q], c->tds[2][q],
                      c->tsb[q])))
            return rc;
    return NYL_OK;
}

// last: the time scheme's rhs call with last = True (RK3 stage 3, Euler
// forward, the LFAM3 corrector), which adds the viscosity
int stage(nyl_ctx *c, int st, double coef, double coef2, bool last) {
    int rc;
    if (c->d.model != NYL_MODEL_LES || c->d.upwind_a || c->d.upwind_vf) {
        // model variants and the linear upwind branches: the tendencies
        // into the slot the stage combination
        // reads last (0 for modes 1 and 4..6, 1 for mode 2, 2 for mode 3),
        // then the combination
        const int slot = st == 2 ? 1 : (st == 3 ? 2 : 0);
        rc = c->mixed ? stage_t<float>(c, 0, slot, 0.0, 0.0, last)
                      : stage_t<double>(c, 0, slot, 0.0, 0.0, last);
        if (rc) return rc;
        rc = c->mixed ? update_t<float>(c, st, coef, coef2) : update_t<double>(c, st, coef, coef2);
        if (rc) return rc;
        if (st != 5) c->cur = 1 - c->cur;
        return NYL_OK;
    }
    const int slot = (st == 1 || st == 2) ? st - 1 : -1;
    rc = c->mixed ? stage_t<float>(c, st, slot, coef, coef2, last)
                  : stage_t<double>(c, st, slot, coef, coef2, last);
    if (rc) return rc;
    if (st != 5) c->cur = 1 - c->cur;
    return NYL_OK;
}

template <typename T>
int reduce_t(nyl_ctx *c, int what) {
    const int s = c->cur;
    double *part = F<double>(c, c->partials), *res = F<double>(c, c->result);
    T *b = F<T>(c, c->state[s][0]), *u = F<T>(c, c->state[s][1]), *v = F<T>(c, c->state[s][2]),
      *w = F<T>(c, c->state[s][3]);
    switch (what) {
    case NYL_RED_MAXU2:
        HIPCHK(launch_maxu2<T>(c->g, u, v, w, c->d.idx2, c->d.idy2, c->d.idz2, part, res,
                               c->stream));
        break;
    case NYL_RED_KE:
        HIPCHK(launch_kesum<T>(c->g, u, v, w, c->d.idx2, c->d.idy2, c->d.idz2, part, res,
                               c->stream));
        break;
    case NYL_RED_BSUM:
        HIPCHK(launch_sum<T>(c->g, b, part, res, c->stream));
        break;
    case NYL_RED_NONFINITE:
        HIPCHK(launch_nonfinite<T>(c->g, b, u, v, w, part, res, c->stream));
        break;
    default:
        return fail(NYL_E_ARG, "unknown reduction %d", what);
    }
    return NYL_OK;
}

}  // namespace

extern "C" {

int nyl_abi_version(void) { return NYL_ABI_VERSION; }

const char *nyl_last_error(void) { return g_err; }

int nyl_ctx_create(const nyl_desc *desc, nyl_ctx **out) {
    if (!desc || !out) return fail(NYL_E_ARG, "null argument");
    *out = nullptr;
    const nyl_desc &d = *desc;
    if (d.precision != NYL_PREC_FP64 && d.precision != NYL_PREC_MIXED)
        return fail(NYL_E_ARG, "unknown precision mode %d", d.precision);
    const int n3[3] = {d.nx, d.ny, d.nz};
    const int p3[3] = {d.periodic_x, d.periodic_y, d.periodic_z};
    for (int a = 0; a < 3; a++) {
        // wall axes need n >= 6 for the flux1d closures (core/weno.f90:116-150)
        if (n3[a] < 8)
            return fail(NYL_E_SIZE, "axis %d has %d cells; at least 8 are needed", a, n3[a]);
        if (p3[a] != 0 && p3[a] != 1) return fail(NYL_E_ARG, "periodic flags must be 0 or 1");
    }
    if (!(d.dx > 0 && d.dy > 0 && d.dz > 0)) return fail(NYL_E_ARG, "grid spacings must be > 0");
    // the stage kernel addresses a field with 32-bit byte offsets
    const int64_t esz = d.precision == NYL_PREC_FP64 ? 8 : 4;
    if ((int64_t)d.nx * d.ny * d.nz * esz >= ((int64_t)1 << 32))
        return fail(NYL_E_SIZE, "grid of %lld cells: a field must stay below 4 GiB "
                    "(%lld cells in FP64, %lld in mixed precision)",
                    (long long)d.nx * d.ny * d.nz, ((long long)1 << 29) - 1,
                    ((long long)1 << 30) - 1);
    if (d.mg_maxite < 0) return fail(NYL_E_ARG, "mg_maxite must be >= 0");
    if (d.n_tracers < 0 || d.n_tracers > NYL_MAX_TRACERS)
        return fail(NYL_E_ARG, "n_tracers must be 0..%d", NYL_MAX_TRACERS);
    if (!(d.viscosity >= 0.0)) return fail(NYL_E_ARG, "viscosity must be >= 0");
    if (d.model < NYL_MODEL_LES || d.model > NYL_MODEL_EULER)
        return fail(NYL_E_ARG, "model must be NYL_MODEL_LES, _LINEAR or _EULER");
    for (int o : {d.upwind_a, d.upwind_vf})
        if (o != 0 && o != 1 && o != 3 && o != 5)
            return fail(NYL_E_ARG, "linear upwind orders are 1, 3 or 5 (0: WENO)");
    if (d.upwind_a || d.upwind_vf) {
        // the reference's closures overlap (and read outside the line) on
        // wall lines shorter than 6 cells
        const int n[3] = {d.nx, d.ny, d.nz}, p[3] = {d.periodic_x, d.periodic_y, d.periodic_z};
        for (int q = 0; q < 3; q++)
            if (n[q] < 6 && !p[q])
                return fail(NYL_E_SIZE, "linear upwind needs 6 or more cells along a wall axis");
    }
    nyl_ctx *c = new nyl_ctx();
    memset(c, 0, sizeof(*c));
    c->d = d;
    c->ntr = d.n_tracers;
    c->mixed = d.precision == NYL_PREC_MIXED;
    c->g = make_geo(d.nx, d.ny, d.nz, d.periodic_x, d.periodic_y, d.periodic_z);
    if (c->mixed && !outer_resid_fits(c->g)) {
        delete c;
        return fail(NYL_E_SIZE, "grid %dx%dx%d is too wide for the mixed-precision solve",
                    d.nx, d.ny, d.nz);
    }
    c->ncell = (int64_t)d.nx * d.ny * d.nz;
    int rc = build_hierarchy(c);
    if (rc) { delete c; return rc; }
    build_tabs(c->tabs);
    build_layout(c);
    c->stream = nullptr;
    c->use_graphs = 1;
    c->g_coarse = nullptr;
    *out = c;
    return NYL_OK;
}

int nyl_ctx_destroy(nyl_ctx *c) {
    if (!c) return NYL_OK;
    drop_graphs(c);
    for (int q = 0; q < c->prof_nev; q++) (void)hipEventDestroy(c->prof_ev[q]);
    if (c->host_out) (void)hipHostFree(c->host_out);
    delete c;
    return NYL_OK;
}

int nyl_workspace_bytes(const nyl_ctx *c, uint64_t *bytes) {
    if (!c || !bytes) return fail(NYL_E_ARG, "null argument");
    *bytes = c->ws_bytes;
    return NYL_OK;
}

int nyl_set_stream(nyl_ctx *c, void *stream) {
    if (!c) return fail(NYL_E_ARG, "null context");
    c->stream = (hipStream_t)stream;
    return NYL_OK;
}

int nyl_set_graphs(nyl_ctx *c, int on) {
    if (!c) return fail(NYL_E_ARG, "null context");
    c->use_graphs = on ? 1 : 0;
    if (!on) drop_graphs(c);
    return NYL_OK;
}

int nyl_bind_workspace(nyl_ctx *c, void *dptr, uint64_t bytes) {
    if (!c || !dptr) return fail(NYL_E_ARG, "null argument");
    drop_graphs(c);
    if (bytes < c->ws_bytes)
        return fail(NYL_E_ARG, "workspace of %llu bytes is smaller than the %llu needed",
                    (unsigned long long)bytes, (unsigned long long)c->ws_bytes);
    if (((uintptr_t)dptr) & 255) return fail(NYL_E_ARG, "workspace must be 256-byte aligned");
    if (!c->host_out) {
        HIPCHK(hipHostMalloc((void **)&c->host_out, 16 * sizeof(double),
                             hipHostMallocMapped | hipHostMallocCoherent));
        HIPCHK(hipHostGetDevicePointer((void **)&c->host_dev, c->host_out, 0));
        *reinterpret_cast<volatile uint32_t *>(c->host_out + 8) = 0;
        c->pub_seq = 0;
    }
    c->ws = (char *)dptr;
    HIPCHK(launch_fill((double *)c->ws, (int64_t)(c->ws_bytes / sizeof(double)), 0.0, c->stream));
    if (c->exact) {
        std::vector<double> tab(coarse_table_doubles());
        coarse_tables(c->lev[c->nlev - 1].g, tab.data());
        HIPCHK(hipMemcpyAsync(c->ws + c->cs_tab, tab.data(), tab.size() * sizeof(double),
                              hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->cur = 0;
    c->total_cycles = 0;
    c->lf_first = 1;
    return NYL_OK;
}

int nyl_field_offset(const nyl_ctx *c, int field, uint64_t *off, int64_t *nelem, int32_t *dtype) {
    if (!c || !off) return fail(NYL_E_ARG, "null argument");
    uint64_t o;
    const int s = c->cur;
    bool f64 = !c->mixed;   // model fields follow the precision mode
    switch (field) {
    case NYL_F_B: case NYL_F_U: case NYL_F_V: case NYL_F_W: o = c->state[s][field - NYL_F_B]; break;
    case NYL_F_KE: o = c->ke; break;
    case NYL_F_VORX: case NYL_F_VORY: case NYL_F_VORZ: o = c->vor[field - NYL_F_VORX]; break;
    case NYL_F_UCX: case NYL_F_UCY: case NYL_F_UCZ: o = c->uc[field - NYL_F_UCX]; break;
    case NYL_F_P: o = c->p; f64 = true; break;
    case NYL_F_DIV: o = c->mgb; f64 = true; break;
    case NYL_F_MGX: o = c->mgx; f64 = true; break;
    case NYL_F_MGB: o = c->mgb; f64 = true; break;
    case NYL_F_MGY: o = c->lev[0].y; f64 = lev64(c, 0); break;
    default:
        if (field >= NYL_F_DS0_B && field <= NYL_F_DS2_W) {
            const int q = field - NYL_F_DS0_B;
            o = c->ds[q / 4][q % 4];
            break;
        }
        if (field >= NYL_F_TRACER(0) && field < NYL_F_TRACER(c->ntr)) {
            o = c->tst[s][field - NYL_F_TRACER(0)];
            break;
        }
        if (field >= NYL_F_DS_TRACER(0, 0) && field < NYL_F_DS_TRACER(3, 0) &&
            (field - NYL_F_DS_TRACER(0, 0)) % NYL_MAX_TRACERS < c->ntr) {
            const int q = field - NYL_F_DS_TRACER(0, 0);
            o = c->tds[q / NYL_MAX_TRACERS][q % NYL_MAX_TRACERS];
            break;
        }
        return fail(NYL_E_ARG, "unknown field %d", field);
    }
    *off = o;
    if (nelem) *nelem = c->ncell;
    if (dtype) *dtype = f64 ? NYL_DT_F64 : NYL_DT_F32;
    return NYL_OK;
}

int nyl_mg_levels(const nyl_ctx *c, int32_t *n) {
    if (!c || !n) return fail(NYL_E_ARG, "null argument");
    *n = c->nlev;
    return NYL_OK;
}

int nyl_mg_level_shape(const nyl_ctx *c, int level, int32_t *nx, int32_t *ny, int32_t *nz) {
    if (!c || !nx || !ny || !nz) return fail(NYL_E_ARG, "null argument");
    if (level < 0 || level >= c->nlev) return fail(NYL_E_ARG, "level %d out of range", level);
    *nx = c->lev[level].g.nx; *ny = c->lev[level].g.ny; *nz = c->lev[level].g.nz;
    return NYL_OK;
}

int nyl_rhs(nyl_ctx *c, int slot, int last) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (slot < 0 || slot > 2) return fail(NYL_E_ARG, "slot must be 0..2");
    return c->mixed ? stage_t<float>(c, 0, slot, 0.0, 0.0, last != 0)
                    : stage_t<double>(c, 0, slot, 0.0, 0.0, last != 0);
}

int nyl_stage_update(nyl_ctx *c, int mode, double coef, double coef2) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (mode < 1 || mode > 6) return fail(NYL_E_ARG, "mode must be 1..6");
    rc = c->mixed ? update_t<float>(c, mode, coef, coef2) : update_t<double>(c, mode, coef, coef2);
    if (rc) return rc;
    c->cur = 1 - c->cur;   // mode 5 wrote over the saved 'now' state, which becomes current
    return NYL_OK;
}

int nyl_rk3_stage(nyl_ctx *c, int st, double coef) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (st < 1 || st > 3) return fail(NYL_E_ARG, "stage must be 1..3");
    return stage(c, st, coef, 0.0, st == 3);
}

int nyl_diagnose(nyl_ctx *c, nyl_mg_stats *st) {
    int rc = check_ready(c);
    if (rc) return rc;
    return diagnose(c, st);
}

int nyl_step_rk3(nyl_ctx *c, double dt, nyl_mg_stats *st) {
    int rc = check_ready(c);
    if (rc) return rc;
    nyl_mg_stats s1, acc;
    memset(&acc, 0, sizeof(acc));
    const double coef[3] = {dt, dt / 4., dt / 12.};
    for (int k = 1; k <= 3; k++) {
        if ((rc = stage(c, k, coef[k - 1], 0.0, k == 3))) return rc;
        if ((rc = diagnose(c, &s1))) return rc;
        acc.nite += s1.nite;
        acc.res = s1.res;
        acc.normb = s1.normb;
    }
    acc.total_cycles = c->total_cycles;
    if (st) *st = acc;
    return NYL_OK;
}

int nyl_step_ef(nyl_ctx *c, double dt, nyl_mg_stats *st) {
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = stage(c, 1, dt, 0.0, true))) return rc;
    return diagnose(c, st);
}

int nyl_step_lfam3(nyl_ctx *c, double dt, nyl_mg_stats *st) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (c->lf_first) {
        if ((rc = stage(c, 6, dt, 0.0, false))) return rc;
        c->lf_first = 0;
        return diagnose(c, st);
    }
    nyl_mg_stats s1, s2;
    // predictor: the current buffer keeps sn, the next one receives s
    if ((rc = stage(c, 4, 0.0, 2. * dt, false))) return rc;
    if ((rc = diagnose(c, &s1))) return rc;
    // corrector: rhs at n+1/2, then s = sn + dt*ds written over sn
    if ((rc = stage(c, 5, dt, 0.0, true))) return rc;
    c->cur = 1 - c->cur;
    if ((rc = diagnose(c, &s2))) return rc;
    if (st) {
        *st = s2;
        st->nite = s1.nite + s2.nite;
    }
    return NYL_OK;
}

int nyl_lfam3_reset(nyl_ctx *c) {
    if (!c) return fail(NYL_E_ARG, "null context");
    c->lf_first = 1;
    return NYL_OK;
}

int nyl_mg_solve(nyl_ctx *c, nyl_mg_stats *st) {
    int rc = check_ready(c);
    if (rc) return rc;
    return mg_solve(c, st);
}

int nyl_apply(nyl_ctx *c, int op) {
    int rc = check_ready(c);
    if (rc) return rc;
    return apply_op(c, op);
}

int nyl_reduce(nyl_ctx *c, int what, double *out) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!out) return fail(NYL_E_ARG, "null argument");
    arm_publish(c, 1);
    if ((rc = c->mixed ? reduce_t<float>(c, what) : reduce_t<double>(c, what))) return rc;
    return read_result(c, out);
}

int nyl_prof_start(nyl_ctx *c, int kernel_class) {
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = prof_drain(c))) return rc;
    c->prof_cls = kernel_class;
    c->prof_ms = 0.0;
    c->prof_bytes = 0.0;
    c->prof_launches = 0;
    return NYL_OK;
}

int nyl_prof_stop(nyl_ctx *c, double *total_ms, int64_t *launches, double *alg_bytes) {
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = prof_drain(c))) return rc;
    if (total_ms) *total_ms = c->prof_ms;
    if (launches) *launches = c->prof_launches;
    if (alg_bytes) *alg_bytes = c->prof_bytes;
    c->prof_cls = 0;
    return NYL_OK;
}

int nyl_debug_mg(nyl_ctx *c, int what) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (c->mixed) return fail(NYL_E_STATE, "nyl_debug_mg works on FP64 contexts");
    Level &f = c->lev[0];
    double *x = F<double>(c, f.x), *b = F<double>(c, f.b), *y = F<double>(c, f.y);
    double *part = F<double>(c, c->partials), *res_dev = F<double>(c, c->result);
    switch (what) {
    case 0:
        HIPCHK(launch_mg_sweep<double>(f.g, x, y, b, c->d.mg_omega, c->tabs, c->stream));
        break;
    case 1:
        if (!sweep25_fits(f.g)) return fail(NYL_E_SIZE, "level 1 does not fit the tiled sweep");
        HIPCHK(launch_mg_sweep25<double>(f.g, x, y, b, c->d.mg_omega, c->tabs, nullptr, nullptr,
                                         c->stream));
        break;
    case 2:
        if (!sweep25_fits(f.g)) return fail(NYL_E_SIZE, "level 1 does not fit the tiled sweep");
        HIPCHK(launch_mg_sweep25<double>(f.g, x, y, b, c->d.mg_omega, c->tabs, part, res_dev,
                                         c->stream));
        break;
    case 3:
        HIPCHK(launch_resnorm2<double>(f.g, x, b, c->tabs, part, res_dev, c->stream));
        break;
    case 4:
    case 5:
    case 6:
        if (!sweep25_fits(f.g)) return fail(NYL_E_SIZE, "level 1 does not fit the tiled sweep");
        if (what == 5 && sweep2x_blocks(f.g, 8) > NYL_RED_BLOCKS)
            return fail(NYL_E_SIZE, "too many blocks for the residual norm");
        HIPCHK(launch_mg_sweep2x25<double>(f.g, what == 6 ? nullptr : x, y, b, c->d.mg_omega,
                                           c->tabs, c->stream, nullptr,
                                           what == 5 ? part : nullptr,
                                           what == 5 ? res_dev : nullptr));
        break;
    default:
        return fail(NYL_E_ARG, "unknown debug operation %d", what);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return NYL_OK;
}

int nyl_reduce_last(nyl_ctx *c, double *out) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!out) return fail(NYL_E_ARG, "null argument");
    return read_result(c, out);
}

int nyl_sync(nyl_ctx *c) {
    if (!c) return fail(NYL_E_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return NYL_OK;
}

}  // extern "C"
