// nyl_internal.h -- host-side launcher declarations shared by the
// translation units of libnyles_hip.so (not part of the public ABI).
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "nyl_device.h"

namespace nyl {

// Events of the kernel-class profiler: when set, the next kernel launched
// through NYL_LAUNCH records its own start and end on them
// (hipExtLaunchKernelGGL places the timestamps in the dispatch itself, so the
// time is the kernel's alone, without the gaps an event record on each side
// of the launch adds) and clears them.
struct LaunchEvents {
    hipEvent_t start, stop;
};
inline thread_local LaunchEvents g_launch_events = {nullptr, nullptr};

#define NYL_LAUNCH(kernel, grid, block, shmem, stream, ...)                                 \
    do {                                                                                    \
        if (nyl::g_launch_events.start) {                                                   \
            hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream,                       \
                                  nyl::g_launch_events.start, nyl::g_launch_events.stop, 0, \
                                  __VA_ARGS__);                                             \
            nyl::g_launch_events.start = nullptr;                                           \
        } else {                                                                            \
            hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);            \
        }                                                                                   \
    } while (0)

// Arguments of the fused RK stage kernel (rhs + update).
template <typename T>
struct StageArgs {
    const T *b, *u, *v, *w;          // current state
    T *bn, *un, *vn, *wn;            // next state (stage update target)
    const T *ox, *oy, *oz;           // vorticity
    const T *ke;                     // kinetic energy
    const T *d0b, *d0u, *d0v, *d0w;  // tendency slot 0 (read in stages 2, 3)
    const T *d1b, *d1u, *d1v, *d1w;  // tendency slot 1 (read in stage 3)
    T *tb, *tu, *tv, *tw;            // tendency output (modes 0, 1, 2)
    T *sbb, *sbu, *sbv, *sbw;        // LFAM3 'before' state (modes 4, 6)
    double idx2, idy2, idz2;         // U = u*idx2 ... (core/cov_to_contra.py:15-17), recomputed
    double dz;                       // for cff = 0.5*dz (core/fortran_bernoulli.f90:47)
    double c;                        // RK coefficient dt, dt/4., dt/12.
    double c2;                       // LFAM3: 2.*dt (mode 4)
    // viscosity at the last stage (core/model_les.py:140-141): when visc is
    // set, du += vc_d * delta_d[delta_d[u]] for d = x, y, z in turn, with
    // vc_d = viscosity * ids2[d] (core/viscosity.py:3-9)
    double vcx, vcy, vcz;
    int visc;
    // model variant, mode 0 only (the variants run tendencies + update):
    // bit 0 no vortex force (linear model), bit 1 no advection of b and no
    // buoyancy (Euler3d)
    int flags;
    // linear upwind orders (1, 3, 5) of the tracer advection and of the
    // vortex force, mode 0 only; 0: WENO (the branch the reference runs)
    int upw_a, upw_vf;
};
#define NYL_SF_NO_VF 1
#define NYL_SF_NO_B 2

// mode 0: tendencies only; 1..3: RK3_SSP stage update;
// 4: LFAM3 predictor; 5: LFAM3 corrector (next = next + c*ds);
// 6: Euler step that also copies the state into the LFAM3 'before' state
template <typename T>
hipError_t launch_stage(const Geo &g, const StageArgs<T> &a, int mode, hipStream_t s);

// One passive tracer (core/model_les.py:36-43): its flux divergence with the
// current U (core/tracer.py:59-72, as for b) and the stage update of the
// launch's mode, the same arithmetic as b's in the stage kernel.
template <typename T>
struct TracerArgs {
    const T *q;                      // current tracer
    T *qn;                           // next (stage update target)
    const T *u, *v, *w;              // current velocity (U = u*idx2 ...)
    const T *d0, *d1;                // tendency slots 0, 1 (stages 2, 3)
    T *tout;                         // tendency output (modes 0, 1, 2)
    T *sb;                           // LFAM3 'before' state (modes 4, 6)
    double idx2, idy2, idz2;
    double c, c2;
    int upw;                         // linear upwind order (mode 0 only), 0: WENO
};
template <typename T>
hipError_t launch_tracer_stage(const Geo &g, const TracerArgs<T> &a, int mode, hipStream_t s);

// The stage combination of a time scheme from stored tendencies, for one
// field (core/timescheme.py:113-221): used when the tendencies are modified
// between the right-hand side and the update (user forcing).  ds0 is the
// stage's own tendency for modes 1 and 4..6; mode 2 reads ds0, ds1 and mode
// 3 ds0, ds1, ds2 (RK3 slots).
template <typename T>
struct UpdateArgs {
    const T *s;                      // current
    T *sn;                           // next
    const T *ds0, *ds1, *ds2;
    T *sb;                           // LFAM3 'before'
    double c, c2;
};
template <typename T>
hipError_t launch_update(int64_t n, const UpdateArgs<T> &a, int mode, hipStream_t s);

template <typename T>
struct DiagArgs {
    const T *u, *v, *w;
    T *Ux, *Uy, *Uz;
    T *ox, *oy, *oz;
    T *ke;
    double idx2, idy2, idz2;
    double fpar;                     // Coriolis parameter added to vor_k when > 0
};
// what: bit 0 = U, bit 1 = vorticity, bit 2 = ke
template <typename T>
hipError_t launch_diag(const Geo &g, const DiagArgs<T> &a, int what, hipStream_t s);

template <typename T, typename TO>
struct DivArgs {
    const T *u, *v, *w;
    double idx2, idy2, idz2;
    TO *out;
};
template <typename T, typename TX>
struct CorrectArgs {
    T *u, *v, *w;
    const TX *x;
    double dx2;
};

// div of U (computed from u) into out (core/projection.py:15-35)
template <typename T, typename TO>
hipError_t launch_div(const Geo &g, const T *u, const T *v, const T *w, double idx2,
                      double idy2, double idz2, TO *out, hipStream_t s);

// u -= grad(p), p = x*dx2 (core/projection.py:82-87)
template <typename T, typename TX>
hipError_t launch_correct(const Geo &g, T *u, T *v, T *w, const TX *x, double dx2,
                          hipStream_t s);

// p = x*dx2
template <typename TX, typename T>
hipError_t launch_pressure(const Geo &g, const TX *x, double dx2, T *p, hipStream_t s);

// ---- multigrid (core/mgfor) ----------------------------------------------
struct MGTabs {
    double idiag[7];     // 1/diag for diag = 0..6 (0 -> 0)
    double diag[7];      // diag as a real
    double pcoef[8];     // 1/((3+mx)(3+my)(3+mz)) indexed mx + 2 my + 4 mz
};

template <typename T>
hipError_t launch_mg_sweep(const Geo &g, const T *src, T *dst, const T *b, double omega,
                           const MGTabs &tab, hipStream_t s);
// copy n doubles to coherent host memory, then store seq into flag
hipError_t launch_publish(const double *src, int n, double *dst, uint32_t *flag, uint32_t seq,
                          hipStream_t s);
// A request to publish n device results at src (copy to dst in coherent host
// memory, then seq into flag at system scope) from the final-sum launch that
// writes the last of them, src[n-1]: host code that polls for the results
// right after the launches arms it, saving the separate k_publish launch.
// The final-sum launch that matches the request by address takes it
// (consumed); an unmatched request is published by k_publish as before.
struct Publish {
    const double *src;
    int n;
    double *dst;
    uint32_t *flag;
    uint32_t seq;
    int consumed;
};
inline thread_local Publish g_publish = {};
// the second stage of a deterministic sum (one block over np partials),
// publishing the result when it completes the armed request
hipError_t launch_final_sum(const double *partials, int np, double *out, hipStream_t s);
// Fused level kernels (below the finest level).  down: x = sweep(y), then
// the residual b - A x restricted into bc and the coarse first sweep from
// zero into yc.  up: y = sweep(sweep(x + P xc)); x and xc are only read.
bool fused_fits(const Geo &gf, const Geo &gc);
template <typename Tf, typename Tc>
hipError_t launch_mg_down(const Geo &gf, const Geo &gc, const Tf *y, const Tf *b, Tf *x, Tc *bc,
                          Tc *yc, double omega, const MGTabs &tab, hipStream_t s);
template <typename Tf, typename Tc>
hipError_t launch_mg_up(const Geo &gf, const Geo &gc, const Tf *x, const Tc *xc, const Tf *b,
                        Tf *y, double omega, const MGTabs &tab, hipStream_t s);
// residual + restriction from a level of type Tf to the next of type Tc;
// yc non-null: also the coarse level's first sweep from x = 0 into yc (coarse
// x is then left as is, the second sweep overwrites it), else xc = 0
template <typename Tf, typename Tc>
hipError_t launch_mg_resid_restrict(const Geo &gf, const Geo &gc, const Tf *x, const Tf *b,
                                    Tc *bc, Tc *xc, const MGTabs &tab, hipStream_t s,
                                    Tc *yc = nullptr, double omega = 0.0);
template <typename Tf, typename Tc>
hipError_t launch_mg_prolong(const Geo &gf, const Geo &gc, Tf *xf, const Tc *xc,
                             const MGTabs &tab, hipStream_t s);

// mixed precision defect correction: outer FP64 residual of x (b - shift)
// into the inner right-hand side r and the inner first sweep y, with the sum
// of r**2 in dev_out[0] and, with_bsq (only where outer_bsq_fits), the sum of
// b**2 in dev_out[1]; x += e
bool outer_resid_fits(const Geo &g);
bool outer_bsq_fits(const Geo &g);
template <typename TI>
hipError_t launch_mg_outer_resid(const Geo &g, const double *x, const double *b, double shift,
                                 TI *r, TI *y, double omega, const MGTabs &tab, double *partials,
                                 double *dev_out, hipStream_t s, bool with_bsq);
template <typename TI>
hipError_t launch_add_correction(const Geo &g, double *x, const TI *e, hipStream_t s);
// dev_out[0] = sum v, dev_out[1] = sum v**2
hipError_t launch_sum_sumsq(const Geo &g, const double *v, double *partials, double *dev_out,
                            hipStream_t s);
// dev_out[0] = sum of the FP64 kinetic-energy field of u, v, w
template <typename T>
hipError_t launch_kesum(const Geo &g, const T *u, const T *v, const T *w, double idx2,
                        double idy2, double idz2, double *partials, double *dev_out,
                        hipStream_t s);

// One fsmoother3d sweep with 2.5-D LDS tiling (levels with nx % 64 == 0 and
// ny % 8 == 0).  partials: also reduce the squared residual of src into
// dev_out (needs sweep25_blocks(g) <= NYL_RED_BLOCKS partials).
bool sweep25_fits(const Geo &g);
int sweep25_blocks(const Geo &g);
// gc/xc non-null: the sweep reads src + prolongation of xc (level gc) in
// place of src (fused k_mg_prolong; src itself is not modified).
// addx non-null: the result is added to the FP64 addx instead of stored.
template <typename T>
hipError_t launch_mg_sweep25(const Geo &g, const T *src, T *dst, const T *b, double omega,
                             const MGTabs &tab, double *partials, double *dev_out,
                             hipStream_t s, const Geo *gc = nullptr, const T *xc = nullptr,
                             double *addx = nullptr);
// residual + restriction and prolongation with 2.5-D tiling (fine level
// with sweep25_fits and even nz)
bool restrict25_fits(const Geo &g);
template <typename Tf, typename Tc>
hipError_t launch_mg_prolong25(const Geo &gf, const Geo &gc, Tf *xf, const Tc *xc,
                               const MGTabs &tab, hipStream_t s, Tf *yf = nullptr);
// The smallest levels of a V-cycle in one workgroup (k_mg_tail): levels
// l0 .. nlev-2 fused (exact halving), the coarsest smoothed (no direct
// solve).  x, b, y: each level's fields in the workspace ws (of level l0,
// b and y are read and y written); lx, lb, ly: their places in the kernel's
// LDS, where all the tail's levels live (tail_layout).
#define NYL_TAIL_MAXLEV 16
#define TAIL_T 512
#define NYL_TAIL_LDS 65536
struct TailLevel {
    Geo g;
    uint64_t x, b, y;
    uint32_t lx, lb, ly;
};
struct TailArgs {
    char *ws;
    TailLevel lev[NYL_TAIL_MAXLEV];
    int l0, nlev;
    double omega;
    MGTabs tab;
};
// LDS places of levels l0 .. nlev-1 (es: bytes per value of the fused
// levels, esc: of the coarsest); returns the bytes used
inline uint64_t tail_layout(TailArgs &a, int es, int esc) {
    uint64_t o = 0;
    for (int l = a.l0; l < a.nlev; l++) {
        const TailLevel &v = a.lev[l];
        const uint64_t n = (uint64_t)v.g.nx * v.g.ny * v.g.nz * (l == a.nlev - 1 ? esc : es);
        const uint64_t step = (n + 15) / 16 * 16;
        a.lev[l].lx = (uint32_t)o;
        a.lev[l].lb = (uint32_t)(o + step);
        a.lev[l].ly = (uint32_t)(o + 2 * step);
        o += 3 * step;
    }
    return o;
}
template <typename T, typename TC>
hipError_t launch_mg_tail(const TailArgs &a, hipStream_t s);

// two consecutive tiled sweeps src -> (tmp) -> dst in one pass (src != dst;
// src null: from src = 0, nothing read); addx: the result added to that
// FP64 field instead of stored; partials: also the residual norm of src into
// dev_out (needs sweep2x_blocks(g, es) <= NYL_RED_BLOCKS partials)
template <typename T>
hipError_t launch_mg_sweep2x25(const Geo &g, const T *src, T *dst, const T *b, double omega,
                               const MGTabs &tab, hipStream_t s, double *addx = nullptr,
                               double *partials = nullptr, double *dev_out = nullptr);
int sweep2x_blocks(const Geo &g, int es);
template <typename Tf, typename Tc>
hipError_t launch_mg_restrict25(const Geo &gf, const Geo &gc, const Tf *x, const Tf *b, Tc *bc,
                                Tc *xc, const MGTabs &tab, hipStream_t s, Tc *yc = nullptr,
                                double omega = 0.0);

// ---- direct solve of a coarsest level that cannot be halved (odd size) ------
// x = A^-1 b with the constant mode removed, by per-axis eigenvector
// transforms (s: scratch of the level's size; tab: coarse_tables()).
#define CS_MAX 32     // cells along y and z
#define CS_MAXX 64    // cells along x
struct CoarseArgs { Geo g; const double *b; double *s; double *x; const double *tab; };
bool coarse_fits(const Geo &g);               // nx <= CS_MAXX, ny and nz <= CS_MAX
int coarse_table_doubles();
void coarse_tables(const Geo &g, double *tab);   // host: eigenvectors and eigenvalues
hipError_t launch_coarse_solve(const CoarseArgs &a, hipStream_t s);

// ---- ensemble batches -------------------------------------------------------
// One member's share of a batched launch: the geometry of the level the
// launch works on, the member's index in the batch, its number of blocks and
// its first block in the flattened grid (items are sorted by start).  Cells
// of a member are flattened in x-fastest order, one thread per cell.
struct BHead {
    Geo g;
    int member;
    int nblk;
    int64_t start;
};
template <typename T> struct StageItem { BHead h; StageArgs<T> a; };
template <typename T> struct DivItem { BHead h; DivArgs<T, double> a; };
template <typename T> struct CorrItem { BHead h; CorrectArgs<T, double> a; };
template <typename T> struct DiagItem { BHead h; DiagArgs<T> a; };
// xo non-null: the sweep result is added to xo (FP64) instead of stored in dst
template <typename T> struct SweepItem { BHead h; const T *src; T *dst; const T *b; double omega;
                                         double *xo; };
template <typename Tf, typename Tc>
struct RestrItem { BHead h; Geo gc; const Tf *x; const Tf *b; Tc *bc; Tc *xc; };
template <typename Tf, typename Tc>
struct ProlItem { BHead h; Geo gc; Tf *xf; const Tc *xc; };
// per-member partial sums: v (norm2 / sum and sum of squares) or x, b (residual)
template <typename T> struct RedItem { BHead h; const T *v; const T *b; double *partials; };
struct OuterItem { BHead h; const double *x; const double *b; float *r; float *y; double omega;
                   double *partials; };
// h.nblk = nz (passes 0, 2) or ny (pass 1) of the member's coarsest level
struct CoarseItem { BHead h; CoarseArgs a; };

// Device-side multigrid control of a batch: the stop rule of mgfor's solve
// (core/mgfor/solvers.f90:17-30) evaluated per member on the device.
struct BMG {
    double *normb, *res, *shift, *tol;
    int *nite, *maxite, *active, *any;
    int *acc;              // V-cycles summed over the stages of a step
    const double *dt;
};
// one member's final reduction: partials[0..nparts) of the member
struct FinalItem { const double *partials; int nparts; int member; double ncell; };
#define NYL_FIN_NORMB 0   // normb = sum
#define NYL_FIN_INIT 1    // res = sum/normb, active = !(res < tol), nite = 0
#define NYL_FIN_STEP 2    // nite++; active = nite < maxite && !(res = sum/normb < tol)
#define NYL_FIN_SHIFT 3   // shift = sum/n, normb = sumsq - n shift^2 (two partial sets)

#define NYL_BATCH_TPB 256
int batch_blocks(int64_t ncell);   // blocks of one member at NYL_BATCH_TPB threads

// Tile of the batched stage kernel: tpb cells, either rows of tpb
// consecutive cells of a member (lx < 0) or a brick of 2^lx x 2^ly x
// (tpb >> (lx + ly)) cells.
struct StageTile {
    int tpb, lx, ly;
};
#define NYL_N_TILES 15
extern const StageTile kStageTiles[NYL_N_TILES];
// tiles of one member for a tile shape
int64_t stage_tiles(const Geo &g, const StageTile &t);

template <typename T>
hipError_t launch_b_stage(const StageItem<T> *items, int n, int64_t nblk, int mode, int perio,
                          const StageTile &tile, const double *dt, hipStream_t s);
template <typename T>
hipError_t launch_b_div(const DivItem<T> *items, int n, int64_t nblk, hipStream_t s);
template <typename T>
hipError_t launch_b_correct(const CorrItem<T> *items, int n, int64_t nblk, hipStream_t s);
template <typename T>
hipError_t launch_b_diag(const DiagItem<T> *items, int n, int64_t nblk, int what, hipStream_t s);
template <typename T>
hipError_t launch_b_sweep(const SweepItem<T> *items, int n, int64_t nblk, const MGTabs &tab,
                          const int *active, hipStream_t s);
template <typename Tf, typename Tc>
hipError_t launch_b_restrict(const RestrItem<Tf, Tc> *items, int n, int64_t nblk,
                             const MGTabs &tab, const int *active, hipStream_t s);
template <typename Tf, typename Tc>
hipError_t launch_b_prolong(const ProlItem<Tf, Tc> *items, int n, int64_t nblk,
                            const MGTabs &tab, const int *active, hipStream_t s);
// kind 0: sum v^2; 1: sum of squared residuals of (v = x, b); 2: sum v and sum v^2
// (second set of partials at partials + nblk of the member)
template <typename T>
hipError_t launch_b_reduce(const RedItem<T> *items, int n, int64_t nblk, int kind,
                           const MGTabs &tab, const int *active, hipStream_t s);
// active null: every member (the initial residual of a solve)
hipError_t launch_b_outer_resid(const OuterItem *items, int n, int64_t nblk, const BMG &mg,
                                const MGTabs &tab, const int *active, hipStream_t s);
hipError_t launch_b_final(const FinalItem *items, int n, int kind, const BMG &mg, hipStream_t s);
hipError_t launch_b_coarse(const CoarseItem *items, int n, int64_t nblk, int pass,
                           const int *active, hipStream_t s);
// mode 0: acc = 0; mode 1: acc += nite
hipError_t launch_b_acc(int n, const BMG &mg, int mode, hipStream_t s);

// reductions: partials then one final block; result written to dev_out[0]
#define NYL_RED_BLOCKS 16384
template <typename T>
hipError_t launch_norm2(const Geo &g, const T *v, double *partials, double *dev_out,
                        hipStream_t s);
template <typename T>
hipError_t launch_resnorm2(const Geo &g, const T *x, const T *b, const MGTabs &tab,
                           double *partials, double *dev_out, hipStream_t s);
template <typename T>
hipError_t launch_maxu2(const Geo &g, const T *u, const T *v, const T *w, double idx2,
                        double idy2, double idz2, double *partials, double *dev_out,
                        hipStream_t s);
template <typename T>
hipError_t launch_sum(const Geo &g, const T *v, double *partials, double *dev_out,
                      hipStream_t s);
template <typename T>
hipError_t launch_nonfinite(const Geo &g, const T *a, const T *b, const T *c, const T *d,
                            double *partials, double *dev_out, hipStream_t s);

hipError_t launch_fill(double *p, int64_t n, double v, hipStream_t s);
// p[0..n) = 0 (a kernel, so it may be captured in a graph)
template <typename T>
hipError_t launch_zero(T *p, int64_t n, hipStream_t s);

}  // namespace nyl
