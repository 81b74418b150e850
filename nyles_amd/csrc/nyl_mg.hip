// nyl_mg.hip -- geometric multigrid of the pressure projection (the mgfor
// library of the reference) on gfx950, and the deterministic reductions.
//
// Level arrays are interior-only [z][y][x].  mgfor's coefficient arrays are
// closed forms of the position here (they are what core/mgfor/operators.f90
// :299-457 computes for a box without land): msk = 1 on the interior,
// diag = number of in-domain neighbours, idiag = 1/diag, Rcoef = 1/2 and
// Pcoef = 1/((3+mx)(3+my)(3+mz)) with m* = 1 when the coarse neighbour used
// by the trilinear stencil is in the domain (or periodic), 0 at a wall.
// Out-of-domain neighbours read as zero, which is what mgfor's unfilled wall
// halos hold; periodic neighbours are wrapped, which is what its filled
// halos hold (core/mgfor/mod_halo.f90:235-262).
#include "nyl_internal.h"

namespace nyl {

#define MG_BX 64
#define MG_BY 4

static inline dim3 mg_grid(const Geo &g) {
    return dim3((g.nx + MG_BX - 1) / MG_BX, (g.ny + MG_BY - 1) / MG_BY, g.nz);
}

// Face-neighbour values of a cell, zero outside a wall axis, and the number
// of in-domain neighbours (mgfor's diag).
template <typename T>
struct Nbrs {
    T xm, xp, ym, yp, zm, zp;
    int nwet;
};

template <typename T>
__device__ __forceinline__ Nbrs<T> neighbours(const Geo &g, const T *__restrict__ a, int i, int j,
                                              int k, int64_t c) {
    Nbrs<T> r;
    r.nwet = 0;
    if (g.px || i > 0) { r.xm = a[c - (g.px ? (int64_t)(i - wrapi(i - 1, g.nx)) : 1)]; r.nwet++; } else r.xm = T(0);
    if (g.px || i < g.nx - 1) { r.xp = a[c + (g.px ? (int64_t)(wrapi(i + 1, g.nx) - i) : 1)]; r.nwet++; } else r.xp = T(0);
    if (g.py || j > 0) { r.ym = a[c - (g.py ? (int64_t)(j - wrapi(j - 1, g.ny)) * g.sy : g.sy)]; r.nwet++; } else r.ym = T(0);
    if (g.py || j < g.ny - 1) { r.yp = a[c + (g.py ? (int64_t)(wrapi(j + 1, g.ny) - j) * g.sy : g.sy)]; r.nwet++; } else r.yp = T(0);
    if (g.pz || k > 0) { r.zm = a[c - (g.pz ? (int64_t)(k - wrapi(k - 1, g.nz)) * g.sz : g.sz)]; r.nwet++; } else r.zm = T(0);
    if (g.pz || k < g.nz - 1) { r.zp = a[c + (g.pz ? (int64_t)(wrapi(k + 1, g.nz) - k) * g.sz : g.sz)]; r.nwet++; } else r.zp = T(0);
    return r;
}

// One damped-Jacobi sweep of fsmoother3d (core/mgfor/basicoperators.f90:363-400):
// dst = cff1*src + omega*(sum6(src) - b)*idiag.
template <typename T>
__device__ __forceinline__ T sweep_value(const Geo &g, const T *__restrict__ src,
                                         const T *__restrict__ b, double omega,
                                         const MGTabs &tab, int i, int j, int k, int64_t c) {
    const Nbrs<T> n = neighbours<T>(g, src, i, j, k, c);
    const T S = n.xm + n.xp + n.ym + n.yp + n.zm + n.zp;
    const T om = T(omega), cff1 = T(1.0 - omega);
    return cff1 * src[c] + om * (S - b[c]) * T(tab.idiag[n.nwet]);
}

template <typename T>
__device__ __forceinline__ void sweep_cell(const Geo &g, const T *__restrict__ src,
                                           T *__restrict__ dst, const T *__restrict__ b,
                                           double omega, const MGTabs &tab, int i, int j, int k) {
    const int64_t c = (int64_t)k * g.sz + (int64_t)j * g.sy + i;
    dst[c] = sweep_value<T>(g, src, b, omega, tab, i, j, k, c);
}

template <typename T>
__global__ __launch_bounds__(256) void k_mg_sweep(Geo g, const T *__restrict__ src,
                                                  T *__restrict__ dst, const T *__restrict__ b,
                                                  double omega, MGTabs tab) {
    const Blk3 bk = xcd_block();
    const int i = bk.x * MG_BX + threadIdx.x;
    const int j = bk.y * MG_BY + threadIdx.y;
    const int k = bk.z;
    if (i >= g.nx || j >= g.ny) return;
    sweep_cell<T>(g, src, dst, b, omega, tab, i, j, k);
}

template <typename T>
hipError_t launch_mg_sweep(const Geo &g, const T *src, T *dst, const T *b, double omega,
                           const MGTabs &tab, hipStream_t s) {
    NYL_LAUNCH((k_mg_sweep<T>), mg_grid(g), dim3(MG_BX, MG_BY, 1), 0, s, g, src, dst, b,
                       omega, tab);
    return hipGetLastError();
}

// fresidual3d (core/mgfor/basicoperators.f90:300-323): r = b + diag*x - sum6(x)
template <typename T>
__device__ __forceinline__ T residual_at(const Geo &g, const T *__restrict__ x,
                                         const T *__restrict__ b, int i, int j, int k,
                                         const MGTabs &tab) {
    const int64_t c = (int64_t)k * g.sz + (int64_t)j * g.sy + i;
    const Nbrs<T> n = neighbours<T>(g, x, i, j, k, c);
    return b[c] + T(tab.diag[n.nwet]) * x[c] - (n.xm + n.xp + n.ym + n.yp + n.zm + n.zp);
}

// residual then restriction (core/mgfor/operators.f90:121-122 of vcycle:
// residual + restriction(rb)); frestrict_centers3d
// (core/mgfor/basicoperators.f90:32-60) with Rcoef = 1/2; coarse x = 0
// (core/mgfor/operators.f90:209).
// The first smoothing sweep of a level whose x is zero (after restriction):
// k_mg_sweep with src = 0, i.e. cff1*0 + omega*(0 - b)*idiag with the six
// zero neighbours summing to +0, evaluated literally (same bits).
template <typename T>
__device__ __forceinline__ T sweep_from_zero(const Geo &g, T bv, double omega, const MGTabs &tab,
                                             int i, int j, int k) {
    const int nw = (g.px || i > 0) + (g.px || i < g.nx - 1) + (g.py || j > 0) +
                   (g.py || j < g.ny - 1) + (g.pz || k > 0) + (g.pz || k < g.nz - 1);
    const T om = T(omega), cff1 = T(1.0 - omega);
    return cff1 * T(0) + om * (T(0) - bv) * T(tab.idiag[nw]);
}

// yc non-null: the coarse level's first sweep from x = 0 is written to yc
// (then coarse x is not needed: the second sweep overwrites it); else xc = 0.
template <typename Tf, typename Tc>
__device__ __forceinline__ void restrict_cell(const Geo &gf, const Geo &gc,
                                              const Tf *__restrict__ x, const Tf *__restrict__ b,
                                              Tc *__restrict__ bc, Tc *__restrict__ xc,
                                              const MGTabs &tab, int ic, int jc, int kc,
                                              Tc *__restrict__ yc = nullptr, double omega = 0.0);

template <typename Tf, typename Tc>
__global__ __launch_bounds__(256) void k_mg_resid_restrict(Geo gf, Geo gc,
                                                           const Tf *__restrict__ x,
                                                           const Tf *__restrict__ b,
                                                           Tc *__restrict__ bc, Tc *__restrict__ xc,
                                                           MGTabs tab, Tc *__restrict__ yc,
                                                           double omega) {
    const Blk3 bk = xcd_block();
    const int ic = bk.x * MG_BX + threadIdx.x;
    const int jc = bk.y * MG_BY + threadIdx.y;
    const int kc = bk.z;
    if (ic >= gc.nx || jc >= gc.ny) return;
    restrict_cell<Tf, Tc>(gf, gc, x, b, bc, xc, tab, ic, jc, kc, yc, omega);
}

template <typename Tf, typename Tc>
__device__ __forceinline__ void restrict_cell(const Geo &gf, const Geo &gc,
                                              const Tf *__restrict__ x, const Tf *__restrict__ b,
                                              Tc *__restrict__ bc, Tc *__restrict__ xc,
                                              const MGTabs &tab, int ic, int jc, int kc,
                                              Tc *__restrict__ yc, double omega) {
    const int i = 2 * ic, j = 2 * jc, k = 2 * kc;
    const Tf r000 = residual_at<Tf>(gf, x, b, i, j, k, tab);
    const Tf r100 = residual_at<Tf>(gf, x, b, i + 1, j, k, tab);
    const Tf r010 = residual_at<Tf>(gf, x, b, i, j + 1, k, tab);
    const Tf r110 = residual_at<Tf>(gf, x, b, i + 1, j + 1, k, tab);
    const Tf r001 = residual_at<Tf>(gf, x, b, i, j, k + 1, tab);
    const Tf r101 = residual_at<Tf>(gf, x, b, i + 1, j, k + 1, tab);
    const Tf r011 = residual_at<Tf>(gf, x, b, i, j + 1, k + 1, tab);
    const Tf r111 = residual_at<Tf>(gf, x, b, i + 1, j + 1, k + 1, tab);
    const int64_t cc = (int64_t)kc * gc.sz + (int64_t)jc * gc.sy + ic;
    const Tc bv = Tc(Tf(0.5) * (r000 + r100 + r010 + r110 + r001 + r101 + r011 + r111));
    bc[cc] = bv;
    if (yc) yc[cc] = sweep_from_zero<Tc>(gc, bv, omega, tab, ic, jc, kc);
    else xc[cc] = Tc(0);
}

template <typename Tf, typename Tc>
hipError_t launch_mg_resid_restrict(const Geo &gf, const Geo &gc, const Tf *x, const Tf *b,
                                    Tc *bc, Tc *xc, const MGTabs &tab, hipStream_t s, Tc *yc,
                                    double omega) {
    NYL_LAUNCH((k_mg_resid_restrict<Tf, Tc>), mg_grid(gc), dim3(MG_BX, MG_BY, 1), 0, s,
                       gf, gc, x, b, bc, xc, tab, yc, omega);
    return hipGetLastError();
}

// fprolongation_centers3d (core/mgfor/basicoperators.f90:173-231) at one
// fine cell: Pcoef*(3*b + a) with the (9,3,3,1) planar weights, b the
// parent's coarse plane and a the coarse plane on the child's side.
template <typename Tf, typename Tc>
__device__ __forceinline__ Tf prolong_term(const Geo &gc, const Tc *__restrict__ xc,
                                           const MGTabs &tab, int i, int j, int k) {
    const int ic = i >> 1, jc = j >> 1, kc = k >> 1;
    // coarse neighbour on the side of the fine cell within its parent
    int icn = (i & 1) ? ic + 1 : ic - 1;
    int jcn = (j & 1) ? jc + 1 : jc - 1;
    int kcn = (k & 1) ? kc + 1 : kc - 1;
    int mx = 1, my = 1, mz = 1;
    if (gc.px) icn = wrapi(icn, gc.nx); else if (icn < 0 || icn >= gc.nx) mx = 0;
    if (gc.py) jcn = wrapi(jcn, gc.ny); else if (jcn < 0 || jcn >= gc.ny) my = 0;
    if (gc.pz) kcn = wrapi(kcn, gc.nz); else if (kcn < 0 || kcn >= gc.nz) mz = 0;
    auto X = [&](int a, int bb, int cc, bool ok) -> Tf {
        return ok ? Tf(xc[(int64_t)cc * gc.sz + (int64_t)bb * gc.sy + a]) : Tf(0);
    };
    auto A = [&](int kk, bool okz) -> Tf {
        return Tf(9) * X(ic, jc, kk, okz) + Tf(3) * X(icn, jc, kk, okz && mx)
             + Tf(3) * X(ic, jcn, kk, okz && my) + X(icn, jcn, kk, okz && mx && my);
    };
    const Tf bmid = A(kc, true);
    const Tf aside = A(kcn, mz != 0);
    return Tf(tab.pcoef[mx + 2 * my + 4 * mz]) * (Tf(3) * bmid + aside);
}

// one thread per fine cell: xf = xf + Pcoef*(3*b + a)
template <typename Tf, typename Tc>
__global__ __launch_bounds__(256) void k_mg_prolong(Geo gf, Geo gc, Tf *__restrict__ xf,
                                                    const Tc *__restrict__ xc, MGTabs tab) {
    const Blk3 bk = xcd_block();
    const int i = bk.x * MG_BX + threadIdx.x;
    const int j = bk.y * MG_BY + threadIdx.y;
    const int k = bk.z;
    if (i >= gf.nx || j >= gf.ny) return;
    const int64_t c = (int64_t)k * gf.sz + (int64_t)j * gf.sy + i;
    xf[c] = xf[c] + prolong_term<Tf, Tc>(gc, xc, tab, i, j, k);
}

template <typename Tf, typename Tc>
hipError_t launch_mg_prolong(const Geo &gf, const Geo &gc, Tf *xf, const Tc *xc,
                             const MGTabs &tab, hipStream_t s) {
    NYL_LAUNCH((k_mg_prolong<Tf, Tc>), mg_grid(gf), dim3(MG_BX, MG_BY, 1), 0, s, gf, gc,
                       xf, xc, tab);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Deterministic two-stage reductions.  The grid size depends only on the
// cell count, each thread sums a fixed set of cells in a fixed order and the
// block trees are fixed, so a result does not change from run to run.

#define RED_T 256

// Fixed-order tree over the RED_T threads of a block; tid is the linear
// thread index (blocks may be 2-D).
__device__ __forceinline__ void block_sum_store(double v, double *out, int tid) {
    __shared__ double sh[RED_T];
    sh[tid] = v;
    __syncthreads();
    for (int s = RED_T / 2; s > 0; s >>= 1) {
        if (tid < s) sh[tid] = sh[tid] + sh[tid + s];
        __syncthreads();
    }
    if (tid == 0) *out = sh[0];
}

__device__ __forceinline__ void block_max_store(double v, double *out, int tid) {
    __shared__ double sh[RED_T];
    sh[tid] = v;
    __syncthreads();
    for (int s = RED_T / 2; s > 0; s >>= 1) {
        if (tid < s) sh[tid] = fmax(sh[tid], sh[tid + s]);
        __syncthreads();
    }
    if (tid == 0) *out = sh[0];
}

// One fsmoother3d sweep with 2.5-D tiling: a block owns a 64 x 8 column of
// cells (two rows per thread) and marches up KZ planes.  The current plane,
// with a one-cell ring, sits in LDS (three buffers, one barrier per plane);
// the z-neighbours of a thread's own cells stay in registers.  Each global
// value is loaded about once.  The loads run one plane ahead of the
// arithmetic: plane k+2 of x and plane k+1 of b are in flight while plane k
// is computed, so a plane costs one memory latency, not two in a row.  The
// arithmetic is that of k_mg_sweep (same operand order), so the result is
// bit-identical; out-of-domain neighbours along a wall axis are zero.
//
// RESNORM: the block also sums the squared residual b + diag*src - sum6(src)
//          of its INPUT over its cells into partials[block] (the convergence
//          test of the previous V-cycle, core/mgfor/operators.f90:81-96).
#define SW_TX 64
#define SW_TY 8
#define SW_R (SW_TY / 4)
#define SW_LX (SW_TX + 2)
#define SW_LY (SW_TY + 2)

// Per-thread plane access of the 2.5-D kernels.  Thread (tx, ty) of a
// SW_TX x 4 block owns R cells of column i0 + tx (LDS rows row[r]) and, for
// the first NRING threads, one cell of the one-cell ring around the tile.
// Every load is issued from a valid address (clamped when the cell is outside
// a wall axis) and its value selected when it is stored to LDS, so the load
// of a plane can be in flight across the barrier and the previous plane's
// arithmetic.
template <int R>
struct Plane25 {
    static constexpr int NRING = 2 * SW_LX + 2 * SW_TY;
    int off[R], row[R];   // in-plane offsets and LDS rows of the own cells
    int roff, rlx, rly;   // ring cell: in-plane offset, LDS column and row
    int rgi, rgj;         // ring cell: (wrapped) grid position
    bool ring, rin;       // thread holds a ring cell; the cell is in the domain

    // row[] is set by the caller first
    __device__ __forceinline__ void init(const Geo &g, int i0, int j0, int tx, int tid) {
#pragma unroll
        for (int r = 0; r < R; r++) off[r] = (j0 + row[r] - 1) * (int)g.sy + i0 + tx;
        ring = tid < NRING;
        rin = false;
        roff = off[0];
        rlx = rly = rgi = rgj = 0;
        if (ring) {
            const int q = tid;
            if (q < SW_LX) { rlx = q; rly = 0; }
            else if (q < 2 * SW_LX) { rlx = q - SW_LX; rly = SW_LY - 1; }
            else { const int t = q - 2 * SW_LX; rly = 1 + (t >> 1); rlx = (t & 1) ? SW_LX - 1 : 0; }
            int gi = i0 + rlx - 1, gj = j0 + rly - 1;
            bool in = true;
            if (g.px) gi = wrapi(gi, g.nx); else in = in && gi >= 0 && gi < g.nx;
            if (g.py) gj = wrapi(gj, g.ny); else in = in && gj >= 0 && gj < g.ny;
            if (in) { roff = gj * (int)g.sy + gi; rgi = gi; rgj = gj; rin = true; }
        }
    }
};

// raw values of one plane as loaded (selected in put25)
template <typename T, int R>
struct Raw25 {
    T own[R], ring;
    int mw;     // plane index loaded (wrapped; 0 when outside a wall axis)
    bool zin;   // the plane is inside the domain
};

template <typename T, int R>
__device__ __forceinline__ void fetch25(const Plane25<R> &p, const Geo &g, const T *__restrict__ src,
                                        int m, Raw25<T, R> &v) {
    v.zin = g.pz || (m >= 0 && m < g.nz);
    v.mw = g.pz ? wrapi(m, g.nz) : (v.zin ? m : 0);
    const T *s = src + (int64_t)v.mw * g.sz;
#pragma unroll
    for (int r = 0; r < R; r++) v.own[r] = s[p.off[r]];
    v.ring = s[p.roff];
}

// Store a fetched plane into an LDS buffer and the own values into own[]:
// zero outside the domain, plus the prolongation of the coarse xc when
// PROLONG (the value k_mg_prolong would store, same expression).
template <typename T, int R, bool PROLONG>
__device__ __forceinline__ void put25(const Plane25<R> &p, const Raw25<T, R> &v,
                                      T (*buf)[SW_LX], int tx, T *own, const Geo &gc,
                                      const T *__restrict__ xc, const MGTabs &tab, int i,
                                      int j0) {
#pragma unroll
    for (int r = 0; r < R; r++) {
        T t = v.zin ? v.own[r] : T(0);
        if (PROLONG && v.zin) t = t + prolong_term<T, T>(gc, xc, tab, i, j0 + p.row[r] - 1, v.mw);
        own[r] = t;
        buf[p.row[r]][1 + tx] = t;
    }
    if (p.ring) {
        T t = (v.zin && p.rin) ? v.ring : T(0);
        if (PROLONG && v.zin && p.rin) t = t + prolong_term<T, T>(gc, xc, tab, p.rgi, p.rgj, v.mw);
        buf[p.rly][p.rlx] = t;
    }
}

//
// PROLONG: src is first corrected by the prolongation of the coarse xc, on
//          the own cells and the ring; the corrected x is only read by this
//          sweep (the next sweep reads dst), so it is not stored.
//
// ADDX: the result is not stored; it is added to the FP64 xo instead (the
//       last sweep of an inner V-cycle of the mixed solve, whose output e is
//       only ever used for x += e).
template <typename T, bool RESNORM, bool PROLONG, bool ADDX>
__global__ __launch_bounds__(256) void k_mg_sweep25(Geo g, const T *__restrict__ src,
                                                    T *__restrict__ dst,
                                                    const T *__restrict__ b, double omega,
                                                    MGTabs tab, int kz,
                                                    double *__restrict__ partials, Geo gc,
                                                    const T *__restrict__ xc,
                                                    double *__restrict__ xo) {
    __shared__ T sp[3][SW_LY][SW_LX];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * SW_TX + tx;
    const Blk3 bk = xcd_block();
    const int i0 = bk.x * SW_TX, j0 = bk.y * SW_TY, k0 = bk.z * kz;
    const T om = T(omega), cff1 = T(1.0 - omega);
    const int i = i0 + tx;

    Plane25<SW_R> P;
#pragma unroll
    for (int r = 0; r < SW_R; r++) P.row[r] = 1 + ty + 4 * r;
    P.init(g, i0, j0, tx, tid);
    // diagonal and its inverse per own cell for 0, 1, 2 in-domain z-neighbours
    // (the x and y parts of the neighbour count are fixed along the march)
    T idg[3][SW_R], dg[3][SW_R];
#pragma unroll
    for (int r = 0; r < SW_R; r++) {
        const int j = j0 + P.row[r] - 1;
        const int nxy = (g.px || i > 0) + (g.px || i < g.nx - 1) + (g.py || j > 0)
                      + (g.py || j < g.ny - 1);
#pragma unroll
        for (int q = 0; q < 3; q++) {
            idg[q][r] = T(tab.idiag[nxy + q]);
            dg[q][r] = RESNORM ? T(tab.diag[nxy + q]) : T(0);
        }
    }

    T zm[SW_R], z0[SW_R], zp[SW_R], bc[SW_R], bn[SW_R];
    double xoc[SW_R], xon[SW_R];
    Raw25<T, SW_R> va, vb;
    double acc = 0.0;
    // plane k lives in buffer (k - k0 + 1) % 3; storing plane k+1 reuses the
    // buffer of plane k-2, whose readers all passed the previous barrier
    fetch25(P, g, src, k0 - 1, va);
    fetch25(P, g, src, k0, vb);
    put25<T, SW_R, PROLONG>(P, va, sp[0], tx, zm, gc, xc, tab, i, j0);
    put25<T, SW_R, PROLONG>(P, vb, sp[1], tx, z0, gc, xc, tab, i, j0);
    fetch25(P, g, src, k0 + 1, va);
#pragma unroll
    for (int r = 0; r < SW_R; r++) {
        bn[r] = b[(int64_t)k0 * g.sz + P.off[r]];
        if (ADDX) xon[r] = xo[(int64_t)k0 * g.sz + P.off[r]];
    }
    for (int k = k0; k < k0 + kz; k++) {
        const int cur = (k - k0 + 1) % 3;
        put25<T, SW_R, PROLONG>(P, va, sp[(k - k0 + 2) % 3], tx, zp, gc, xc, tab, i, j0);
#pragma unroll
        for (int r = 0; r < SW_R; r++) {
            bc[r] = bn[r];
            if (ADDX) xoc[r] = xon[r];
        }
        if (k + 1 < k0 + kz) {
            fetch25(P, g, src, k + 2, va);
#pragma unroll
            for (int r = 0; r < SW_R; r++) {
                bn[r] = b[(int64_t)(k + 1) * g.sz + P.off[r]];
                if (ADDX) xon[r] = xo[(int64_t)(k + 1) * g.sz + P.off[r]];
            }
        }
        __syncthreads();
        const int nz_ = (g.pz || k > 0) + (g.pz || k < g.nz - 1);
#pragma unroll
        for (int r = 0; r < SW_R; r++) {
            const int ly = P.row[r], lx = 1 + tx;
            const T S = sp[cur][ly][lx - 1] + sp[cur][ly][lx + 1] + sp[cur][ly - 1][lx]
                      + sp[cur][ly + 1][lx] + zm[r] + zp[r];
            const int64_t c = (int64_t)k * g.sz + P.off[r];
            const T bv = bc[r];
            const T id = nz_ == 2 ? idg[2][r] : (nz_ == 1 ? idg[1][r] : idg[0][r]);
            if (RESNORM) {
                const T d = nz_ == 2 ? dg[2][r] : (nz_ == 1 ? dg[1][r] : dg[0][r]);
                const double rr = (double)(bv + d * z0[r] - S);
                acc = acc + rr * rr;
            }
            const T v = cff1 * z0[r] + om * (S - bv) * id;
            if (ADDX) xo[c] = xoc[r] + (double)v;
            else dst[c] = v;
            zm[r] = z0[r];
            z0[r] = zp[r];
        }
    }
    if (RESNORM) {
        __syncthreads();
        const int nb = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
        block_sum_store(acc, partials + nb, tid);
    }
}

// Residual + restriction with 2.5-D tiling: a block owns fine 64 x 8 cells
// (coarse 32 x 4) and marches up kz fine planes (kz even); thread (tx, ty)
// holds the fine rows j0 + 2 ty and j0 + 2 ty + 1 of column i0 + tx, so the
// y pair of a coarse cell is in one thread, the x pair in two adjacent lanes
// and the z pair in consecutive planes.  Loads run one plane ahead, as in
// k_mg_sweep25.  The residuals are those of residual_at and the eight are
// summed in the order of k_mg_resid_restrict, so the coarse b is bitwise the
// same.
template <typename Tf, typename Tc>
__global__ __launch_bounds__(256) void k_mg_restrict25(Geo g, Geo gc, const Tf *__restrict__ x,
                                                       const Tf *__restrict__ b,
                                                       Tc *__restrict__ bc, Tc *__restrict__ xc,
                                                       MGTabs tab, int kz, Tc *__restrict__ yc,
                                                       double omega) {
    __shared__ Tf sp[3][SW_LY][SW_LX];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * SW_TX + tx;
    const Blk3 bk = xcd_block();
    const int i0 = bk.x * SW_TX, j0 = bk.y * SW_TY, k0 = bk.z * kz;
    const int i = i0 + tx;

    Plane25<2> P;
    P.row[0] = 1 + 2 * ty;
    P.row[1] = 2 + 2 * ty;
    P.init(g, i0, j0, tx, tid);
    Tf dg[3][2];
#pragma unroll
    for (int r = 0; r < 2; r++) {
        const int j = j0 + 2 * ty + r;
        const int nxy = (g.px || i > 0) + (g.px || i < g.nx - 1) + (g.py || j > 0)
                      + (g.py || j < g.ny - 1);
#pragma unroll
        for (int q = 0; q < 3; q++) dg[q][r] = Tf(tab.diag[nxy + q]);
    }
    Tf zm[2], z0[2], zp[2], bcur[2], bn[2];
    Raw25<Tf, 2> va, vb;
    Tf acc = Tf(0);
    fetch25(P, g, x, k0 - 1, va);
    fetch25(P, g, x, k0, vb);
    put25<Tf, 2, false>(P, va, sp[0], tx, zm, g, x, tab, i, j0);
    put25<Tf, 2, false>(P, vb, sp[1], tx, z0, g, x, tab, i, j0);
    fetch25(P, g, x, k0 + 1, va);
#pragma unroll
    for (int r = 0; r < 2; r++) bn[r] = b[(int64_t)k0 * g.sz + P.off[r]];
    for (int k = k0; k < k0 + kz; k++) {
        const int cur = (k - k0 + 1) % 3;
        put25<Tf, 2, false>(P, va, sp[(k - k0 + 2) % 3], tx, zp, g, x, tab, i, j0);
#pragma unroll
        for (int r = 0; r < 2; r++) bcur[r] = bn[r];
        if (k + 1 < k0 + kz) {
            fetch25(P, g, x, k + 2, va);
#pragma unroll
            for (int r = 0; r < 2; r++) bn[r] = b[(int64_t)(k + 1) * g.sz + P.off[r]];
        }
        __syncthreads();
        const int nz_ = (g.pz || k > 0) + (g.pz || k < g.nz - 1);
        Tf r[2];
#pragma unroll
        for (int rr = 0; rr < 2; rr++) {
            const int ly = P.row[rr], lx = 1 + tx;
            const Tf S = sp[cur][ly][lx - 1] + sp[cur][ly][lx + 1] + sp[cur][ly - 1][lx]
                       + sp[cur][ly + 1][lx] + zm[rr] + zp[rr];
            const Tf d = nz_ == 2 ? dg[2][rr] : (nz_ == 1 ? dg[1][rr] : dg[0][rr]);
            r[rr] = bcur[rr] + d * z0[rr] - S;
            zm[rr] = z0[rr];
            z0[rr] = zp[rr];
        }
        const Tf rx0 = __shfl_down(r[0], 1), rx1 = __shfl_down(r[1], 1);
        if (((k - k0) & 1) == 0) {
            acc = ((r[0] + rx0) + r[1]) + rx1;
        } else {
            acc = (((acc + r[0]) + rx0) + r[1]) + rx1;
            if ((tx & 1) == 0) {
                const int kc = k >> 1, jc = (j0 >> 1) + ty, ic = i >> 1;
                const int64_t cc = (int64_t)kc * gc.sz + (int64_t)jc * gc.sy + ic;
                const Tc bv = Tc(Tf(0.5) * acc);
                bc[cc] = bv;
                if (yc) yc[cc] = sweep_from_zero<Tc>(gc, bv, omega, tab, ic, jc, kc);
                else xc[cc] = Tc(0);
            }
        }
    }
}

static inline int red_blocks(int64_t n) {
    int64_t b = (n + RED_T - 1) / RED_T;
    return (int)(b < NYL_RED_BLOCKS ? (b < 1 ? 1 : b) : NYL_RED_BLOCKS);
}

// pub.dst non-null: thread 0, which stored *out = pub.src[pub.n - 1], copies
// pub.src[0 .. n) to coherent host memory and then stores the sequence
// number (the other results were written by earlier kernels)
__global__ __launch_bounds__(RED_T) void k_final_sum(const double *__restrict__ partials, int np,
                                                     double *out, Publish pub) {
    double acc = 0.0;
    for (int q = threadIdx.x; q < np; q += RED_T) acc = acc + partials[q];
    block_sum_store(acc, out, threadIdx.x);
    if (pub.dst && threadIdx.x == 0) {
        for (int q = 0; q < pub.n; q++) pub.dst[q] = pub.src[q];
        __threadfence_system();
        __hip_atomic_store(pub.flag, pub.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

hipError_t launch_final_sum(const double *partials, int np, double *out, hipStream_t s) {
    Publish pub = {};
    Publish &r = g_publish;
    if (r.dst && !r.consumed && r.n >= 1 && out == r.src + (r.n - 1)) {
        pub = r;
        r.consumed = 1;
    }
    NYL_LAUNCH(k_final_sum, dim3(1), dim3(RED_T), 0, s, partials, np, out, pub);
    return hipGetLastError();
}

__global__ __launch_bounds__(RED_T) void k_final_max(const double *__restrict__ partials, int np,
                                                     double *__restrict__ out) {
    double acc = 0.0;
    for (int q = threadIdx.x; q < np; q += RED_T) acc = fmax(acc, partials[q]);
    block_max_store(acc, out, threadIdx.x);
}

template <typename T>
__global__ __launch_bounds__(RED_T) void k_norm2(int64_t n, const T *__restrict__ v,
                                                 double *__restrict__ partials) {
    double acc = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * RED_T + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * RED_T) {
        const double x = (double)v[q];
        acc = acc + x * x;
    }
    block_sum_store(acc, partials + blockIdx.x, threadIdx.x);
}

template <typename T>
hipError_t launch_norm2(const Geo &g, const T *v, double *partials, double *dev_out,
                        hipStream_t s) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int nb = red_blocks(n);
    NYL_LAUNCH((k_norm2<T>), dim3(nb), dim3(RED_T), 0, s, n, v, partials);
    (void)launch_final_sum(partials, nb, dev_out, s);
    return hipGetLastError();
}

template <typename T>
__global__ __launch_bounds__(RED_T) void k_resnorm2(Geo g, const T *__restrict__ x,
                                                    const T *__restrict__ b, MGTabs tab,
                                                    double *__restrict__ partials) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    double acc = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * RED_T + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * RED_T) {
        const int i = (int)(q % g.nx);
        const int64_t t = q / g.nx;
        const int j = (int)(t % g.ny);
        const int k = (int)(t / g.ny);
        const double r = (double)residual_at<T>(g, x, b, i, j, k, tab);
        acc = acc + r * r;
    }
    block_sum_store(acc, partials + blockIdx.x, threadIdx.x);
}

template <typename T>
hipError_t launch_resnorm2(const Geo &g, const T *x, const T *b, const MGTabs &tab,
                           double *partials, double *dev_out, hipStream_t s) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int nb = red_blocks(n);
    NYL_LAUNCH((k_resnorm2<T>), dim3(nb), dim3(RED_T), 0, s, g, x, b, tab, partials);
    (void)launch_final_sum(partials, nb, dev_out, s);
    return hipGetLastError();
}

// max of U**2 + V**2 + W**2 at equal indices (core/nyles.py:249)
template <typename T>
__global__ __launch_bounds__(RED_T) void k_maxu2(int64_t n, const T *__restrict__ u,
                                                 const T *__restrict__ v,
                                                 const T *__restrict__ w, double idx2,
                                                 double idy2, double idz2,
                                                 double *__restrict__ partials) {
    double acc = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * RED_T + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * RED_T) {
        const double U = (double)(u[q] * T(idx2)), V = (double)(v[q] * T(idy2)),
                     W = (double)(w[q] * T(idz2));
        const double s2 = U * U + V * V + W * W;
        acc = fmax(acc, s2);
    }
    block_max_store(acc, partials + blockIdx.x, threadIdx.x);
}

template <typename T>
hipError_t launch_maxu2(const Geo &g, const T *u, const T *v, const T *w, double idx2,
                        double idy2, double idz2, double *partials, double *dev_out,
                        hipStream_t s) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int nb = red_blocks(n);
    NYL_LAUNCH((k_maxu2<T>), dim3(nb), dim3(RED_T), 0, s, n, u, v, w, idx2, idy2, idz2,
                       partials);
    NYL_LAUNCH(k_final_max, dim3(1), dim3(RED_T), 0, s, partials, nb, dev_out);
    return hipGetLastError();
}

template <typename T>
__global__ __launch_bounds__(RED_T) void k_sum(int64_t n, const T *__restrict__ v,
                                               double *__restrict__ partials) {
    double acc = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * RED_T + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * RED_T)
        acc = acc + (double)v[q];
    block_sum_store(acc, partials + blockIdx.x, threadIdx.x);
}

template <typename T>
hipError_t launch_sum(const Geo &g, const T *v, double *partials, double *dev_out,
                      hipStream_t s) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int nb = red_blocks(n);
    NYL_LAUNCH((k_sum<T>), dim3(nb), dim3(RED_T), 0, s, n, v, partials);
    (void)launch_final_sum(partials, nb, dev_out, s);
    return hipGetLastError();
}

template <typename T>
__global__ __launch_bounds__(RED_T) void k_nonfinite(int64_t n, const T *__restrict__ a,
                                                     const T *__restrict__ b,
                                                     const T *__restrict__ c,
                                                     const T *__restrict__ d,
                                                     double *__restrict__ partials) {
    double acc = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * RED_T + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * RED_T) {
        acc += isfinite((double)a[q]) ? 0.0 : 1.0;
        acc += isfinite((double)b[q]) ? 0.0 : 1.0;
        acc += isfinite((double)c[q]) ? 0.0 : 1.0;
        acc += isfinite((double)d[q]) ? 0.0 : 1.0;
    }
    block_sum_store(acc, partials + blockIdx.x, threadIdx.x);
}

template <typename T>
hipError_t launch_nonfinite(const Geo &g, const T *a, const T *b, const T *c, const T *d,
                            double *partials, double *dev_out, hipStream_t s) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int nb = red_blocks(n);
    NYL_LAUNCH((k_nonfinite<T>), dim3(nb), dim3(RED_T), 0, s, n, a, b, c, d, partials);
    (void)launch_final_sum(partials, nb, dev_out, s);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Mixed precision: multigrid as defect correction.  The outer iteration keeps
// x and b in FP64; each correction e solves A e = r with one FP32 V-cycle
// from e = 0, and x += e in FP64.  A V-cycle with linear smoothers is the
// affine map x -> x + B(b - A x), so this is the same iteration as the FP64
// solve up to the precision of B.

// Outer residual r = (b - shift) + diag*x - sum6(x) in FP64 (shift: the mean
// of b removed for compatibility), stored in the inner precision as the
// right-hand side of the inner V-cycle; y receives that V-cycle's first
// pre-smoothing sweep from e = 0 (k_mg_sweep with src = 0:
// cff1*0 + omega*(0 - r)*idiag), and the blocks sum r**2 in FP64.  Each
// block takes the planes blockIdx.z, blockIdx.z + gridDim.z, ... so that the
// grid stays within NYL_RED_BLOCKS partial sums.
template <typename TI>
__global__ __launch_bounds__(256) void k_mg_outer_resid(Geo g, const double *__restrict__ x,
                                                        const double *__restrict__ b,
                                                        double shift, TI *__restrict__ r_out,
                                                        TI *__restrict__ y_out, double omega,
                                                        MGTabs tab,
                                                        double *__restrict__ partials) {
    const int i = blockIdx.x * MG_BX + threadIdx.x;
    const int j = blockIdx.y * MG_BY + threadIdx.y;
    const int tid = threadIdx.y * MG_BX + threadIdx.x;
    const TI om = TI(omega), cff1 = TI(1.0 - omega);
    double acc = 0.0;
    if (i < g.nx && j < g.ny) {
        for (int k = blockIdx.z; k < g.nz; k += gridDim.z) {
            const int64_t c = (int64_t)k * g.sz + (int64_t)j * g.sy + i;
            const Nbrs<double> n = neighbours<double>(g, x, i, j, k, c);
            const double r = (b[c] - shift) + tab.diag[n.nwet] * x[c]
                           - (n.xm + n.xp + n.ym + n.yp + n.zm + n.zp);
            acc = acc + r * r;
            const TI rf = TI(r);
            r_out[c] = rf;
            if (y_out) y_out[c] = cff1 * TI(0) + om * (TI(0) - rf) * TI(tab.idiag[n.nwet]);
        }
    }
    const int nb = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    block_sum_store(acc, partials + nb, tid);
}

static dim3 outer_grid(const Geo &g) {
    dim3 grd = mg_grid(g);
    const int64_t cols = (int64_t)grd.x * grd.y;
    int64_t gz = NYL_RED_BLOCKS / cols;
    if (gz > g.nz) gz = g.nz;
    if (gz < 1) gz = 1;
    grd.z = (unsigned)gz;
    return grd;
}

bool outer_resid_fits(const Geo &g) {
    const dim3 grd = mg_grid(g);
    return (int64_t)grd.x * grd.y <= NYL_RED_BLOCKS;
}

// x = x + e (the outer correction, in FP64)
template <typename TI>
__global__ __launch_bounds__(256) void k_add_correction(int64_t n, double *__restrict__ x,
                                                        const TI *__restrict__ e) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q < n) x[q] = x[q] + (double)e[q];
}

template <typename TI>
hipError_t launch_add_correction(const Geo &g, double *x, const TI *e, hipStream_t s) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    NYL_LAUNCH((k_add_correction<TI>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       n, x, e);
    return hipGetLastError();
}

// sum and sum of squares in one pass: dev_out[0] = sum v, dev_out[1] = sum v**2
__global__ __launch_bounds__(RED_T) void k_sum_sumsq(int64_t n, const double *__restrict__ v,
                                                     double *__restrict__ partials, int nb) {
    double a1 = 0.0, a2 = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * RED_T + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * RED_T) {
        const double x = v[q];
        a1 = a1 + x;
        a2 = a2 + x * x;
    }
    block_sum_store(a1, partials + blockIdx.x, threadIdx.x);
    __syncthreads();
    block_sum_store(a2, partials + nb + blockIdx.x, threadIdx.x);
}

hipError_t launch_sum_sumsq(const Geo &g, const double *v, double *partials, double *dev_out,
                            hipStream_t s) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    int nb = red_blocks(n);
    if (nb > NYL_RED_BLOCKS / 2) nb = NYL_RED_BLOCKS / 2;
    NYL_LAUNCH(k_sum_sumsq, dim3(nb), dim3(RED_T), 0, s, n, v, partials, nb);
    (void)launch_final_sum(partials, nb, dev_out, s);
    (void)launch_final_sum(partials + nb, nb, dev_out + 1, s);
    return hipGetLastError();
}

// n results from device memory to coherent host memory, then the sequence
// number (after a system-scope fence, so the host sees the values first)
__global__ __launch_bounds__(64) void k_publish(const double *__restrict__ src, int n,
                                                double *dst, uint32_t *flag, uint32_t seq) {
    if (threadIdx.x != 0) return;
    for (int q = 0; q < n; q++) dst[q] = src[q];
    __threadfence_system();
    __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

hipError_t launch_publish(const double *src, int n, double *dst, uint32_t *flag, uint32_t seq,
                          hipStream_t s) {
    NYL_LAUNCH(k_publish, dim3(1), dim3(64), 0, s, src, n, dst, flag, seq);
    return hipGetLastError();
}

// Sum of the kinetic energy field evaluated in FP64 from u, v, w with the
// formula of core/fortran_kinenergy.f90:47-55 (as the diagnostic kernel):
// the energy diagnostic of both precision modes.
template <typename T>
__global__ __launch_bounds__(RED_T) void k_kesum(Geo g, const T *__restrict__ u,
                                                 const T *__restrict__ v,
                                                 const T *__restrict__ w, double idx2,
                                                 double idy2, double idz2,
                                                 double *__restrict__ partials) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const double cx = 0.5 * idx2, cy = 0.5 * idy2, cz = 0.5 * idz2;
    double acc = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * RED_T + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * RED_T) {
        const int i = (int)(q % g.nx);
        const int64_t t = q / g.nx;
        const int j = (int)(t % g.ny);
        const int k = (int)(t / g.ny);
        double ke = 0.0;
        if (g.px || i > 0) {
            const double a = u[q], m = u[q - (g.px ? (int64_t)(i - wrapi(i - 1, g.nx)) : 1)];
            ke = ke + cx * 0.5 * (a * a + m * m);
        }
        if (g.py || j > 0) {
            const double a = v[q], m = v[q - (g.py ? (int64_t)(j - wrapi(j - 1, g.ny)) * g.sy : g.sy)];
            ke = ke + cy * 0.5 * (a * a + m * m);
        }
        if (g.pz || k > 0) {
            const double a = w[q], m = w[q - (g.pz ? (int64_t)(k - wrapi(k - 1, g.nz)) * g.sz : g.sz)];
            ke = ke + cz * 0.5 * (a * a + m * m);
        }
        acc = acc + ke;
    }
    block_sum_store(acc, partials + blockIdx.x, threadIdx.x);
}

template <typename T>
hipError_t launch_kesum(const Geo &g, const T *u, const T *v, const T *w, double idx2,
                        double idy2, double idz2, double *partials, double *dev_out,
                        hipStream_t s) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int nb = red_blocks(n);
    NYL_LAUNCH((k_kesum<T>), dim3(nb), dim3(RED_T), 0, s, g, u, v, w, idx2, idy2, idz2,
                       partials);
    (void)launch_final_sum(partials, nb, dev_out, s);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Ensemble batches (items as in the stencil kernels; `active` skips the
// blocks of members whose solve has stopped).
#define BT NYL_BATCH_TPB

template <typename T>
__global__ __launch_bounds__(BT) void k_b_sweep(const SweepItem<T> *__restrict__ items, int n,
                                                MGTabs tab, const int *__restrict__ active) {
    const SweepItem<T> &it = find_item(items, n, blockIdx.x);
    if (active && !active[it.h.member]) return;
    int i, j, k;
    if (!cell_of(it.h.g, ((int64_t)blockIdx.x - it.h.start) * BT + threadIdx.x, i, j, k)) return;
    const int64_t c = (int64_t)k * it.h.g.sz + (int64_t)j * it.h.g.sy + i;
    const T v = sweep_value<T>(it.h.g, it.src, it.b, it.omega, tab, i, j, k, c);
    if (it.xo) it.xo[c] = it.xo[c] + (double)v;   // the mixed solve's x += e, e not stored
    else it.dst[c] = v;
}

template <typename T>
hipError_t launch_b_sweep(const SweepItem<T> *items, int n, int64_t nblk, const MGTabs &tab,
                          const int *active, hipStream_t s) {
    NYL_LAUNCH((k_b_sweep<T>), dim3((unsigned)nblk), BT, 0, s, items, n, tab, active);
    return hipGetLastError();
}

// flattened over the coarse cells
template <typename Tf, typename Tc>
__global__ __launch_bounds__(BT) void k_b_restrict(const RestrItem<Tf, Tc> *__restrict__ items,
                                                   int n, MGTabs tab,
                                                   const int *__restrict__ active) {
    const RestrItem<Tf, Tc> &it = find_item(items, n, blockIdx.x);
    if (active && !active[it.h.member]) return;
    int i, j, k;
    if (!cell_of(it.gc, ((int64_t)blockIdx.x - it.h.start) * BT + threadIdx.x, i, j, k)) return;
    restrict_cell<Tf, Tc>(it.h.g, it.gc, it.x, it.b, it.bc, it.xc, tab, i, j, k);
}

template <typename Tf, typename Tc>
hipError_t launch_b_restrict(const RestrItem<Tf, Tc> *items, int n, int64_t nblk,
                             const MGTabs &tab, const int *active, hipStream_t s) {
    NYL_LAUNCH((k_b_restrict<Tf, Tc>), dim3((unsigned)nblk), BT, 0, s, items, n, tab,
                       active);
    return hipGetLastError();
}

template <typename Tf, typename Tc>
__global__ __launch_bounds__(BT) void k_b_prolong(const ProlItem<Tf, Tc> *__restrict__ items,
                                                  int n, MGTabs tab,
                                                  const int *__restrict__ active) {
    const ProlItem<Tf, Tc> &it = find_item(items, n, blockIdx.x);
    if (active && !active[it.h.member]) return;
    int i, j, k;
    if (!cell_of(it.h.g, ((int64_t)blockIdx.x - it.h.start) * BT + threadIdx.x, i, j, k)) return;
    const int64_t c = (int64_t)k * it.h.g.sz + (int64_t)j * it.h.g.sy + i;
    it.xf[c] = it.xf[c] + prolong_term<Tf, Tc>(it.gc, it.xc, tab, i, j, k);
}

template <typename Tf, typename Tc>
hipError_t launch_b_prolong(const ProlItem<Tf, Tc> *items, int n, int64_t nblk,
                            const MGTabs &tab, const int *active, hipStream_t s) {
    NYL_LAUNCH((k_b_prolong<Tf, Tc>), dim3((unsigned)nblk), BT, 0, s, items, n, tab,
                       active);
    return hipGetLastError();
}

// Per-member partial sums: each member owns h.nblk blocks that stride over
// its cells, so its partials (and their order) do not depend on the batch.
template <typename T, int KIND>
__global__ __launch_bounds__(BT) void k_b_reduce(const RedItem<T> *__restrict__ items, int n,
                                                 MGTabs tab, const int *__restrict__ active) {
    const RedItem<T> &it = find_item(items, n, blockIdx.x);
    if (active && !active[it.h.member]) return;
    const Geo &g = it.h.g;
    const int64_t nc = (int64_t)g.nx * g.ny * g.nz;
    const int64_t lb = (int64_t)blockIdx.x - it.h.start;
    double a1 = 0.0, a2 = 0.0;
    for (int64_t q = lb * BT + threadIdx.x; q < nc; q += (int64_t)it.h.nblk * BT) {
        if (KIND == 1) {
            const int i = (int)(q % g.nx);
            const int64_t t = q / g.nx;
            const double r = (double)residual_at<T>(g, it.v, it.b, i, (int)(t % g.ny),
                                                    (int)(t / g.ny), tab);
            a1 = a1 + r * r;
        } else {
            const double x = (double)it.v[q];
            if (KIND == 2) {
                a1 = a1 + x;
                a2 = a2 + x * x;
            } else {
                a1 = a1 + x * x;
            }
        }
    }
    block_sum_store(a1, it.partials + lb, threadIdx.x);
    if (KIND == 2) {
        __syncthreads();
        block_sum_store(a2, it.partials + it.h.nblk + lb, threadIdx.x);
    }
}

template <typename T>
hipError_t launch_b_reduce(const RedItem<T> *items, int n, int64_t nblk, int kind,
                           const MGTabs &tab, const int *active, hipStream_t s) {
    const dim3 grd((unsigned)nblk);
    if (kind == 0) NYL_LAUNCH((k_b_reduce<T, 0>), grd, BT, 0, s, items, n, tab, active);
    else if (kind == 1) NYL_LAUNCH((k_b_reduce<T, 1>), grd, BT, 0, s, items, n, tab, active);
    else NYL_LAUNCH((k_b_reduce<T, 2>), grd, BT, 0, s, items, n, tab, active);
    return hipGetLastError();
}

// the outer FP64 residual of the mixed solve (k_mg_outer_resid per member)
__global__ __launch_bounds__(BT) void k_b_outer_resid(const OuterItem *__restrict__ items, int n,
                                                      BMG mg, MGTabs tab,
                                                      const int *__restrict__ active) {
    const OuterItem &it = find_item(items, n, blockIdx.x);
    if (active && !active[it.h.member]) return;
    const Geo &g = it.h.g;
    const int64_t nc = (int64_t)g.nx * g.ny * g.nz;
    const int64_t lb = (int64_t)blockIdx.x - it.h.start;
    const double shift = mg.shift[it.h.member];
    const float om = float(it.omega), cff1 = float(1.0 - it.omega);
    double acc = 0.0;
    for (int64_t q = lb * BT + threadIdx.x; q < nc; q += (int64_t)it.h.nblk * BT) {
        const int i = (int)(q % g.nx);
        const int64_t t = q / g.nx;
        const int j = (int)(t % g.ny), k = (int)(t / g.ny);
        const Nbrs<double> nb = neighbours<double>(g, it.x, i, j, k, q);
        const double r = (it.b[q] - shift) + tab.diag[nb.nwet] * it.x[q]
                       - (nb.xm + nb.xp + nb.ym + nb.yp + nb.zm + nb.zp);
        acc = acc + r * r;
        const float rf = float(r);
        it.r[q] = rf;
        it.y[q] = cff1 * 0.0f + om * (0.0f - rf) * float(tab.idiag[nb.nwet]);
    }
    block_sum_store(acc, it.partials + lb, threadIdx.x);
}

hipError_t launch_b_outer_resid(const OuterItem *items, int n, int64_t nblk, const BMG &mg,
                                const MGTabs &tab, const int *active, hipStream_t s) {
    NYL_LAUNCH(k_b_outer_resid, dim3((unsigned)nblk), BT, 0, s, items, n, mg, tab,
                       active);
    return hipGetLastError();
}

// One block per member: its partials summed in a fixed order, then the stop
// rule of mgfor's solve loop (core/mgfor/solvers.f90:17-30).
template <int KIND>
__global__ __launch_bounds__(RED_T) void k_b_final(const FinalItem *__restrict__ items, BMG mg) {
    const FinalItem &it = items[blockIdx.x];
    const int m = it.member;
    if ((KIND == NYL_FIN_STEP) && !mg.active[m]) return;
    __shared__ double tot;
    double acc = 0.0;
    for (int q = threadIdx.x; q < it.nparts; q += RED_T) acc = acc + it.partials[q];
    block_sum_store(acc, &tot, threadIdx.x);
    double acc2 = 0.0, tot2 = 0.0;
    if (KIND == NYL_FIN_SHIFT) {
        __syncthreads();
        for (int q = threadIdx.x; q < it.nparts; q += RED_T) acc2 = acc2 + it.partials[it.nparts + q];
        __shared__ double t2;
        block_sum_store(acc2, &t2, threadIdx.x);
        __syncthreads();
        tot2 = t2;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double s1 = tot;
    if (KIND == NYL_FIN_NORMB) {
        mg.normb[m] = s1;
    } else if (KIND == NYL_FIN_SHIFT) {
        const double shift = s1 / it.ncell;
        double nb = tot2 - it.ncell * shift * shift;
        mg.shift[m] = shift;
        mg.normb[m] = nb < 0.0 ? 0.0 : nb;
    } else if (KIND == NYL_FIN_INIT) {
        const double nb = mg.normb[m];
        const double res = nb > 0.0 ? s1 / nb : 0.0;
        mg.res[m] = res;
        mg.nite[m] = 0;
        mg.active[m] = !(res < mg.tol[m]) ? 1 : 0;   // mgfor runs one V-cycle even if maxite = 0
    } else {   // NYL_FIN_STEP
        const int nite = mg.nite[m] + 1;
        mg.nite[m] = nite;
        if (nite >= mg.maxite[m]) {
            mg.active[m] = 0;
        } else {
            const double nb = mg.normb[m];
            const double res = nb > 0.0 ? s1 / nb : 0.0;
            mg.res[m] = res;
            mg.active[m] = res < mg.tol[m] ? 0 : 1;
        }
    }
}

// mg.any[0] = number of members still iterating, mg.any[1] = the most
// V-cycles any member has run in the current solve
__global__ __launch_bounds__(RED_T) void k_b_count(int n, BMG mg) {
    __shared__ double tot, most;
    double acc = 0.0, mx = 0.0;
    for (int q = threadIdx.x; q < n; q += RED_T) {
        acc = acc + (double)mg.active[q];
        mx = fmax(mx, (double)mg.nite[q]);
    }
    block_sum_store(acc, &tot, threadIdx.x);
    __syncthreads();
    block_max_store(mx, &most, threadIdx.x);
    __syncthreads();
    if (threadIdx.x == 0) {
        mg.any[0] = (int)tot;
        mg.any[1] = (int)most;
    }
}

hipError_t launch_b_final(const FinalItem *items, int n, int kind, const BMG &mg, hipStream_t s) {
    switch (kind) {
    case NYL_FIN_NORMB: NYL_LAUNCH((k_b_final<NYL_FIN_NORMB>), dim3(n), RED_T, 0, s, items, mg); break;
    case NYL_FIN_INIT: NYL_LAUNCH((k_b_final<NYL_FIN_INIT>), dim3(n), RED_T, 0, s, items, mg); break;
    case NYL_FIN_STEP: NYL_LAUNCH((k_b_final<NYL_FIN_STEP>), dim3(n), RED_T, 0, s, items, mg); break;
    default: NYL_LAUNCH((k_b_final<NYL_FIN_SHIFT>), dim3(n), RED_T, 0, s, items, mg); break;
    }
    if (kind == NYL_FIN_INIT || kind == NYL_FIN_STEP)
        NYL_LAUNCH(k_b_count, dim3(1), RED_T, 0, s, n, mg);
    return hipGetLastError();
}

__global__ __launch_bounds__(RED_T) void k_b_acc(int n, BMG mg, int mode) {
    for (int q = threadIdx.x; q < n; q += RED_T) mg.acc[q] = mode ? mg.acc[q] + mg.nite[q] : 0;
}

hipError_t launch_b_acc(int n, const BMG &mg, int mode, hipStream_t s) {
    NYL_LAUNCH(k_b_acc, dim3(1), RED_T, 0, s, n, mg, mode);
    return hipGetLastError();
}

#undef BT

// ---------------------------------------------------------------------------
// Fused level kernels for the levels below the finest, where a launch costs
// more than its data: one launch does a level's second pre-smoothing sweep
// and the residual restriction to the next level (k_mg_down), one does the
// prolongation and both post-smoothing sweeps (k_mg_up).  A block owns an
// FT^3 tile of fine cells and recomputes a halo in LDS (two cells for two
// sweeps, one for sweep + residual); every value is computed by the
// expression of the per-cell kernels (sweep_cell, residual_at,
// restrict_cell, prolong_term), so the results are bitwise those of the
// separate launches.  k_mg_up reads x and writes its result to y (other
// blocks read x in their halos), so after it the level's correction is in y.

#ifndef NYL_FUSED_TILE
#define NYL_FUSED_TILE 4
#endif
#define FT NYL_FUSED_TILE    // fine tile edge (cells)
#define FH2 (FT + 4)         // tile + two-cell halo
#define FH1 (FT + 2)         // tile + one-cell halo
#define FUSED_T 256

// periodic wrap for any offset (tiles may exceed a small level)
__device__ __forceinline__ int wrapg(int m, int n) {
    const int r = m % n;
    return r < 0 ? r + n : r;
}

// in-domain test of an unwrapped coordinate; w = the cell it stands for
__device__ __forceinline__ bool axis_in(int m, int n, int per, int &w) {
    if (per) {
        w = wrapg(m, n);
        return true;
    }
    w = m;
    return m >= 0 && m < n;
}

// sweep_cell with the source in an E^3 LDS box; (a, bb, cc) the box
// position of the cell, (wi, wj, wk) the cell (for the wall tests)
template <typename T, int E>
__device__ __forceinline__ T sweep_box(const Geo &g, const T *src, int a, int bb, int cc, int wi,
                                       int wj, int wk, T bv, T om, T cff1, const MGTabs &tab) {
    const int q = (cc * E + bb) * E + a;
    int nw = 0;
    T xm = T(0), xp = T(0), ym = T(0), yp = T(0), zm = T(0), zp = T(0);
    if (g.px || wi > 0) { xm = src[q - 1]; nw++; }
    if (g.px || wi < g.nx - 1) { xp = src[q + 1]; nw++; }
    if (g.py || wj > 0) { ym = src[q - E]; nw++; }
    if (g.py || wj < g.ny - 1) { yp = src[q + E]; nw++; }
    if (g.pz || wk > 0) { zm = src[q - E * E]; nw++; }
    if (g.pz || wk < g.nz - 1) { zp = src[q + E * E]; nw++; }
    const T S = xm + xp + ym + yp + zm + zp;
    return cff1 * src[q] + om * (S - bv) * T(tab.idiag[nw]);
}

// residual_at with x in an E^3 LDS box
template <typename T, int E>
__device__ __forceinline__ T resid_box(const Geo &g, const T *src, int a, int bb, int cc, int wi,
                                       int wj, int wk, T bv, const MGTabs &tab) {
    const int q = (cc * E + bb) * E + a;
    int nw = 0;
    T xm = T(0), xp = T(0), ym = T(0), yp = T(0), zm = T(0), zp = T(0);
    if (g.px || wi > 0) { xm = src[q - 1]; nw++; }
    if (g.px || wi < g.nx - 1) { xp = src[q + 1]; nw++; }
    if (g.py || wj > 0) { ym = src[q - E]; nw++; }
    if (g.py || wj < g.ny - 1) { yp = src[q + E]; nw++; }
    if (g.pz || wk > 0) { zm = src[q - E * E]; nw++; }
    if (g.pz || wk < g.nz - 1) { zp = src[q + E * E]; nw++; }
    return bv + T(tab.diag[nw]) * src[q] - (xm + xp + ym + yp + zm + zp);
}

// Loads of an E^3 box of a level field at unwrapped offsets from (i0, j0,
// k0), zero outside a wall axis: every load of a thread is issued before
// the first is stored, so a block waits for one memory latency, not one per
// round.
template <typename T, int E>
__device__ __forceinline__ void load_box(const Geo &g, const T *__restrict__ src, T *box, int i0,
                                         int j0, int k0, int tid) {
    constexpr int N = E * E * E, R = (N + FUSED_T - 1) / FUSED_T;
    T v[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int q = tid + r * FUSED_T;
        const int a = q % E, t = q / E, bb = t % E, cc = t / E;
        int wi, wj, wk;
        const bool ii = axis_in(i0 + a, g.nx, g.px, wi);
        const bool ij = axis_in(j0 + bb, g.ny, g.py, wj);
        const bool ik = axis_in(k0 + cc, g.nz, g.pz, wk);
        v[r] = (q < N && ii && ij && ik) ? src[(int64_t)wk * g.sz + (int64_t)wj * g.sy + wi] : T(0);
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int q = tid + r * FUSED_T;
        if (q < N) box[q] = v[r];
    }
}

// x = sweep(y) on the level; b_c = restriction of b - A x; y_c = the
// coarse level's first sweep from zero
template <typename Tf, typename Tc>
__global__ __launch_bounds__(FUSED_T) void k_mg_down(Geo gf, Geo gc, const Tf *__restrict__ y,
                                                     const Tf *__restrict__ b,
                                                     Tf *__restrict__ x, Tc *__restrict__ bc,
                                                     Tc *__restrict__ yc, double omega,
                                                     MGTabs tab) {
    __shared__ Tf s2[FH2 * FH2 * FH2];
    __shared__ Tf s1[FH1 * FH1 * FH1];
    __shared__ Tf sb[FH1 * FH1 * FH1];
    const int tid = threadIdx.x;
    const int i0 = blockIdx.x * FT, j0 = blockIdx.y * FT, k0 = blockIdx.z * FT;
    const Tf om = Tf(omega), cff1 = Tf(1.0 - omega);
    load_box<Tf, FH2>(gf, y, s2, i0 - 2, j0 - 2, k0 - 2, tid);
    load_box<Tf, FH1>(gf, b, sb, i0 - 1, j0 - 1, k0 - 1, tid);
    __syncthreads();
    for (int q = tid; q < FH1 * FH1 * FH1; q += FUSED_T) {
        const int a = q % FH1, t = q / FH1, bb = t % FH1, cc = t / FH1;
        const int gi = i0 - 1 + a, gj = j0 - 1 + bb, gk = k0 - 1 + cc;
        int wi, wj, wk;
        const bool ii = axis_in(gi, gf.nx, gf.px, wi);
        const bool ij = axis_in(gj, gf.ny, gf.py, wj);
        const bool ik = axis_in(gk, gf.nz, gf.pz, wk);
        Tf v = Tf(0);
        if (ii && ij && ik) {
            v = sweep_box<Tf, FH2>(gf, s2, a + 1, bb + 1, cc + 1, wi, wj, wk, sb[q], om, cff1, tab);
            const bool own = a >= 1 && a <= FT && bb >= 1 && bb <= FT && cc >= 1 && cc <= FT &&
                             gi < gf.nx && gj < gf.ny && gk < gf.nz;
            if (own) x[(int64_t)wk * gf.sz + (int64_t)wj * gf.sy + wi] = v;
        }
        s1[q] = v;
    }
    __syncthreads();
    if (tid < (FT / 2) * (FT / 2) * (FT / 2)) {
        const int ca = tid % (FT / 2), cb = (tid / (FT / 2)) % (FT / 2), cz = tid / ((FT / 2) * (FT / 2));
        const int ic = i0 / 2 + ca, jc = j0 / 2 + cb, kc = k0 / 2 + cz;
        if (ic < gc.nx && jc < gc.ny && kc < gc.nz) {
            auto R = [&](int dx, int dy, int dz) -> Tf {
                const int a = 1 + 2 * ca + dx, bb = 1 + 2 * cb + dy, cc = 1 + 2 * cz + dz;
                return resid_box<Tf, FH1>(gf, s1, a, bb, cc, 2 * ic + dx, 2 * jc + dy, 2 * kc + dz,
                                          sb[(cc * FH1 + bb) * FH1 + a], tab);
            };
            const Tf r000 = R(0, 0, 0), r100 = R(1, 0, 0), r010 = R(0, 1, 0), r110 = R(1, 1, 0);
            const Tf r001 = R(0, 0, 1), r101 = R(1, 0, 1), r011 = R(0, 1, 1), r111 = R(1, 1, 1);
            const int64_t ccell = (int64_t)kc * gc.sz + (int64_t)jc * gc.sy + ic;
            const Tc bv = Tc(Tf(0.5) * (r000 + r100 + r010 + r110 + r001 + r101 + r011 + r111));
            bc[ccell] = bv;
            yc[ccell] = sweep_from_zero<Tc>(gc, bv, omega, tab, ic, jc, kc);
        }
    }
}

#define FC (FT / 2 + 4)      // coarse box edge of k_mg_up

// prolong_term with the coarse values in an FC^3 LDS box whose origin is
// the coarse cell (ci0, cj0, ck0); (gi, gj, gk) an unwrapped fine cell (on a
// periodic axis the same coarse values, wrapped, as for the cell it stands
// for; on a wall axis the cell itself)
template <typename Tf, typename Tc>
__device__ __forceinline__ Tf prolong_box(const Geo &gc, const Tc *cbx, int ci0, int cj0, int ck0,
                                          int gi, int gj, int gk, const MGTabs &tab) {
    const int ic = gi >> 1, jc = gj >> 1, kc = gk >> 1;
    const int icn = (gi & 1) ? ic + 1 : ic - 1;
    const int jcn = (gj & 1) ? jc + 1 : jc - 1;
    const int kcn = (gk & 1) ? kc + 1 : kc - 1;
    const int mx = (gc.px || (icn >= 0 && icn < gc.nx)) ? 1 : 0;
    const int my = (gc.py || (jcn >= 0 && jcn < gc.ny)) ? 1 : 0;
    const int mz = (gc.pz || (kcn >= 0 && kcn < gc.nz)) ? 1 : 0;
    auto X = [&](int a, int bb, int cc, bool ok) -> Tf {
        return ok ? Tf(cbx[((cc - ck0) * FC + (bb - cj0)) * FC + (a - ci0)]) : Tf(0);
    };
    auto A = [&](int kk, bool okz) -> Tf {
        return Tf(9) * X(ic, jc, kk, okz) + Tf(3) * X(icn, jc, kk, okz && mx)
             + Tf(3) * X(ic, jcn, kk, okz && my) + X(icn, jcn, kk, okz && mx && my);
    };
    const Tf bmid = A(kc, true);
    const Tf aside = A(kcn, mz != 0);
    return Tf(tab.pcoef[mx + 2 * my + 4 * mz]) * (Tf(3) * bmid + aside);
}

// y = sweep(sweep(x + P xc)) on the level (x and xc are not written)
template <typename Tf, typename Tc>
__global__ __launch_bounds__(FUSED_T) void k_mg_up(Geo gf, Geo gc, const Tf *__restrict__ x,
                                                   const Tc *__restrict__ xc,
                                                   const Tf *__restrict__ b, Tf *__restrict__ y,
                                                   double omega, MGTabs tab) {
    __shared__ Tf s2[FH2 * FH2 * FH2];
    __shared__ Tf s1[FH1 * FH1 * FH1];
    __shared__ Tf sb[FH1 * FH1 * FH1];
    __shared__ Tc scx[FC * FC * FC];
    const int tid = threadIdx.x;
    const int i0 = blockIdx.x * FT, j0 = blockIdx.y * FT, k0 = blockIdx.z * FT;
    const int ci0 = i0 / 2 - 2, cj0 = j0 / 2 - 2, ck0 = k0 / 2 - 2;
    const Tf om = Tf(omega), cff1 = Tf(1.0 - omega);
    load_box<Tf, FH2>(gf, x, s2, i0 - 2, j0 - 2, k0 - 2, tid);
    load_box<Tf, FH1>(gf, b, sb, i0 - 1, j0 - 1, k0 - 1, tid);
    load_box<Tc, FC>(gc, xc, scx, ci0, cj0, ck0, tid);
    __syncthreads();
    for (int q = tid; q < FH2 * FH2 * FH2; q += FUSED_T) {
        const int a = q % FH2, t = q / FH2, bb = t % FH2, cc = t / FH2;
        const int gi = i0 - 2 + a, gj = j0 - 2 + bb, gk = k0 - 2 + cc;
        int wi, wj, wk;
        const bool ii = axis_in(gi, gf.nx, gf.px, wi);
        const bool ij = axis_in(gj, gf.ny, gf.py, wj);
        const bool ik = axis_in(gk, gf.nz, gf.pz, wk);
        if (ii && ij && ik) s2[q] = s2[q] + prolong_box<Tf, Tc>(gc, scx, ci0, cj0, ck0, gi, gj, gk, tab);
    }
    __syncthreads();
    for (int q = tid; q < FH1 * FH1 * FH1; q += FUSED_T) {
        const int a = q % FH1, t = q / FH1, bb = t % FH1, cc = t / FH1;
        int wi, wj, wk;
        const bool ii = axis_in(i0 - 1 + a, gf.nx, gf.px, wi);
        const bool ij = axis_in(j0 - 1 + bb, gf.ny, gf.py, wj);
        const bool ik = axis_in(k0 - 1 + cc, gf.nz, gf.pz, wk);
        Tf v = Tf(0);
        if (ii && ij && ik)
            v = sweep_box<Tf, FH2>(gf, s2, a + 1, bb + 1, cc + 1, wi, wj, wk, sb[q], om, cff1, tab);
        s1[q] = v;
    }
    __syncthreads();
    for (int q = tid; q < FT * FT * FT; q += FUSED_T) {
        const int a = q % FT, t = q / FT, bb = t % FT, cc = t / FT;
        const int i = i0 + a, j = j0 + bb, k = k0 + cc;
        if (i < gf.nx && j < gf.ny && k < gf.nz) {
            const int64_t c = (int64_t)k * gf.sz + (int64_t)j * gf.sy + i;
            const Tf bv = sb[((cc + 1) * FH1 + bb + 1) * FH1 + a + 1];
            y[c] = sweep_box<Tf, FH1>(gf, s1, a + 1, bb + 1, cc + 1, i, j, k, bv, om, cff1, tab);
        }
    }
}

static inline dim3 fused_grid(const Geo &gf) {
    return dim3((gf.nx + FT - 1) / FT, (gf.ny + FT - 1) / FT, (gf.nz + FT - 1) / FT);
}

bool fused_fits(const Geo &gf, const Geo &gc) {
    return gf.nx == 2 * gc.nx && gf.ny == 2 * gc.ny && gf.nz == 2 * gc.nz && gc.nx >= 1 &&
           gc.ny >= 1 && gc.nz >= 1;
}

template <typename Tf, typename Tc>
hipError_t launch_mg_down(const Geo &gf, const Geo &gc, const Tf *y, const Tf *b, Tf *x, Tc *bc,
                          Tc *yc, double omega, const MGTabs &tab, hipStream_t s) {
    NYL_LAUNCH((k_mg_down<Tf, Tc>), fused_grid(gf), dim3(FUSED_T), 0, s, gf, gc, y, b, x,
                       bc, yc, omega, tab);
    return hipGetLastError();
}

template <typename Tf, typename Tc>
hipError_t launch_mg_up(const Geo &gf, const Geo &gc, const Tf *x, const Tc *xc, const Tf *b,
                        Tf *y, double omega, const MGTabs &tab, hipStream_t s) {
    NYL_LAUNCH((k_mg_up<Tf, Tc>), fused_grid(gf), dim3(FUSED_T), 0, s, gf, gc, x, xc, b,
                       y, omega, tab);
    return hipGetLastError();
}

#define NYL_FUSED_INST(Tf, Tc)                                                                  \
    template hipError_t launch_mg_down<Tf, Tc>(const Geo &, const Geo &, const Tf *,            \
                                               const Tf *, Tf *, Tc *, Tc *, double,             \
                                               const MGTabs &, hipStream_t);                     \
    template hipError_t launch_mg_up<Tf, Tc>(const Geo &, const Geo &, const Tf *, const Tc *,  \
                                             const Tf *, Tf *, double, const MGTabs &,          \
                                             hipStream_t);
NYL_FUSED_INST(double, double)
NYL_FUSED_INST(float, float)
NYL_FUSED_INST(float, double)
#undef NYL_FUSED_INST

// ---------------------------------------------------------------------------
// The tail's cell forms on a dense level in LDS: every neighbour index is a
// valid cell (periodic axes wrapped, a wall neighbour clamped to the cell
// itself) and the wall's zero is selected after the load, so a cell issues
// all its loads at once instead of one branch per neighbour.  The values and
// the order of the sums are those of neighbours(), sweep_value(),
// residual_at() and prolong_term().
template <typename T>
struct Six {
    T v[6];   // xm, xp, ym, yp, zm, zp
    int nw;
};
template <typename T>
__device__ __forceinline__ Six<T> six(const Geo &g, const T *a, int i, int j, int k, int c) {
    const int sy = (int)g.sy, sz = (int)g.sz;
    const bool bxm = g.px || i > 0, bxp = g.px || i < g.nx - 1;
    const bool bym = g.py || j > 0, byp = g.py || j < g.ny - 1;
    const bool bzm = g.pz || k > 0, bzp = g.pz || k < g.nz - 1;
    const int im = g.px ? wrapi(i - 1, g.nx) : (bxm ? i - 1 : i);
    const int ip = g.px ? wrapi(i + 1, g.nx) : (bxp ? i + 1 : i);
    const int jm = g.py ? wrapi(j - 1, g.ny) : (bym ? j - 1 : j);
    const int jp = g.py ? wrapi(j + 1, g.ny) : (byp ? j + 1 : j);
    const int km = g.pz ? wrapi(k - 1, g.nz) : (bzm ? k - 1 : k);
    const int kp = g.pz ? wrapi(k + 1, g.nz) : (bzp ? k + 1 : k);
    const T xm = a[c + im - i], xp = a[c + ip - i];
    const T ym = a[c + (jm - j) * sy], yp = a[c + (jp - j) * sy];
    const T zm = a[c + (km - k) * sz], zp = a[c + (kp - k) * sz];
    Six<T> r;
    r.v[0] = bxm ? xm : T(0);
    r.v[1] = bxp ? xp : T(0);
    r.v[2] = bym ? ym : T(0);
    r.v[3] = byp ? yp : T(0);
    r.v[4] = bzm ? zm : T(0);
    r.v[5] = bzp ? zp : T(0);
    r.nw = (int)bxm + (int)bxp + (int)bym + (int)byp + (int)bzm + (int)bzp;
    return r;
}

template <typename T>
__device__ __forceinline__ T tail_sweep(const Geo &g, const T *src, const T *b, double omega,
                                        const MGTabs &tab, int i, int j, int k, int c) {
    const Six<T> n = six<T>(g, src, i, j, k, c);
    const T S = n.v[0] + n.v[1] + n.v[2] + n.v[3] + n.v[4] + n.v[5];
    const T om = T(omega), cff1 = T(1.0 - omega);
    return cff1 * src[c] + om * (S - b[c]) * T(tab.idiag[n.nw]);
}

template <typename T>
__device__ __forceinline__ T tail_resid(const Geo &g, const T *x, const T *b, const MGTabs &tab,
                                        int i, int j, int k) {
    const int c = k * (int)g.sz + j * (int)g.sy + i;
    const Six<T> n = six<T>(g, x, i, j, k, c);
    return b[c] + T(tab.diag[n.nw]) * x[c] - (n.v[0] + n.v[1] + n.v[2] + n.v[3] + n.v[4] + n.v[5]);
}

template <typename Tf, typename Tc>
__device__ __forceinline__ void tail_restrict(const Geo &gf, const Geo &gc, const Tf *x,
                                              const Tf *b, Tc *bc, Tc *yc, double omega,
                                              const MGTabs &tab, int ic, int jc, int kc) {
    const int i = 2 * ic, j = 2 * jc, k = 2 * kc;
    const Tf r000 = tail_resid<Tf>(gf, x, b, tab, i, j, k);
    const Tf r100 = tail_resid<Tf>(gf, x, b, tab, i + 1, j, k);
    const Tf r010 = tail_resid<Tf>(gf, x, b, tab, i, j + 1, k);
    const Tf r110 = tail_resid<Tf>(gf, x, b, tab, i + 1, j + 1, k);
    const Tf r001 = tail_resid<Tf>(gf, x, b, tab, i, j, k + 1);
    const Tf r101 = tail_resid<Tf>(gf, x, b, tab, i + 1, j, k + 1);
    const Tf r011 = tail_resid<Tf>(gf, x, b, tab, i, j + 1, k + 1);
    const Tf r111 = tail_resid<Tf>(gf, x, b, tab, i + 1, j + 1, k + 1);
    const int cc = kc * (int)gc.sz + jc * (int)gc.sy + ic;
    const Tc bv = Tc(Tf(0.5) * (r000 + r100 + r010 + r110 + r001 + r101 + r011 + r111));
    bc[cc] = bv;
    yc[cc] = sweep_from_zero<Tc>(gc, bv, omega, tab, ic, jc, kc);
}

template <typename Tf, typename Tc>
__device__ __forceinline__ Tf tail_prolong(const Geo &gc, const Tc *xc, const MGTabs &tab, int i,
                                           int j, int k) {
    const int ic = i >> 1, jc = j >> 1, kc = k >> 1;
    int icn = (i & 1) ? ic + 1 : ic - 1;
    int jcn = (j & 1) ? jc + 1 : jc - 1;
    int kcn = (k & 1) ? kc + 1 : kc - 1;
    int mx = 1, my = 1, mz = 1;
    if (gc.px) icn = wrapi(icn, gc.nx); else if (icn < 0 || icn >= gc.nx) mx = 0;
    if (gc.py) jcn = wrapi(jcn, gc.ny); else if (jcn < 0 || jcn >= gc.ny) my = 0;
    if (gc.pz) kcn = wrapi(kcn, gc.nz); else if (kcn < 0 || kcn >= gc.nz) mz = 0;
    const int il = mx ? icn : ic, jl = my ? jcn : jc, kl = mz ? kcn : kc;
    const int sy = (int)gc.sy, sz = (int)gc.sz;
    const Tf a0 = Tf(xc[kc * sz + jc * sy + ic]), a1 = Tf(xc[kc * sz + jc * sy + il]);
    const Tf a2 = Tf(xc[kc * sz + jl * sy + ic]), a3 = Tf(xc[kc * sz + jl * sy + il]);
    const Tf c0 = Tf(xc[kl * sz + jc * sy + ic]), c1 = Tf(xc[kl * sz + jc * sy + il]);
    const Tf c2 = Tf(xc[kl * sz + jl * sy + ic]), c3 = Tf(xc[kl * sz + jl * sy + il]);
    const bool bx = mx != 0, by = my != 0, bz = mz != 0;
    const Tf bmid = Tf(9) * a0 + Tf(3) * (bx ? a1 : Tf(0)) + Tf(3) * (by ? a2 : Tf(0)) +
                    ((bx && by) ? a3 : Tf(0));
    const Tf aside = Tf(9) * (bz ? c0 : Tf(0)) + Tf(3) * ((bz && bx) ? c1 : Tf(0)) +
                     Tf(3) * ((bz && by) ? c2 : Tf(0)) + ((bz && bx && by) ? c3 : Tf(0));
    return Tf(tab.pcoef[mx + 2 * my + 4 * mz]) * (Tf(3) * bmid + aside);
}

// The smallest levels of a V-cycle in one workgroup (k_mg_tail): the down
// legs of the fused levels l0 .. L-2 (second pre-sweep y -> x, residual
// restriction into level l+1's b and its first sweep from zero into y), the
// coarsest level's second sweep (y -> x), and the up legs (x + P(xc) -> y,
// then the two post-sweeps y -> x -> y, so the level's correction ends in y
// as after k_mg_up).  All the levels live in LDS: level l0's b and y are
// read once, its correction y written once, and nothing else of these
// levels touches memory.  Each phase runs the per-cell kernels' expressions (sweep_cell,
// restrict_cell, prolong_term, in the branch-free forms above) over the
// level and a barrier separates the phases, so the result is bitwise that
// of the separate launches: one launch instead of 2 per level + 1, which
// dominate at these sizes.  One workgroup issues 64 lanes a cycle, so the
// tail pays only for levels of at most a few hundred cells (a 16^3 level
// costs more here than in its two launches across the chip).  T: the fused
// levels' type, TC: the coarsest level's.
template <typename T, typename TC>
__global__ __launch_bounds__(TAIL_T) void k_mg_tail(TailArgs a) {
    __shared__ double lds_d[NYL_TAIL_LDS / sizeof(double)];
    __shared__ MGTabs tab;   // the coefficient tables, indexed per lane
    char *lds = reinterpret_cast<char *>(lds_d);
    const int tid = threadIdx.x;
    if (tid < 7) {
        tab.idiag[tid] = a.tab.idiag[tid];
        tab.diag[tid] = a.tab.diag[tid];
    }
    if (tid < 8) tab.pcoef[tid] = a.tab.pcoef[tid];
    auto S = [&](uint32_t off) { return reinterpret_cast<T *>(lds + off); };
    auto SC = [&](uint32_t off) { return reinterpret_cast<TC *>(lds + off); };
    auto each = [&](const Geo &g, auto fn) {
        const int nx = g.nx, nxy = g.nx * g.ny, n = nxy * g.nz;
#pragma nounroll
        for (int q = tid; q < n; q += TAIL_T) {
            const int k = q / nxy, r = q - k * nxy, j = r / nx, i = r - j * nx;
            fn(i, j, k, k * (int)g.sz + j * (int)g.sy + i);
        }
    };
    const int L = a.nlev;
    {
        const TailLevel &t = a.lev[a.l0];
        const T *gy = reinterpret_cast<const T *>(a.ws + t.y);
        const T *gb = reinterpret_cast<const T *>(a.ws + t.b);
        each(t.g, [&](int, int, int, int c) {
            const T yv = gy[c], bv = gb[c];
            S(t.ly)[c] = yv;
            S(t.lb)[c] = bv;
        });
        __syncthreads();
    }
    for (int l = a.l0; l <= L - 2; l++) {
        const TailLevel &f = a.lev[l], &q = a.lev[l + 1];
        T *x = S(f.lx);
        const T *y = S(f.ly), *b = S(f.lb);
        each(f.g, [&](int i, int j, int k, int c) {
            x[c] = tail_sweep<T>(f.g, y, b, a.omega, tab, i, j, k, c);
        });
        __syncthreads();
        if (l + 1 == L - 1) {
            each(q.g, [&](int i, int j, int k, int) {
                tail_restrict<T, TC>(f.g, q.g, x, b, SC(q.lb), SC(q.ly), a.omega, tab, i, j, k);
            });
        } else {
            each(q.g, [&](int i, int j, int k, int) {
                tail_restrict<T, T>(f.g, q.g, x, b, S(q.lb), S(q.ly), a.omega, tab, i, j, k);
            });
        }
        __syncthreads();
    }
    {
        const TailLevel &c = a.lev[L - 1];
        TC *cx = SC(c.lx);
        each(c.g, [&](int i, int j, int k, int q) {
            cx[q] = tail_sweep<TC>(c.g, SC(c.ly), SC(c.lb), a.omega, tab, i, j, k, q);
        });
        __syncthreads();
    }
    for (int l = L - 2; l >= a.l0; l--) {
        const TailLevel &f = a.lev[l], &q = a.lev[l + 1];
        T *x = S(f.lx), *y = S(f.ly);
        const T *b = S(f.lb);
        if (l + 1 == L - 1) {
            const TC *xc = SC(q.lx);
            each(f.g, [&](int i, int j, int k, int c) {
                y[c] = x[c] + tail_prolong<T, TC>(q.g, xc, tab, i, j, k);
            });
        } else {
            const T *xc = S(q.ly);
            each(f.g, [&](int i, int j, int k, int c) {
                y[c] = x[c] + tail_prolong<T, T>(q.g, xc, tab, i, j, k);
            });
        }
        __syncthreads();
        each(f.g, [&](int i, int j, int k, int c) {
            x[c] = tail_sweep<T>(f.g, y, b, a.omega, tab, i, j, k, c);
        });
        __syncthreads();
        if (l > a.l0) {
            each(f.g, [&](int i, int j, int k, int c) {
                y[c] = tail_sweep<T>(f.g, x, b, a.omega, tab, i, j, k, c);
            });
            __syncthreads();
        }
    }
    // level l0's last sweep goes straight to memory
    const TailLevel &t = a.lev[a.l0];
    T *gy = reinterpret_cast<T *>(a.ws + t.y);
    each(t.g, [&](int i, int j, int k, int c) {
        gy[c] = tail_sweep<T>(t.g, S(t.lx), S(t.lb), a.omega, tab, i, j, k, c);
    });
}

template <typename T, typename TC>
hipError_t launch_mg_tail(const TailArgs &a, hipStream_t s) {
    if (a.nlev > NYL_TAIL_MAXLEV || a.l0 < 0 || a.l0 > a.nlev - 2) return hipErrorInvalidValue;
    TailArgs t = a;
    if (tail_layout(t, sizeof(T), sizeof(TC)) > NYL_TAIL_LDS) return hipErrorInvalidValue;
    NYL_LAUNCH((k_mg_tail<T, TC>), dim3(1), dim3(TAIL_T), 0, s, t);
    return hipGetLastError();
}
template hipError_t launch_mg_tail<double, double>(const TailArgs &, hipStream_t);
template hipError_t launch_mg_tail<float, double>(const TailArgs &, hipStream_t);
template hipError_t launch_mg_tail<float, float>(const TailArgs &, hipStream_t);

bool sweep25_fits(const Geo &g) {
    return g.nx % SW_TX == 0 && g.ny % SW_TY == 0 && g.nx >= SW_TX;
}

// planes per block: up to 8, fewer when that leaves the chip short of blocks
static int sweep25_kz(const Geo &g, int kmin) {
    const int64_t cols = (int64_t)(g.nx / SW_TX) * (g.ny / SW_TY);
    int kz = 8;
    while (kz > kmin && (g.nz % kz != 0 || cols * (g.nz / kz) < 1024)) kz >>= 1;
    return kz;
}

// planes per block of the restriction and the prolongation (even): up to 16
// in FP32 (256^3: 28.4 and 26.8 us against 29.4 and 28.0 at 8 planes), 8 in
// FP64, keeping at least 1024 blocks
static int rp25_kz(const Geo &g, int es) {
    const int64_t cols = (int64_t)(g.nx / SW_TX) * (g.ny / SW_TY);
    int kz = es == 4 ? 16 : 8;
    while (kz > 2 && (g.nz % kz != 0 || cols * (g.nz / kz) < 1024)) kz >>= 1;
    return kz;
}

int sweep25_blocks(const Geo &g) {
    return (g.nx / SW_TX) * (g.ny / SW_TY) * (g.nz / sweep25_kz(g, 1));
}

template <typename T>
hipError_t launch_mg_sweep25(const Geo &g, const T *src, T *dst, const T *b, double omega,
                             const MGTabs &tab, double *partials, double *dev_out,
                             hipStream_t s, const Geo *gc, const T *xc, double *addx) {
    const int kz = sweep25_kz(g, 1);
    dim3 grd(g.nx / SW_TX, g.ny / SW_TY, g.nz / kz), blk(SW_TX, 4);
    const Geo gcv = gc ? *gc : g;
    if (partials) {
        NYL_LAUNCH((k_mg_sweep25<T, true, false, false>), grd, blk, 0, s, g, src, dst, b,
                           omega, tab, kz, partials, gcv, xc, addx);
        const int nb = grd.x * grd.y * grd.z;
        (void)launch_final_sum(partials, nb, dev_out, s);
    } else if (addx) {
        NYL_LAUNCH((k_mg_sweep25<T, false, false, true>), grd, blk, 0, s, g, src, dst, b,
                           omega, tab, kz, partials, gcv, xc, addx);
    } else if (xc) {
        NYL_LAUNCH((k_mg_sweep25<T, false, true, false>), grd, blk, 0, s, g, src, dst, b,
                           omega, tab, kz, partials, gcv, xc, addx);
    } else {
        NYL_LAUNCH((k_mg_sweep25<T, false, false, false>), grd, blk, 0, s, g, src, dst,
                           b, omega, tab, kz, partials, gcv, xc, addx);
    }
    return hipGetLastError();
}

// Outer residual of the mixed solve with the 2.5-D tiling of k_mg_sweep25
// (FP64 planes of x in LDS, loads one plane ahead): per cell the values of
// k_mg_outer_resid (same operand order), the block sums of r**2 into
// partials[block] (and of b**2 into partials_b[block] when given).
template <typename TI>
__global__ __launch_bounds__(256) void k_mg_outer25(Geo g, const double *__restrict__ x,
                                                    const double *__restrict__ b, double shift,
                                                    TI *__restrict__ r_out,
                                                    TI *__restrict__ y_out, double omega,
                                                    MGTabs tab, int kz,
                                                    double *__restrict__ partials,
                                                    double *__restrict__ partials_b) {
    __shared__ double sp[3][SW_LY][SW_LX];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * SW_TX + tx;
    const Blk3 bk = xcd_block();
    const int i0 = bk.x * SW_TX, j0 = bk.y * SW_TY, k0 = bk.z * kz;
    const int i = i0 + tx;
    const TI om = TI(omega), cff1 = TI(1.0 - omega);

    Plane25<SW_R> P;
#pragma unroll
    for (int r = 0; r < SW_R; r++) P.row[r] = 1 + ty + 4 * r;
    P.init(g, i0, j0, tx, tid);
    double dg[3][SW_R];
    TI idg[3][SW_R];
#pragma unroll
    for (int r = 0; r < SW_R; r++) {
        const int j = j0 + P.row[r] - 1;
        const int nxy = (g.px || i > 0) + (g.px || i < g.nx - 1) + (g.py || j > 0)
                      + (g.py || j < g.ny - 1);
#pragma unroll
        for (int q = 0; q < 3; q++) {
            dg[q][r] = tab.diag[nxy + q];
            idg[q][r] = TI(tab.idiag[nxy + q]);
        }
    }
    double zm[SW_R], z0[SW_R], zp[SW_R], bcur[SW_R], bn[SW_R];
    Raw25<double, SW_R> va, vb;
    double acc = 0.0, accb = 0.0;
    fetch25(P, g, x, k0 - 1, va);
    fetch25(P, g, x, k0, vb);
    put25<double, SW_R, false>(P, va, sp[0], tx, zm, g, x, tab, i, j0);
    put25<double, SW_R, false>(P, vb, sp[1], tx, z0, g, x, tab, i, j0);
    fetch25(P, g, x, k0 + 1, va);
#pragma unroll
    for (int r = 0; r < SW_R; r++) bn[r] = b[(int64_t)k0 * g.sz + P.off[r]];
    for (int k = k0; k < k0 + kz; k++) {
        const int cur = (k - k0 + 1) % 3;
        put25<double, SW_R, false>(P, va, sp[(k - k0 + 2) % 3], tx, zp, g, x, tab, i, j0);
#pragma unroll
        for (int r = 0; r < SW_R; r++) bcur[r] = bn[r];
        if (k + 1 < k0 + kz) {
            fetch25(P, g, x, k + 2, va);
#pragma unroll
            for (int r = 0; r < SW_R; r++) bn[r] = b[(int64_t)(k + 1) * g.sz + P.off[r]];
        }
        __syncthreads();
        const int nz_ = (g.pz || k > 0) + (g.pz || k < g.nz - 1);
#pragma unroll
        for (int r = 0; r < SW_R; r++) {
            const int ly = P.row[r], lx = 1 + tx;
            const double S = sp[cur][ly][lx - 1] + sp[cur][ly][lx + 1] + sp[cur][ly - 1][lx]
                           + sp[cur][ly + 1][lx] + zm[r] + zp[r];
            const int64_t c = (int64_t)k * g.sz + P.off[r];
            const double d = nz_ == 2 ? dg[2][r] : (nz_ == 1 ? dg[1][r] : dg[0][r]);
            const TI id = nz_ == 2 ? idg[2][r] : (nz_ == 1 ? idg[1][r] : idg[0][r]);
            const double bv = bcur[r];
            const double rr = (bv - shift) + d * z0[r] - S;
            acc = acc + rr * rr;
            accb = accb + bv * bv;
            const TI rf = TI(rr);
            r_out[c] = rf;
            if (y_out) y_out[c] = cff1 * TI(0) + om * (TI(0) - rf) * id;
            zm[r] = z0[r];
            z0[r] = zp[r];
        }
    }
    __syncthreads();
    const int nb = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    block_sum_store(acc, partials + nb, tid);
    if (partials_b) {
        __syncthreads();
        block_sum_store(accb, partials_b + nb, tid);
    }
}

// planes per block of the tiled outer residual: as many as divide nz up to 8
static int outer25_kz(const Geo &g) {
    int kz = 8;
    while (kz > 1 && g.nz % kz != 0) kz >>= 1;
    return kz;
}

bool outer25_fits(const Geo &g) {
    if (!sweep25_fits(g)) return false;
    const int64_t nb = (int64_t)(g.nx / SW_TX) * (g.ny / SW_TY) * (g.nz / outer25_kz(g));
    return nb <= NYL_RED_BLOCKS;
}

bool outer_bsq_fits(const Geo &g) {
    if (!sweep25_fits(g)) return false;
    const int64_t nb = (int64_t)(g.nx / SW_TX) * (g.ny / SW_TY) * (g.nz / outer25_kz(g));
    return 2 * nb <= NYL_RED_BLOCKS;
}

template <typename TI>
hipError_t launch_mg_outer_resid(const Geo &g, const double *x, const double *b, double shift,
                                 TI *r, TI *y, double omega, const MGTabs &tab, double *partials,
                                 double *dev_out, hipStream_t s, bool with_bsq) {
    if (outer25_fits(g)) {
        const int kz = outer25_kz(g);
        const dim3 grd(g.nx / SW_TX, g.ny / SW_TY, g.nz / kz), blk(SW_TX, 4);
        const int nb = grd.x * grd.y * grd.z;
        if (with_bsq && 2 * nb > NYL_RED_BLOCKS) return hipErrorInvalidValue;
        double *pb = with_bsq ? partials + nb : nullptr;
        NYL_LAUNCH((k_mg_outer25<TI>), grd, blk, 0, s, g, x, b, shift, r, y, omega, tab,
                           kz, partials, pb);
        (void)launch_final_sum(partials, nb, dev_out, s);
        if (with_bsq)
            (void)launch_final_sum(pb, nb, dev_out + 1, s);
        return hipGetLastError();
    }
    if (with_bsq) return hipErrorInvalidValue;
    const dim3 grd = outer_grid(g);
    NYL_LAUNCH((k_mg_outer_resid<TI>), grd, dim3(MG_BX, MG_BY, 1), 0, s, g, x, b, shift,
                       r, y, omega, tab, partials);
    const int nb = grd.x * grd.y * grd.z;
    (void)launch_final_sum(partials, nb, dev_out, s);
    return hipGetLastError();
}

// Prolongation with 2.5-D tiling: a block owns fine 64 x 8 cells (rows
// j0 + ty + 4r of column i0 + tx) and marches up kz fine planes (kz even).
// The coarse values its cells use (a 34 x 6 footprint per coarse plane,
// zero outside a wall axis, wrapped along a periodic one: what the
// guarded reads of prolong_term return) sit in a five-plane LDS ring: the
// fine pair (2kc, 2kc+1) reads coarse planes kc-1, kc, kc+1 while plane
// kc+2 is stored, so one barrier per coarse plane separates a slot's last
// reader from its next writer.  The next coarse value and the next fine
// pair are loaded one coarse plane ahead.  Each fine cell evaluates
// prolong_term's expression in its order: bitwise the same.
#define PR_CX (SW_TX / 2 + 2)
#define PR_CY (SW_TY / 2 + 2)
#define PR_RING 5
template <typename Tf, typename Tc>
__global__ __launch_bounds__(256) void k_mg_prolong25(Geo gf, Geo gc, Tf *__restrict__ xf,
                                                      const Tc *__restrict__ xc, MGTabs tab,
                                                      int kz, Tf *__restrict__ yf) {
    __shared__ Tf cp[PR_RING][PR_CY][PR_CX];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * SW_TX + tx;
    const Blk3 bk = xcd_block();
    const int i0 = bk.x * SW_TX, j0 = bk.y * SW_TY, k0 = bk.z * kz;
    const int i = i0 + tx;
    const int ci0 = (i0 >> 1) - 1, cj0 = (j0 >> 1) - 1;
    // coarse footprint cell loaded by this thread (one per coarse plane)
    const bool has_c = tid < PR_CX * PR_CY;
    int coff = 0;   // a valid offset; cin: the cell is in the domain
    bool cin = false;
    int clx = 0, cly = 0;
    if (has_c) {
        clx = tid % PR_CX;
        cly = tid / PR_CX;
        int gi = ci0 + clx, gj = cj0 + cly;
        bool in = true;
        if (gc.px) gi = wrapi(gi, gc.nx); else in = in && gi >= 0 && gi < gc.nx;
        if (gc.py) gj = wrapi(gj, gc.ny); else in = in && gj >= 0 && gj < gc.ny;
        if (in) { coff = gj * (int)gc.sy + gi; cin = true; }
    }
    auto czin = [&](int m) { return gc.pz || (m >= 0 && m < gc.nz); };
    // the coarse value of plane m (from a valid address; zero outside)
    auto ldc = [&](int m) -> Tf {
        const bool zi = czin(m);
        const int mw = gc.pz ? wrapg(m, gc.nz) : (zi ? m : 0);
        const Tf v = Tf(xc[(int64_t)mw * gc.sz + coff]);
        return (zi && cin) ? v : Tf(0);
    };
    const int kc0 = k0 >> 1;
    auto slot = [&](int m) { return (m - kc0 + 1) % PR_RING; };   // m >= kc0 - 1
    auto putc = [&](int m, Tf v) {
        if (has_c) cp[slot(m)][cly][clx] = v;
    };
    // x and y parts of the stencil, fixed per thread and row; the two
    // coefficients for the z side in or out of the domain
    const int ic = i >> 1;
    const int icn = (i & 1) ? ic + 1 : ic - 1;
    const int mx = (gc.px || (icn >= 0 && icn < gc.nx)) ? 1 : 0;
    const int lx = ic - ci0, lxn = icn - ci0;
    int ly[SW_R], lyn[SW_R], off[SW_R];
    Tf pc0[SW_R], pc1[SW_R];
#pragma unroll
    for (int r = 0; r < SW_R; r++) {
        const int j = j0 + ty + 4 * r;
        const int jc = j >> 1;
        const int jcn = (j & 1) ? jc + 1 : jc - 1;
        const int my = (gc.py || (jcn >= 0 && jcn < gc.ny)) ? 1 : 0;
        ly[r] = jc - cj0;
        lyn[r] = jcn - cj0;
        off[r] = j * (int)gf.sy + i;
        pc0[r] = Tf(tab.pcoef[mx + 2 * my]);
        pc1[r] = Tf(tab.pcoef[mx + 2 * my + 4]);
    }
    const int kce = kc0 + kz / 2;
    putc(kc0 - 1, ldc(kc0 - 1));
    putc(kc0, ldc(kc0));
    putc(kc0 + 1, ldc(kc0 + 1));
    Tf cnext = ldc(kc0 + 2);
    Tf fv[2][SW_R];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int r = 0; r < SW_R; r++) fv[h][r] = xf[(int64_t)(k0 + h) * gf.sz + off[r]];
    for (int kc = kc0; kc < kce; kc++) {
        const bool more = kc + 1 < kce;
        // coarse plane kc+2 into the slot of kc-3, last read two planes ago
        if (more) putc(kc + 2, cnext);
        Tf fc[2][SW_R];
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int r = 0; r < SW_R; r++) fc[h][r] = fv[h][r];
        if (more) {
            if (kc + 2 < kce) cnext = ldc(kc + 3);
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int r = 0; r < SW_R; r++)
                    fv[h][r] = xf[(int64_t)(2 * kc + 2 + h) * gf.sz + off[r]];
        }
        __syncthreads();
        const int sm = slot(kc - 1), s0 = slot(kc), sp = slot(kc + 1);
#pragma unroll
        for (int h = 0; h < 2; h++) {              // fine planes 2kc (side kc-1), 2kc+1 (side kc+1)
            const int k = 2 * kc + h;
            const int kcn = h ? kc + 1 : kc - 1;
            const bool mz = gc.pz || (kcn >= 0 && kcn < gc.nz);
            const int sn = h ? sp : sm;
#pragma unroll
            for (int r = 0; r < SW_R; r++) {
                auto A = [&](int s) -> Tf {
                    return Tf(9) * cp[s][ly[r]][lx] + Tf(3) * cp[s][ly[r]][lxn]
                         + Tf(3) * cp[s][lyn[r]][lx] + cp[s][lyn[r]][lxn];
                };
                const Tf bmid = A(s0);
                const Tf aside = A(sn);
                const Tf term = (mz ? pc1[r] : pc0[r]) * (Tf(3) * bmid + aside);
                const int64_t c = (int64_t)k * gf.sz + off[r];
                const Tf v = fc[h][r] + term;
                if (yf) yf[c] = v;
                else xf[c] = v;
            }
        }
    }
}

template <typename Tf, typename Tc>
hipError_t launch_mg_prolong25(const Geo &gf, const Geo &gc, Tf *xf, const Tc *xc,
                               const MGTabs &tab, hipStream_t s, Tf *yf) {
    const int kz = rp25_kz(gf, (int)sizeof(Tf));
    dim3 grd(gf.nx / SW_TX, gf.ny / SW_TY, gf.nz / kz), blk(SW_TX, 4);
    NYL_LAUNCH((k_mg_prolong25<Tf, Tc>), grd, blk, 0, s, gf, gc, xf, xc, tab, kz, yf);
    return hipGetLastError();
}

// Two consecutive smoothing sweeps of k_mg_sweep25 (src -> tmp -> dst) in
// one 2.5-D march.  The first sweep is evaluated on the 64 x 8 tile and a
// one-cell ring (its input held with a two-cell ring in LDS) and kept for
// three planes; the second sweep is evaluated on the tile.  Each value has
// the per-cell sweep's expression (zero for an absent neighbour), so dst is
// bitwise that of the two separate sweeps.  src and dst must differ.
#define S2_LX (SW_TX + 4)
#define S2_LY (SW_TY + 4)
#define S2_MX (SW_TX + 2)
#define S2_MY (SW_TY + 2)
template <typename T>
__global__ __launch_bounds__(256) void k_mg_sweep2x25(Geo g, const T *__restrict__ src,
                                                      T *__restrict__ dst,
                                                      const T *__restrict__ b, double omega,
                                                      MGTabs tab, int kz,
                                                      double *__restrict__ xo) {
    constexpr int NIN = S2_LY * S2_LX, NMID = S2_MY * S2_MX;
    constexpr int RIN = (NIN + 255) / 256, RMID = (NMID + 255) / 256;
    __shared__ T pin[3][NIN];
    __shared__ T mid[3][NMID];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * SW_TX + tx;
    const Blk3 bk = xcd_block();
    const int i0 = bk.x * SW_TX, j0 = bk.y * SW_TY, k0 = bk.z * kz;
    const T om = T(omega), cff1 = T(1.0 - omega);
    // per-thread positions, fixed for the march: input ring cells (in-plane
    // offset, or -1 outside a wall axis) ...
    int in_off[RIN];
#pragma unroll
    for (int r = 0; r < RIN; r++) {
        const int q = tid + 256 * r;
        const int lx = q % S2_LX, ly = q / S2_LX;
        int gi = i0 - 2 + lx, gj = j0 - 2 + ly;
        bool in = q < NIN;
        if (g.px) gi = wrapg(gi, g.nx); else in = in && gi >= 0 && gi < g.nx;
        if (g.py) gj = wrapg(gj, g.ny); else in = in && gj >= 0 && gj < g.ny;
        in_off[r] = in ? gj * (int)g.sy + gi : -1;
    }
    // ... and first-sweep cells (in-plane offset or -1, in-plane neighbour count)
    int mid_off[RMID], mid_nxy[RMID];
#pragma unroll
    for (int r = 0; r < RMID; r++) {
        const int q = tid + 256 * r;
        const int lx = q % S2_MX, ly = q / S2_MX;
        int gi = i0 - 1 + lx, gj = j0 - 1 + ly;
        bool in = q < NMID;
        if (g.px) gi = wrapg(gi, g.nx); else in = in && gi >= 0 && gi < g.nx;
        if (g.py) gj = wrapg(gj, g.ny); else in = in && gj >= 0 && gj < g.ny;
        mid_off[r] = in ? gj * (int)g.sy + gi : -1;
        mid_nxy[r] = (g.px || gi > 0) + (g.px || gi < g.nx - 1) + (g.py || gj > 0) +
                     (g.py || gj < g.ny - 1);
    }
    int out_nxy[SW_R];
#pragma unroll
    for (int r = 0; r < SW_R; r++) {
        const int i = i0 + tx, j = j0 + ty + 4 * r;
        out_nxy[r] = (g.px || i > 0) + (g.px || i < g.nx - 1) + (g.py || j > 0) +
                     (g.py || j < g.ny - 1);
    }
    // inverse diagonals for 0, 1, 2 in-domain z-neighbours (hoisted out of
    // the march: the table is a kernel argument, read per lane otherwise)
    T mid_id[3][RMID], out_id[3][SW_R];
#pragma unroll
    for (int q = 0; q < 3; q++) {
#pragma unroll
        for (int r = 0; r < RMID; r++) mid_id[q][r] = T(tab.idiag[mid_nxy[r] + q]);
#pragma unroll
        for (int r = 0; r < SW_R; r++) out_id[q][r] = T(tab.idiag[out_nxy[r] + q]);
    }
    auto zin = [&](int m) { return g.pz || (m >= 0 && m < g.nz); };
    auto zw = [&](int m) { return g.pz ? wrapg(m, g.nz) : (zin(m) ? m : 0); };
    auto nzw_of = [&](int m) {   // in-domain z-neighbours of plane m (wrapped)
        const int mw = g.pz ? wrapg(m, g.nz) : m;
        return (g.pz || mw > 0) + (g.pz || mw < g.nz - 1);
    };
    auto pick = [](int nz_, const T (&t)[3]) { return nz_ == 2 ? t[2] : (nz_ == 1 ? t[1] : t[0]); };
    auto slot = [&](int m) { return (m - k0 + 3) % 3; };   // m >= k0 - 2
    // Loads run ahead of their use: every load is issued from a valid
    // address (in-plane offset 0, plane 0 outside a wall axis) and the value
    // selected when it is used, so no branch forces a wait.
    int in_ld[RIN], mid_ld[RMID];
#pragma unroll
    for (int r = 0; r < RIN; r++) in_ld[r] = in_off[r] >= 0 ? in_off[r] : 0;
#pragma unroll
    for (int r = 0; r < RMID; r++) mid_ld[r] = mid_off[r] >= 0 ? mid_off[r] : 0;
    const int out_ld0 = (j0 + ty) * (int)g.sy + i0 + tx;
    auto fetch_in = [&](int m, T (&v)[RIN]) {
        const T *p = src + (int64_t)zw(m) * g.sz;
#pragma unroll
        for (int r = 0; r < RIN; r++) v[r] = p[in_ld[r]];
    };
    auto put_in = [&](int m, const T (&v)[RIN]) {
        T *P = pin[slot(m)];
        const bool zi = zin(m);
#pragma unroll
        for (int r = 0; r < RIN; r++) {
            const int q = tid + 256 * r;
            if (q < NIN) P[q] = (zi && in_off[r] >= 0) ? v[r] : T(0);
        }
    };
    auto fetch_bmid = [&](int m, T (&v)[RMID]) {
        const T *p = b + (int64_t)zw(m) * g.sz;
#pragma unroll
        for (int r = 0; r < RMID; r++) v[r] = p[mid_ld[r]];
    };
    auto fetch_bout = [&](int k, T (&v)[SW_R], double (&xv)[SW_R]) {
        const int64_t o = (int64_t)k * g.sz + out_ld0;
#pragma unroll
        for (int r = 0; r < SW_R; r++) {
            v[r] = b[o + (int64_t)4 * r * g.sy];
            if (xo) xv[r] = xo[o + (int64_t)4 * r * g.sy];
        }
    };
    auto comp_mid = [&](int m, const T (&bv)[RMID]) {
        T *M = mid[slot(m)];
        const T *P0 = pin[slot(m)], *Pm = pin[slot(m - 1)], *Pp = pin[slot(m + 1)];
        const bool zi = zin(m);
        const int nzw = nzw_of(m);
#pragma unroll
        for (int r = 0; r < RMID; r++) {
            const int q = tid + 256 * r;
            if (q >= NMID) continue;
            T v = T(0);
            if (zi && mid_off[r] >= 0) {
                const int lx = q % S2_MX, ly = q / S2_MX;
                const int p = (ly + 1) * S2_LX + (lx + 1);
                const T S = P0[p - 1] + P0[p + 1] + P0[p - S2_LX] + P0[p + S2_LX] + Pm[p] + Pp[p];
                T idg[3] = {mid_id[0][r], mid_id[1][r], mid_id[2][r]};
                v = cff1 * P0[p] + om * (S - bv[r]) * pick(nzw, idg);
            }
            M[q] = v;
        }
    };
    auto comp_out = [&](int k, const T (&bv)[SW_R], const double (&xv)[SW_R]) {
        const T *M0 = mid[slot(k)], *Mm = mid[slot(k - 1)], *Mp = mid[slot(k + 1)];
        const int nzw = (g.pz || k > 0) + (g.pz || k < g.nz - 1);
        const int i = i0 + tx;
#pragma unroll
        for (int r = 0; r < SW_R; r++) {
            const int j = j0 + ty + 4 * r;
            const int p = (ty + 4 * r + 1) * S2_MX + (tx + 1);
            const int64_t c = (int64_t)k * g.sz + (int64_t)j * g.sy + i;
            const T S = M0[p - 1] + M0[p + 1] + M0[p - S2_MX] + M0[p + S2_MX] + Mm[p] + Mp[p];
            T idg[3] = {out_id[0][r], out_id[1][r], out_id[2][r]};
            const T v = cff1 * M0[p] + om * (S - bv[r]) * pick(nzw, idg);
            if (xo) xo[c] = xv[r] + (double)v;   // ADDX of k_mg_sweep25
            else dst[c] = v;
        }
    };
    // prologue: input planes k0-2 .. k0+1 in LDS, first sweep of k0-1, k0
    T vin[RIN], bm[RMID], bo[SW_R];
    double xv[SW_R] = {};
    {
        T v0[RIN], v1[RIN], v2[RIN];
        fetch_in(k0 - 2, v0);
        fetch_in(k0 - 1, v1);
        fetch_in(k0, v2);
        fetch_in(k0 + 1, vin);
        T b0[RMID];
        fetch_bmid(k0 - 1, b0);
        fetch_bmid(k0, bm);
        put_in(k0 - 2, v0);
        put_in(k0 - 1, v1);
        put_in(k0, v2);
        __syncthreads();
        comp_mid(k0 - 1, b0);
        __syncthreads();          // plane k0+1 replaces plane k0-2, just read
        put_in(k0 + 1, vin);
        __syncthreads();
        comp_mid(k0, bm);
    }
    fetch_in(k0 + 2, vin);
    fetch_bmid(k0 + 1, bm);
    fetch_bout(k0, bo, xv);
    __syncthreads();              // comp_mid(k0) read plane k0-1, replaced next
    for (int k = k0; k < k0 + kz; k++) {
        // plane k+2 of the input replaces plane k-1, whose last reader (the
        // first sweep of plane k) passed the barrier before comp_out(k-1);
        // two barriers per plane
        put_in(k + 2, vin);
        T bmc[RMID], boc[SW_R];
        double xvc[SW_R];
#pragma unroll
        for (int r = 0; r < RMID; r++) bmc[r] = bm[r];
#pragma unroll
        for (int r = 0; r < SW_R; r++) {
            boc[r] = bo[r];
            xvc[r] = xv[r];
        }
        if (k + 1 < k0 + kz) {
            fetch_in(k + 3, vin);
            fetch_bmid(k + 2, bm);
            fetch_bout(k + 1, bo, xv);
        }
        __syncthreads();
        comp_mid(k + 1, bmc);     // replaces plane k - 2 of the first sweep
        __syncthreads();
        comp_out(k, boc, xvc);
    }
}

// Two consecutive sweeps (src -> tmp -> dst) in one 2.5-D march with the
// structure of k_mg_sweep25 (k_mg_sweep2t).  A block owns a 64 x 8 tile and
// thread (tx, ty) its rows ty and ty + 4; each thread keeps in registers the
// z-windows of its input (planes k, k+1, k+2) and of the first sweep
// (k-1, k, k+1), so only in-plane neighbours go through LDS.  The first sweep
// is also needed on the one-cell ring around the tile: its 148 cells belong
// to threads 0..147, which keep the input's z-window there too; the input's
// second ring (156 cells, read in-plane only) belongs to threads 0..155.
// Per plane: input plane k+1 (tile and both rings) into one of two LDS
// slots, one barrier, the first sweep of plane k+1 on the tile and the ring
// into one of two LDS slots, the second sweep of plane k on the tile (its
// in-plane neighbours from the slot written one plane earlier).  Cells
// outside the domain (wall axes) hold zero in both stages, as neighbours()
// reads them, and every value is the per-cell sweep's expression, so dst is
// bitwise that of two sweeps.  ZSRC: the input is zero and src is not read
// (the first sweep is then sweep_from_zero's expression: the pre-smoothing of
// a solve from x = 0).  ADDX: the result is added to the FP64 xo instead of
// stored.  RESNORM: the block's partial sum of the input's squared
// residuals (b + diag x - sum6(x), the first sweep's terms) is stored in
// partials, as k_mg_sweep25 does (the speculative pre-smoothing of the next
// V-cycle, launched with the convergence test).
#define S2T_LX (SW_TX + 4)                       // input plane, two-cell ring
#define S2T_LY (SW_TY + 4)
#define S2T_MX (SW_TX + 2)                       // first-sweep plane, one-cell ring
#define S2T_MY (SW_TY + 2)
#define S2T_RING1 (2 * S2T_MX + 2 * SW_TY)       // 148 cells
#define S2T_RING2 (2 * S2T_LX + 2 * (SW_TY + 2))  // 156 cells
template <typename T, bool ADDX, int PZ, bool RESNORM, bool ZSRC>
__global__ __launch_bounds__(256) void k_mg_sweep2t(Geo g, const T *__restrict__ src,
                                                    T *__restrict__ dst,
                                                    const T *__restrict__ b, double omega,
                                                    MGTabs tab, int kz,
                                                    double *__restrict__ xo,
                                                    double *__restrict__ partials) {
    __shared__ T sin_[2][S2T_LY][S2T_LX];
    __shared__ T smid[2][S2T_MY][S2T_MX];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * SW_TX + tx;
    const Blk3 bk = xcd_block();
    const int i0 = bk.x * SW_TX, j0 = bk.y * SW_TY, k0 = bk.z * kz;
    const T om = T(omega), cff1 = T(1.0 - omega);
    const int i = i0 + tx, sy = (int)g.sy;
    // own cells: in-plane offsets, rows, inverse diagonals for 0, 1, 2
    // in-domain z-neighbours
    int off[SW_R], row[SW_R];
    T id0[SW_R], id1[SW_R], id2[SW_R], dg0[SW_R], dg1[SW_R], dg2[SW_R];
#pragma unroll
    for (int r = 0; r < SW_R; r++) {
        row[r] = ty + 4 * r;
        const int j = j0 + row[r];
        off[r] = j * sy + i;
        const int nxy = (g.px || i > 0) + (g.px || i < g.nx - 1) + (g.py || j > 0) +
                        (g.py || j < g.ny - 1);
        id0[r] = T(tab.idiag[nxy]);
        id1[r] = T(tab.idiag[nxy + 1]);
        id2[r] = T(tab.idiag[nxy + 2]);
        dg0[r] = RESNORM ? T(tab.diag[nxy]) : T(0);
        dg1[r] = RESNORM ? T(tab.diag[nxy + 1]) : T(0);
        dg2[r] = RESNORM ? T(tab.diag[nxy + 2]) : T(0);
    }
    double acc = 0.0;   // RESNORM: sum of the input's squared residuals
    // ring cell of the first sweep (threads 0..147): position (mx, my) in the
    // first-sweep plane, (mx + 1, my + 1) in the input plane
    const bool has1 = tid < S2T_RING1;
    int mx = 0, my = 0;
    if (tid < S2T_MX) { mx = tid; my = 0; }
    else if (tid < 2 * S2T_MX) { mx = tid - S2T_MX; my = S2T_MY - 1; }
    else { const int t = tid - 2 * S2T_MX; my = 1 + (t >> 1); mx = (t & 1) ? S2T_MX - 1 : 0; }
    int off1 = 0;
    bool in1 = false;
    T rid0 = T(0), rid1 = T(0), rid2 = T(0);
    if (has1) {
        int gi = i0 - 1 + mx, gj = j0 - 1 + my;
        bool in = true;
        if (g.px) gi = wrapg(gi, g.nx); else in = in && gi >= 0 && gi < g.nx;
        if (g.py) gj = wrapg(gj, g.ny); else in = in && gj >= 0 && gj < g.ny;
        if (in) {
            off1 = gj * sy + gi;
            in1 = true;
            const int nxy = (g.px || gi > 0) + (g.px || gi < g.nx - 1) + (g.py || gj > 0) +
                            (g.py || gj < g.ny - 1);
            rid0 = T(tab.idiag[nxy]);
            rid1 = T(tab.idiag[nxy + 1]);
            rid2 = T(tab.idiag[nxy + 2]);
        }
    }
    // second ring of the input (threads 0..155): position (lx, ly)
    const bool has2 = tid < S2T_RING2;
    int lx2 = 0, ly2 = 0;
    if (tid < S2T_LX) { lx2 = tid; ly2 = 0; }
    else if (tid < 2 * S2T_LX) { lx2 = tid - S2T_LX; ly2 = S2T_LY - 1; }
    else { const int t = tid - 2 * S2T_LX; ly2 = 1 + (t >> 1); lx2 = (t & 1) ? S2T_LX - 1 : 0; }
    int off2 = 0;
    bool in2 = false;
    if (has2) {
        int gi = i0 - 2 + lx2, gj = j0 - 2 + ly2;
        bool in = true;
        if (g.px) gi = wrapg(gi, g.nx); else in = in && gi >= 0 && gi < g.nx;
        if (g.py) gj = wrapg(gj, g.ny); else in = in && gj >= 0 && gj < g.ny;
        if (in) { off2 = gj * sy + gi; in2 = true; }
    }

    // z bookkeeping (PZ: periodic z, a compile-time flag so that the march
    // carries no wrap arithmetic): in-domain test and in-domain z-neighbour
    // count of plane m (the count only matters for in-domain planes)
    auto zin = [&](int m) { return PZ || (m >= 0 && m < g.nz); };
    auto nzw = [&](int m) { return PZ ? 2 : (m > 0) + (m < g.nz - 1); };
    // a plane that can be loaded for plane m (the wrapped one; any plane
    // outside a wall axis, whose values are not used), and the one for m + 1
    auto pl = [&](int m) { return PZ ? wrapg(m, g.nz) : (zin(m) ? m : 0); };
    auto nextp = [&](int q) { return PZ ? (q + 1 == g.nz ? 0 : q + 1) : (q + 1 < g.nz ? q + 1 : q); };
    auto pick = [](int nz_, T a0, T a1, T a2) { return nz_ == 2 ? a2 : (nz_ == 1 ? a1 : a0); };
    // loads from valid addresses; the values of out-of-domain cells are
    // replaced by zero where they are used
    // loads of plane q (a loadable plane index)
    auto ld_in = [&](int q, T (&o)[SW_R], T &v1) {
        if (ZSRC) {   // the input is zero: nothing is read
#pragma unroll
            for (int r = 0; r < SW_R; r++) o[r] = T(0);
            v1 = T(0);
            return;
        }
        const T *p = src + (int64_t)q * g.sz;
#pragma unroll
        for (int r = 0; r < SW_R; r++) o[r] = p[off[r]];
        v1 = p[off1];
    };
    auto ld_b = [&](int q, T (&o)[SW_R], T &v1) {
        const T *p = b + (int64_t)q * g.sz;
#pragma unroll
        for (int r = 0; r < SW_R; r++) o[r] = p[off[r]];
        v1 = p[off1];
    };
    auto ld_in2 = [&](int q) { return ZSRC ? T(0) : src[(int64_t)q * g.sz + off2]; };
    // input plane m into its LDS slot
    auto put_in = [&](int m, const T (&o)[SW_R], T v1, T v2, bool ring2) {
        const bool zi = zin(m);
        T (*P)[S2T_LX] = sin_[m & 1];
#pragma unroll
        for (int r = 0; r < SW_R; r++) P[row[r] + 2][tx + 2] = zi ? o[r] : T(0);
        if (has1) P[my + 1][mx + 1] = (zi && in1) ? v1 : T(0);
        if (ring2 && has2) P[ly2][lx2] = (zi && in2) ? v2 : T(0);
    };
    // first sweep of plane m on the own cells (zm_, zp_: input planes m -+ 1);
    // own: m is one of the block's planes (RESNORM: its residuals are summed,
    // as in k_mg_sweep25)
    auto mid_own = [&](int m, const T (&c_)[SW_R], const T (&zm_)[SW_R], const T (&zp_)[SW_R],
                       const T (&bv)[SW_R], T (&out)[SW_R], bool store, bool own) {
        const bool zi = zin(m), zl = zin(m - 1), zh = zin(m + 1);
        const int nz_ = nzw(m);
        const T (*P)[S2T_LX] = sin_[m & 1];
#pragma unroll
        for (int r = 0; r < SW_R; r++) {
            const int ly = row[r] + 2, lx = tx + 2;
            const T S = P[ly][lx - 1] + P[ly][lx + 1] + P[ly - 1][lx] + P[ly + 1][lx] +
                        (zl ? zm_[r] : T(0)) + (zh ? zp_[r] : T(0));
            if (RESNORM && own) {
                const T d = pick(nz_, dg0[r], dg1[r], dg2[r]);
                const double rr = (double)(bv[r] + d * c_[r] - S);
                acc = acc + rr * rr;
            }
            const T v = cff1 * c_[r] + om * (S - bv[r]) * pick(nz_, id0[r], id1[r], id2[r]);
            out[r] = zi ? v : T(0);
            if (store) smid[m & 1][row[r] + 1][tx + 1] = out[r];
        }
    };
    // first sweep of plane m on the ring cell
    auto mid_ring = [&](int m, T c_, T zm_, T zp_, T bv) {
        if (!has1) return;
        const bool zi = zin(m), zl = zin(m - 1), zh = zin(m + 1);
        const int nz_ = nzw(m);
        const T (*P)[S2T_LX] = sin_[m & 1];
        const int ly = my + 1, lx = mx + 1;
        const T S = P[ly][lx - 1] + P[ly][lx + 1] + P[ly - 1][lx] + P[ly + 1][lx] +
                    (zl ? zm_ : T(0)) + (zh ? zp_ : T(0));
        const T v = cff1 * c_ + om * (S - bv) * pick(nz_, rid0, rid1, rid2);
        smid[m & 1][my][mx] = (zi && in1) ? v : T(0);
    };

    // prologue: input planes k0-2 .. k0+2, first sweep of k0-1 (own) and k0
    T A0[SW_R], A1[SW_R], A2[SW_R], A3[SW_R], A4[SW_R];
    T a0, a1, a2, a3, a4;
    ld_in(pl(k0 - 2), A0, a0);
    ld_in(pl(k0 - 1), A1, a1);
    ld_in(pl(k0), A2, a2);
    ld_in(pl(k0 + 1), A3, a3);
    ld_in(pl(k0 + 2), A4, a4);
    const T o2a = ld_in2(pl(k0));
    T o2 = ld_in2(pl(k0 + 1));
    T Bp[SW_R], Bo[SW_R], Bm[SW_R], bp1, bo1, bm1;
    ld_b(pl(k0 - 1), Bp, bp1);
    ld_b(pl(k0), Bo, bo1);
    ld_b(pl(k0 + 1), Bm, bm1);
    int q3 = pl(k0 + 3), q2 = pl(k0 + 2);   // planes of the loads in the march
    put_in(k0 - 1, A1, a1, T(0), false);
    put_in(k0, A2, a2, o2a, true);
    __syncthreads();
    T Mm[SW_R], M0[SW_R], Mp[SW_R];
    mid_own(k0 - 1, A1, A0, A2, Bp, Mm, false, false);
    mid_own(k0, A2, A1, A3, Bo, M0, true, true);
    mid_ring(k0, a2, a1, a3, bo1);
    __syncthreads();              // plane k0+1 replaces plane k0-1, just read
    // z-windows: input k, k+1, k+2 (own and ring), b of k (output) and k+1
    T I0[SW_R], I1[SW_R], I2[SW_R];
    T r0 = a2, r1 = a3, r2 = a4;
#pragma unroll
    for (int r = 0; r < SW_R; r++) { I0[r] = A2[r]; I1[r] = A3[r]; I2[r] = A4[r]; }
    double xv[SW_R] = {};
    if (ADDX) {
#pragma unroll
        for (int r = 0; r < SW_R; r++) xv[r] = xo[(int64_t)k0 * g.sz + off[r]];
    }
    for (int k = k0; k < k0 + kz; k++) {
        put_in(k + 1, I1, r1, o2, true);
        // loads for the next plane: input k+3, second ring k+2, b of k+2
        T I3[SW_R] = {}, Bn[SW_R] = {}, r3 = T(0), bn1 = T(0), o2n = T(0);
        double xvn[SW_R] = {};
        const bool more = k + 1 < k0 + kz;
        if (more) {
            ld_in(q3, I3, r3);
            o2n = ld_in2(q2);
            ld_b(q2, Bn, bn1);
            if (ADDX) {
#pragma unroll
                for (int r = 0; r < SW_R; r++) xvn[r] = xo[(int64_t)(k + 1) * g.sz + off[r]];
            }
        }
        __syncthreads();
        mid_own(k + 1, I1, I0, I2, Bm, Mp, true, k + 1 < k0 + kz);
        mid_ring(k + 1, r1, r0, r2, bm1);
        // second sweep of plane k (in-plane neighbours from the slot of k)
        {
            const int nz_ = nzw(k);
            const T (*Q)[S2T_MX] = smid[k & 1];
#pragma unroll
            for (int r = 0; r < SW_R; r++) {
                const int ly = row[r] + 1, lx = tx + 1;
                const T S = Q[ly][lx - 1] + Q[ly][lx + 1] + Q[ly - 1][lx] + Q[ly + 1][lx] +
                            Mm[r] + Mp[r];
                const T v = cff1 * M0[r] + om * (S - Bo[r]) * pick(nz_, id0[r], id1[r], id2[r]);
                const int64_t c = (int64_t)k * g.sz + off[r];
                if (ADDX) xo[c] = xv[r] + (double)v;
                else dst[c] = v;
            }
        }
#pragma unroll
        for (int r = 0; r < SW_R; r++) {
            Mm[r] = M0[r];
            M0[r] = Mp[r];
            I0[r] = I1[r];
            I1[r] = I2[r];
            I2[r] = I3[r];
            Bo[r] = Bm[r];
            Bm[r] = Bn[r];
            if (ADDX) xv[r] = xvn[r];
        }
        r0 = r1;
        r1 = r2;
        r2 = r3;
        bm1 = bn1;
        o2 = o2n;
        q3 = nextp(q3);
        q2 = nextp(q2);
    }
    if (RESNORM) {
        __syncthreads();
        const int nb = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
        block_sum_store(acc, partials + nb, tid);
    }
}

// planes per block of the two-sweep pass: the longest march (at most 16
// planes in FP32, 32 in FP64, dividing nz) that keeps at least 1024 blocks;
// each block re-reads four input planes.  256^3 (k_mg_sweep2t): FP32 37.8 us
// at 16 planes (41.0 at 32), FP64 93.4 us at 32 (95.6 at 16).
static int sweep2x_kz(const Geo &g, int es) {
    const int64_t cols = (int64_t)(g.nx / SW_TX) * (g.ny / SW_TY);
    int kz = es == 4 ? 16 : 32;
    while (kz > 1 && (g.nz % kz != 0 || cols * (g.nz / kz) < 1024)) kz >>= 1;
    return kz;
}

// the two-sweep pass with register z-windows (k_mg_sweep2t) in place of
// k_mg_sweep2x25
#ifndef NYL_SWEEP2T
#define NYL_SWEEP2T 1
#endif
template <typename T>
hipError_t launch_mg_sweep2x25(const Geo &g, const T *src, T *dst, const T *b, double omega,
                               const MGTabs &tab, hipStream_t s, double *addx,
                               double *partials, double *dev_out) {
    if (src == dst || (addx && partials) || (!src && (addx || partials)))
        return hipErrorInvalidValue;
    const int kz = sweep2x_kz(g, (int)sizeof(T));
    dim3 grd(g.nx / SW_TX, g.ny / SW_TY, g.nz / kz), blk(SW_TX, 4);
    if (NYL_SWEEP2T) {
#define S2T_GO(A, Z, R)                                                                    \
    NYL_LAUNCH((k_mg_sweep2t<T, A, Z, R, false>), grd, blk, 0, s, g, src, dst, b, omega, tab, \
               kz, addx, partials)
        if (!src) {   // from x = 0
            if (g.pz)
                NYL_LAUNCH((k_mg_sweep2t<T, false, 1, false, true>), grd, blk, 0, s, g, src, dst,
                           b, omega, tab, kz, addx, partials);
            else
                NYL_LAUNCH((k_mg_sweep2t<T, false, 0, false, true>), grd, blk, 0, s, g, src, dst,
                           b, omega, tab, kz, addx, partials);
        } else if (partials) {
            if (g.pz) S2T_GO(false, 1, true);
            else S2T_GO(false, 0, true);
            (void)launch_final_sum(partials, (int)(grd.x * grd.y * grd.z), dev_out, s);
        } else if (addx) {
            if (g.pz) S2T_GO(true, 1, false);
            else S2T_GO(true, 0, false);
        } else {
            if (g.pz) S2T_GO(false, 1, false);
            else S2T_GO(false, 0, false);
        }
#undef S2T_GO
        return hipGetLastError();
    }
    if (partials || !src) return hipErrorInvalidValue;   // these need k_mg_sweep2t
    NYL_LAUNCH((k_mg_sweep2x25<T>), grd, blk, 0, s, g, src, dst, b, omega, tab, kz, addx);
    return hipGetLastError();
}

// blocks of the two-sweep pass (its residual norm needs as many partials)
int sweep2x_blocks(const Geo &g, int es) {
    return (g.nx / SW_TX) * (g.ny / SW_TY) * (g.nz / sweep2x_kz(g, es));
}
template hipError_t launch_mg_sweep2x25<double>(const Geo &, const double *, double *,
                                                const double *, double, const MGTabs &,
                                                hipStream_t, double *, double *, double *);
template hipError_t launch_mg_sweep2x25<float>(const Geo &, const float *, float *,
                                               const float *, double, const MGTabs &,
                                               hipStream_t, double *, double *, double *);

bool restrict25_fits(const Geo &g) {
    return sweep25_fits(g) && g.nz % 2 == 0;
}

template <typename Tf, typename Tc>
hipError_t launch_mg_restrict25(const Geo &gf, const Geo &gc, const Tf *x, const Tf *b, Tc *bc,
                                Tc *xc, const MGTabs &tab, hipStream_t s, Tc *yc, double omega) {
    const int kz = rp25_kz(gf, (int)sizeof(Tf));
    dim3 grd(gf.nx / SW_TX, gf.ny / SW_TY, gf.nz / kz), blk(SW_TX, 4);
    NYL_LAUNCH((k_mg_restrict25<Tf, Tc>), grd, blk, 0, s, gf, gc, x, b, bc, xc, tab, kz,
                       yc, omega);
    return hipGetLastError();
}

#define NYL_MG_INST(T)                                                                          \
    template hipError_t launch_mg_sweep25<T>(const Geo &, const T *, T *, const T *, double,    \
                                             const MGTabs &, double *, double *, hipStream_t,   \
                                             const Geo *, const T *, double *);                 \
    template hipError_t launch_mg_prolong25<T, T>(const Geo &, const Geo &, T *, const T *,     \
                                                  const MGTabs &, hipStream_t, T *);            \
    template hipError_t launch_mg_restrict25<T, T>(const Geo &, const Geo &, const T *,         \
                                                   const T *, T *, T *, const MGTabs &,         \
                                                   hipStream_t, T *, double);                   \
    template hipError_t launch_mg_sweep<T>(const Geo &, const T *, T *, const T *, double,      \
                                           const MGTabs &, hipStream_t);                        \
    template hipError_t launch_mg_resid_restrict<T, T>(const Geo &, const Geo &, const T *,     \
                                                       const T *, T *, T *, const MGTabs &,     \
                                                       hipStream_t, T *, double);               \
    template hipError_t launch_mg_prolong<T, T>(const Geo &, const Geo &, T *, const T *,       \
                                                const MGTabs &, hipStream_t);                   \
    template hipError_t launch_norm2<T>(const Geo &, const T *, double *, double *, hipStream_t); \
    template hipError_t launch_maxu2<T>(const Geo &, const T *, const T *, const T *, double,   \
                                        double, double, double *, double *, hipStream_t);       \
    template hipError_t launch_sum<T>(const Geo &, const T *, double *, double *, hipStream_t); \
    template hipError_t launch_nonfinite<T>(const Geo &, const T *, const T *, const T *,       \
                                            const T *, double *, double *, hipStream_t);        \
    template hipError_t launch_kesum<T>(const Geo &, const T *, const T *, const T *, double,   \
                                        double, double, double *, double *, hipStream_t);
NYL_MG_INST(double)
NYL_MG_INST(float)
#undef NYL_MG_INST

#define NYL_B_MG_INST(T)                                                                         \
    template hipError_t launch_b_sweep<T>(const SweepItem<T> *, int, int64_t, const MGTabs &,   \
                                          const int *, hipStream_t);                            \
    template hipError_t launch_b_restrict<T, T>(const RestrItem<T, T> *, int, int64_t,          \
                                                const MGTabs &, const int *, hipStream_t);      \
    template hipError_t launch_b_prolong<T, T>(const ProlItem<T, T> *, int, int64_t,            \
                                               const MGTabs &, const int *, hipStream_t);       \
    template hipError_t launch_b_reduce<T>(const RedItem<T> *, int, int64_t, int,               \
                                           const MGTabs &, const int *, hipStream_t);
NYL_B_MG_INST(double)
NYL_B_MG_INST(float)
#undef NYL_B_MG_INST
template hipError_t launch_b_restrict<float, double>(const RestrItem<float, double> *, int,
                                                     int64_t, const MGTabs &, const int *,
                                                     hipStream_t);
template hipError_t launch_b_prolong<float, double>(const ProlItem<float, double> *, int, int64_t,
                                                    const MGTabs &, const int *, hipStream_t);

template hipError_t launch_resnorm2<double>(const Geo &, const double *, const double *,
                                            const MGTabs &, double *, double *, hipStream_t);
template hipError_t launch_resnorm2<float>(const Geo &, const float *, const float *,
                                           const MGTabs &, double *, double *, hipStream_t);
// FP32 inner levels above an FP64 coarsest level
template hipError_t launch_mg_resid_restrict<float, double>(const Geo &, const Geo &,
                                                            const float *, const float *,
                                                            double *, double *, const MGTabs &,
                                                            hipStream_t, double *, double);
template hipError_t launch_mg_prolong<float, double>(const Geo &, const Geo &, float *,
                                                     const double *, const MGTabs &, hipStream_t);
template hipError_t launch_mg_outer_resid<float>(const Geo &, const double *, const double *,
                                                 double, float *, float *, double,
                                                 const MGTabs &, double *, double *, hipStream_t,
                                                 bool);
template hipError_t launch_add_correction<float>(const Geo &, double *, const float *,
                                                 hipStream_t);

}  // namespace nyl
