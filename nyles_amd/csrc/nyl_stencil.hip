// nyl_stencil.hip -- explicit right-hand side, RK3 stage update and the
// diagnostic stencils of the Nyles time step on gfx950.
//
// Every FP64 instantiation reproduces the reference's arithmetic operation
// by operation (the library is compiled with -ffp-contract=off), so its
// results are bit-identical to the oracle's.
#include "nyl_internal.h"
#include <initializer_list>

namespace nyl {

#define NYL_BX 64
#define NYL_BY 4

// Minimum waves per SIMD for the stage kernel (caps its registers).
#ifndef NYL_STAGE_WAVES
#define NYL_STAGE_WAVES 1
#endif
// Sequencing of the flux evaluations of the stage kernel.  Its stencil loads
// are invariant (const __restrict__), so the compiler is free to issue the
// loads of all twelve reconstructions before any arithmetic, which needs
// ~180 registers and leaves 2 waves per SIMD.  STAGE_SEQ(idx, dep) passes
// the line index of the next evaluation through an empty asm that also takes
// the previous evaluation's result: the next loads cannot be issued before
// that result exists, each reconstruction holds only its own operands and
// occupancy hides the load latency.  No instruction is emitted.
#ifndef NYL_STAGE_SEQ
#define NYL_STAGE_SEQ 1
#endif
// block order of the stage kernel: XCD y slabs marched in z (1) or z slabs (0)
#ifndef NYL_STAGE_YSLAB
#define NYL_STAGE_YSLAB 1
#endif
#if NYL_STAGE_SEQ
#define STAGE_SEQ(idx, dep) asm volatile("" : "+v"(idx) : "v"(dep))
#else
#define STAGE_SEQ(idx, dep) ((void)0)
#endif
// the sequencing points between the two faces of one tracer axis (off: the
// two reconstructions of an axis overlap; FP32 stage 6 % faster)
#ifndef NYL_STAGE_SEQ_FACES
#define NYL_STAGE_SEQ_FACES 0
#endif
// the sequencing points between the vortex-force terms in FP32: every
// term (1), every second term (2) or none (0: FP32 stage 579 -> 533 us; the
// FP32 operands are half the registers, so the overlap fits)
#ifndef NYL_STAGE_SEQ_VF32
#define NYL_STAGE_SEQ_VF32 0
#endif
// and in FP64 (0: FP64 stage modes 1 and 3 1214 -> 1152 us, mode 2 unchanged)
#ifndef NYL_STAGE_SEQ_VF64
#define NYL_STAGE_SEQ_VF64 0
#endif
#define STAGE_SEQ_V(idx, dep, odd)                                             \
    do {                                                                       \
        if constexpr ((sizeof(T) == 8 && (NYL_STAGE_SEQ_VF64 == 1 ||           \
                                          (NYL_STAGE_SEQ_VF64 == 2 && (odd)))) || \
                      (sizeof(T) == 4 && (NYL_STAGE_SEQ_VF32 == 1 ||           \
                                          (NYL_STAGE_SEQ_VF32 == 2 && (odd)))))  \
            STAGE_SEQ(idx, dep);                                               \
    } while (0)
// the sequencing points between the tracer axes: kept for FP64 (overlapping
// them made the FP64 stage 2 % slower), dropped for FP32 (2 % faster)
#define STAGE_SEQ_A(idx, dep)                      \
    do {                                           \
        if constexpr (sizeof(T) == 8) STAGE_SEQ(idx, dep); \
    } while (0)
#if NYL_STAGE_SEQ_FACES
#define STAGE_SEQ_F(idx, dep) STAGE_SEQ(idx, dep)
#else
#define STAGE_SEQ_F(idx, dep) ((void)0)
#endif

static inline dim3 cell_grid(const Geo &g) {
    return dim3((g.nx + NYL_BX - 1) / NYL_BX, (g.ny + NYL_BY - 1) / NYL_BY, g.nz);
}

// The per-cell diagnostic kernels (div, correct, diag): each thread takes
// NYL_CELL_KZ consecutive planes of its column (when nz is a multiple), in
// the XCD y-slab order of the stage kernel.
#ifndef NYL_CELL_KZ
#define NYL_CELL_KZ 1
#endif
#ifndef NYL_CELL_YSLAB
#define NYL_CELL_YSLAB 0
#endif
static inline int cell_kz(const Geo &g) { return g.nz % NYL_CELL_KZ == 0 ? NYL_CELL_KZ : 1; }
static inline dim3 cell_grid_kz(const Geo &g) {
    return dim3((g.nx + NYL_BX - 1) / NYL_BX, (g.ny + NYL_BY - 1) / NYL_BY, g.nz / cell_kz(g));
}
#if NYL_CELL_YSLAB
#define CELL_BLOCK() xcd_block_yslab()
#else
#define CELL_BLOCK() xcd_block()
#endif

// Stage-kernel loads and stores at a 32-bit byte offset from the field's
// base pointer (the base is uniform, so the load takes it from scalar
// registers and the offset as one 32-bit register: no 64-bit address
// arithmetic per load).  A field is below 4 GiB (checked at context
// creation); offsets wrap modulo 2^32 to the right address.
template <typename T>
__device__ __forceinline__ T ldb(const T *__restrict__ p, uint32_t boff) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + boff);
}
template <typename T>
__device__ __forceinline__ void stb(T *__restrict__ p, uint32_t boff, T v) {
    *reinterpret_cast<T *>(reinterpret_cast<char *>(p) + boff) = v;
}
template <typename T>
__device__ __forceinline__ void stb_stream(T *__restrict__ p, uint32_t boff, T v) {
    store_stream(reinterpret_cast<T *>(reinterpret_cast<char *>(p) + boff), v);
}

// Vortex-force flux of one term at a cell (core/fortran_vortex_force.f90
// WENO branch, :66-81 and :145-159): the vorticity component `om` is
// upwinded along axis d with the contravariant velocity U_d = ud*cu
// averaged along the target axis e (UU) and then along d (u1d).  `based` is
// the index of the cell's line start along d, `sd` the stride of d, `md` the
// cell's position on that line, `oe` the offset to the e+1 neighbour.  U is
// recomputed from the covariant u with the product U_from_u stores
// (core/cov_to_contra.py:15-17), so it is the same double.
template <typename T, int pd>
__device__ __forceinline__ T vf_flux(const T *__restrict__ ud, T cu, const T *__restrict__ om,
                                     int64_t based, int64_t sd, int md, int nd, int64_t oe) {
    // 32-bit byte offsets from the field bases (a field is below 4 GiB)
    constexpr uint32_t ES = sizeof(T);
    const uint32_t bb = (uint32_t)based * ES, sb = (uint32_t)sd * ES, eb = (uint32_t)oe * ES;
    auto UU = [&](int m) -> T {
        const uint32_t p = bb + (uint32_t)m * sb;
        return T(0.5) * (ldb(ud, p) * cu + ldb(ud, p + eb) * cu);
    };
    T uu_prev;
    if (!pd && md == 0) uu_prev = T(0);   // UU_0 = 0. at the line start
    else uu_prev = UU(pd ? wrapi(md - 1, nd) : md - 1);
    const T u1d = T(0.5) * (uu_prev + UU(md));
    auto q = [&](int m) -> T {            // q(1) = 0., q(k) = vort(k-1)
        if (pd) return ldb(om, bb + (uint32_t)wrapi(m - 1, nd) * sb);
        if (m == 0) return T(0);
        return ldb(om, bb + (uint32_t)(m - 1) * sb);
    };
    return face_flux<T, pd>(md, nd, u1d, q);
}

// Tracer flux divergence along one axis (core/fortran_upwind.f90:77-81):
// d = d + flux(m-1) - flux(m), or d - flux(1) at a wall start.
// U_d = ud*cu as in vf_flux.
template <typename T, int pd>
__device__ __forceinline__ T tracer_axis(T d, const T *__restrict__ q_arr,
                                         const T *__restrict__ ud, T cu, int64_t based,
                                         int64_t sd, int md, int nd) {
    // 32-bit byte offsets from the field bases (a field is below 4 GiB)
    constexpr uint32_t ES = sizeof(T);
    uint32_t bb = (uint32_t)based * ES;
    const uint32_t sb = (uint32_t)sd * ES;
    auto q = [&](int m) -> T { return ldb(q_arr, bb + (uint32_t)(pd ? wrapi(m, nd) : m) * sb); };
    const T fr = face_flux<T, pd>(md, nd, ldb(ud, bb + (uint32_t)md * sb) * cu, q);
    STAGE_SEQ_F(bb, fr);
    if (!pd && md == 0) return d - fr;
    const int ml = pd ? wrapi(md - 1, nd) : md - 1;
    const T fl = face_flux<T, pd>(ml, nd, ldb(ud, bb + (uint32_t)ml * sb) * cu, q);
    return (d + fl) - fr;
}

// ---------------------------------------------------------------------------
// The reference's linear upwind branches (linear=.false. in its build:
// core/fortran_upwind.f90:30, core/fortran_vortex_force.f90:28,104), odd
// orders 1, 3, 5.  REAL(4) literals widened; sums left to right as written.

// Reconstruction at line position j of the values q(.) (a functor over
// positions, wrapped on a periodic axis): RIGHT = false the left-biased qp,
// true the right-biased qm.  On a wall line the order drops near the ends:
// tracer flavour (core/interpolate_tracer.f90:45-80) first order at 0 and
// n-1, third at 1 and n-2; vortex-force flavour (core/interpolate.f90:35-79)
// qp(-1) = 0, first order at 0 and n-2 (and qm at n-1), third at 1 and n-3.
template <typename T, bool VF, bool RIGHT, typename QF>
__device__ __forceinline__ T lin_rec(const QF &q, int j, int n, int order, bool per) {
    const T c1 = T(-(1.0f / 6.0f)), c2 = T(5.0f / 6.0f), c3 = T(2.0f / 6.0f);
    const T b1 = T(2.0f / 60.0f), b2 = T(-(13.0f / 60.0f)), b3 = T(47.0f / 60.0f);
    const T b4 = T(27.0f / 60.0f), b5 = T(-(3.0f / 60.0f));
    int ord = order;
    if (!per) {
        if (VF && !RIGHT && j < 0) return T(0);
        const int last1 = VF ? n - 2 : n - 1;          // first order at 0 and here
        if (j <= 0 || j >= last1) ord = 1;
        else if (order == 5 && (j == 1 || j == last1 - 1)) ord = 3;
    }
    if (ord == 1) return q(j);
    if (ord == 3) {
        if (RIGHT) return c3 * q(j - 1) + c2 * q(j) + c1 * q(j + 1);
        return c1 * q(j - 1) + c2 * q(j) + c3 * q(j + 1);
    }
    if (RIGHT) return b5 * q(j - 2) + b4 * q(j - 1) + b3 * q(j) + b2 * q(j + 1) + b1 * q(j + 2);
    return b1 * q(j - 2) + b2 * q(j - 1) + b3 * q(j) + b4 * q(j + 1) + b5 * q(j + 2);
}

// core/fortran_upwind.f90:34-67: d = (d + fx(m-1)) - fx(m), the face flux
// fx(i) = up(i)*qp(i) + um(i)*qm(i+1) with up/um = 0.5*(U -+ |U|), no flux
// through the ends of a wall line.  U_d = ud*cu as in tracer_axis.
template <typename T, int PD>
__device__ __forceinline__ T lin_tracer_axis(T d, const T *__restrict__ q_arr,
                                             const T *__restrict__ ud, T cu, int64_t based,
                                             int64_t sd, int m, int n, int order) {
    auto pos = [&](int p) -> int64_t { return based + (int64_t)(PD ? wrapi(p, n) : p) * sd; };
    auto q = [&](int p) -> T { return q_arr[pos(p)]; };
    auto face = [&](int i) -> T {
        const T ui = ud[pos(i)] * cu, au = fabs(ui);
        const T up = T(0.5) * (ui + au), um = T(0.5) * (ui - au);
        return up * lin_rec<T, false, false>(q, i, n, order, PD != 0)
             + um * lin_rec<T, false, true>(q, i + 1, n, order, PD != 0);
    };
    const T fxm = (PD || m > 0) ? face(m - 1) : T(0);
    const T fx = (PD || m < n - 1) ? face(m) : T(0);
    return (d + fxm) - fx;
}

// core/fortran_vortex_force.f90:43-65 (SIGN < 0, direc) and :121-144
// (SIGN > 0, flip): res = (res -+ qp(md-1)*up) -+ qm(md)*um, where the
// vorticity is not shifted (vU(k) = vort(k)) and up/um split U_interp, the
// WENO branch's u1d (UU averaged along e, then along d from UU_0 = 0).
template <typename T, int PD, int SIGN>
__device__ __forceinline__ T vf_lin(T res, const T *__restrict__ ud, T cu,
                                    const T *__restrict__ om, int64_t based, int64_t sd, int md,
                                    int nd, int64_t oe, int order) {
    auto pos = [&](int p) -> int64_t { return based + (int64_t)(PD ? wrapi(p, nd) : p) * sd; };
    auto UU = [&](int m) -> T {
        const int64_t p = pos(m);
        return T(0.5) * (ud[p] * cu + ud[p + oe] * cu);
    };
    const T uu_prev = (!PD && md == 0) ? T(0) : UU(md - 1);
    const T ui = T(0.5) * (uu_prev + UU(md)), au = fabs(ui);
    const T up = T(0.5) * (ui + au), um = T(0.5) * (ui - au);
    auto v = [&](int p) -> T { return om[pos(p)]; };
    const T a = lin_rec<T, true, false>(v, md - 1, nd, order, PD != 0) * up;
    const T b = lin_rec<T, true, true>(v, md, nd, order, PD != 0) * um;
    if (SIGN < 0) return (res - a) - b;
    return (res + a) + b;
}

// Periodic axes: the byte offsets of the cells -3..+3 away along the axis
// (wrapped), computed once per cell; every stencil load of that axis is then
// at cb + off[..].  The loads and the arithmetic are those of the general
// path (face_flux's interior stencil), so the results are the same.
struct PerLine {
    int m3, m2, m1, z, p1, p2, p3;   // byte offsets of the cells -3..+3 away
};

__device__ __forceinline__ PerLine per_line(int m, int n, int stride) {
    PerLine L;
    L.m3 = (wrapi(m - 3, n) - m) * stride;
    L.m2 = (wrapi(m - 2, n) - m) * stride;
    L.m1 = (wrapi(m - 1, n) - m) * stride;
    L.z = 0;
    L.p1 = (wrapi(m + 1, n) - m) * stride;
    L.p2 = (wrapi(m + 2, n) - m) * stride;
    L.p3 = (wrapi(m + 3, n) - m) * stride;
    return L;
}

// offset of relative position d (-3..3; a constant at every call site, so
// the offsets stay in registers)
__device__ __forceinline__ int lo(const PerLine &L, int d) {
    switch (d) {
    case -3: return L.m3;
    case -2: return L.m2;
    case -1: return L.m1;
    case 0: return L.z;
    case 1: return L.p1;
    case 2: return L.p2;
    default: return L.p3;
    }
}

// u times the WENO5 reconstruction at the face between relative positions r
// and r + 1, upwinded by the sign of u; the value at relative position p is
// q[c + lo(L, p + qs)] (qs = -1: the vortex force's q(k) = vort(k-1)).
template <typename T>
__device__ __forceinline__ T per_face(T u, const T *__restrict__ q, uint32_t cb, const PerLine &L,
                                      int r, int qs) {
    const uint32_t om2 = cb + lo(L, r - 2 + qs), om1 = cb + lo(L, r - 1 + qs);
    const uint32_t o0 = cb + lo(L, r + qs), op1 = cb + lo(L, r + 1 + qs);
    const uint32_t op2 = cb + lo(L, r + 2 + qs), op3 = cb + lo(L, r + 3 + qs);
    // upwind stencil: offsets selected, then one set of five loads
    const bool pos = u > T(0);
    const uint32_t s0 = pos ? om2 : op3, s1 = pos ? om1 : op2, s2 = pos ? o0 : op1;
    const uint32_t s3 = pos ? op1 : o0, s4 = pos ? op2 : om1;
    return u * weno5<T>(ldb(q, s0), ldb(q, s1), ldb(q, s2), ldb(q, s3), ldb(q, s4));
}

// tracer_axis on a periodic axis (cb: the cell's byte offset)
template <typename T>
__device__ __forceinline__ T tracer_axis_per(T d, const T *__restrict__ q, const T *__restrict__ ud,
                                             T cu, uint32_t cb, const PerLine &L) {
    const T fr = per_face<T>(ldb(ud, cb) * cu, q, cb, L, 0, 0);
    STAGE_SEQ_F(cb, fr);
    const T fl = per_face<T>(ldb(ud, cb + L.m1) * cu, q, cb, L, -1, 0);
    STAGE_SEQ(cb, fl);
    return (d + fl) - fr;
}

// vf_flux on a periodic axis (cb, oe: byte offsets)
template <typename T>
__device__ __forceinline__ T vf_flux_per(const T *__restrict__ ud, T cu, const T *__restrict__ om,
                                         uint32_t cb, const PerLine &L, uint32_t oe) {
    auto UU = [&](uint32_t p) -> T { return T(0.5) * (ldb(ud, p) * cu + ldb(ud, p + oe) * cu); };
    const T uu_prev = UU(cb + L.m1);
    const T u1d = T(0.5) * (uu_prev + UU(cb));
    return per_face<T>(u1d, om, cb, L, 0, -1);
}

// One direction of core/fortran_dissipation.f90:2-34 add_laplacian at a
// cell: d + coef*(fx - fxm), fx = f(m+1) - f(m), fxm = f(m) - f(m-1); on a
// wall axis fxm = 0 at the first cell and fx = 0 at the last (no flux
// through the ends of the line).  om, op: byte offsets of the m-1 and m+1
// neighbours (wrapped on a periodic axis).
template <typename T, int P>
__device__ __forceinline__ T lap_axis(T d, const T *__restrict__ f, uint32_t cb, int m, int n,
                                      int32_t om, int32_t op, T coef) {
    const T c0 = ldb(f, cb);
    const T fx = (P || m < n - 1) ? ldb(f, cb + (uint32_t)op) - c0 : T(0);
    const T fxm = (P || m > 0) ? c0 - ldb(f, cb + (uint32_t)om) : T(0);
    return d + coef * (fx - fxm);
}

// add_viscosity (core/viscosity.py:3-9) for the three components of a cell:
// for each component, the x, y and z directions in that order
template <typename T, int PX, int PY, int PZ>
__device__ __forceinline__ void viscosity_cell(const Geo &g, const StageArgs<T> &a, uint32_t cb,
                                               int i, int j, int k, T &du, T &dv, T &dw) {
    constexpr int ES = (int)sizeof(T);
    const int32_t sx = ES, sy = (int32_t)g.sy * ES, sz = (int32_t)g.sz * ES;
    const int32_t oxm = PX ? (wrapi(i - 1, g.nx) - i) * sx : -sx;
    const int32_t oxp = PX ? (wrapi(i + 1, g.nx) - i) * sx : sx;
    const int32_t oym = PY ? (wrapi(j - 1, g.ny) - j) * sy : -sy;
    const int32_t oyp = PY ? (wrapi(j + 1, g.ny) - j) * sy : sy;
    const int32_t ozm = PZ ? (wrapi(k - 1, g.nz) - k) * sz : -sz;
    const int32_t ozp = PZ ? (wrapi(k + 1, g.nz) - k) * sz : sz;
    const T cx = T(a.vcx), cy = T(a.vcy), cz = T(a.vcz);
    auto lap3 = [&](T d, const T *__restrict__ f) -> T {
        d = lap_axis<T, PX>(d, f, cb, i, g.nx, oxm, oxp, cx);
        d = lap_axis<T, PY>(d, f, cb, j, g.ny, oym, oyp, cy);
        return lap_axis<T, PZ>(d, f, cb, k, g.nz, ozm, ozp, cz);
    };
    du = lap3(du, a.u);
    dv = lap3(dv, a.v);
    dw = lap3(dw, a.w);
}

// Fused LES.rhs (core/model_les.py:129-144) + RK3_SSP stage update
// (core/timescheme.py:199-220) for one cell.
template <typename T, int MODE, int PX, int PY, int PZ>
__device__ __forceinline__ void stage_cell(const Geo &g, const StageArgs<T> &a, int i, int j,
                                           int k) {
    const int64_t c = (int64_t)k * g.sz + (int64_t)j * g.sy + i;
    constexpr int ES = (int)sizeof(T);
    const uint32_t cb = (uint32_t)c * (uint32_t)ES;   // the cell's byte offset
    PerLine Lx, Ly, Lz;
    if constexpr (PX != 0) Lx = per_line(i, g.nx, ES);
    if constexpr (PY != 0) Ly = per_line(j, g.ny, (int)g.sy * ES);
    if constexpr (PZ != 0) Lz = per_line(k, g.nz, (int)g.sz * ES);

    // model variants (tendencies-only mode): the linear model has no vortex
    // force (core/model_les.py:135-136), Euler3d neither advects b
    // (core/model_les_euler.py:131) nor adds buoyancy (core/bernoulli.py:26-32)
    const bool vf = MODE != 0 || !(a.flags & NYL_SF_NO_VF);
    const bool with_b = MODE != 0 || !(a.flags & NYL_SF_NO_B);
    // the reference's linear upwind branches (tendencies-only mode)
    const bool lin_a = MODE == 0 && a.upw_a != 0;
    const bool lin_vf = MODE == 0 && a.upw_vf != 0;

    // tracer b: rhstrac, directions i, j, k (core/tracer.py:63-72)
    const T cx = T(a.idx2), cy = T(a.idy2), cz = T(a.idz2);
    T db = T(0);
    int64_t bx = c - i, by = c - (int64_t)j * g.sy, bz = c - (int64_t)k * g.sz;
    uint32_t cs = cb;
    if (with_b && lin_a) {
        db = lin_tracer_axis<T, PX>(db, a.b, a.u, cx, bx, 1, i, g.nx, a.upw_a);
        db = lin_tracer_axis<T, PY>(db, a.b, a.v, cy, by, g.sy, j, g.ny, a.upw_a);
        db = lin_tracer_axis<T, PZ>(db, a.b, a.w, cz, bz, g.sz, k, g.nz, a.upw_a);
    } else if (with_b) {
        if constexpr (PX != 0) db = tracer_axis_per<T>(db, a.b, a.u, cx, cs, Lx);
        else db = tracer_axis<T, 0>(db, a.b, a.u, cx, bx, 1, i, g.nx);
        STAGE_SEQ_A(by, db);
        STAGE_SEQ_A(cs, db);
        if constexpr (PY != 0) db = tracer_axis_per<T>(db, a.b, a.v, cy, cs, Ly);
        else db = tracer_axis<T, 0>(db, a.b, a.v, cy, by, g.sy, j, g.ny);
        STAGE_SEQ_A(bz, db);
        STAGE_SEQ_A(cs, db);
        if constexpr (PZ != 0) db = tracer_axis_per<T>(db, a.b, a.w, cz, cs, Lz);
        else db = tracer_axis<T, 0>(db, a.b, a.w, cz, bz, g.sz, k, g.nz);
        STAGE_SEQ(bx, db);
        STAGE_SEQ(cs, db);
    }

    // vortex force, triplets 'ikj', 'jik', 'kji' (core/vortex_force.py:69-81)
    const int64_t oex = PX ? (int64_t)(wrapi(i + 1, g.nx) - i) : 1;
    const int64_t oey = PY ? (int64_t)(wrapi(j + 1, g.ny) - j) * g.sy : g.sy;
    const int64_t oez = PZ ? (int64_t)(wrapi(k + 1, g.nz) - k) * g.sz : g.sz;
    const uint32_t bex = (uint32_t)oex * (uint32_t)ES, bey = (uint32_t)oey * (uint32_t)ES;
    const uint32_t bez = (uint32_t)oez * (uint32_t)ES;
    const bool ex = PX || i < g.nx - 1;   // loops stop at n-1 on a wall axis
    const bool ey = PY || j < g.ny - 1;
    const bool ez = PZ || k < g.nz - 1;
    T du = T(0), dv = T(0), dw = T(0);
    if (vf && lin_vf) {
        // the six terms in the order below, with the linear branch
        const int o = a.upw_vf;
        if (ey) dv = vf_lin<T, PX, -1>(dv, a.u, cx, a.oz, bx, 1, i, g.nx, oey, o);
        if (ex) du = vf_lin<T, PY, +1>(du, a.v, cy, a.oz, by, g.sy, j, g.ny, oex, o);
        if (ez) dw = vf_lin<T, PY, -1>(dw, a.v, cy, a.ox, by, g.sy, j, g.ny, oez, o);
        if (ey) dv = vf_lin<T, PZ, +1>(dv, a.w, cz, a.ox, bz, g.sz, k, g.nz, oey, o);
        if (ex) du = vf_lin<T, PZ, -1>(du, a.w, cz, a.oy, bz, g.sz, k, g.nz, oex, o);
        if (ez) dw = vf_lin<T, PX, +1>(dw, a.u, cx, a.oy, bx, 1, i, g.nx, oez, o);
    } else if (vf) {
    // direc(U_i, vor_k, du_j)
    if (ey) {
        if constexpr (PX != 0) dv = dv - vf_flux_per<T>(a.u, cx, a.oz, cs, Lx, bey);
        else dv = dv - vf_flux<T, 0>(a.u, cx, a.oz, bx, 1, i, g.nx, oey);
    }
    STAGE_SEQ_V(by, dv, 0);
    STAGE_SEQ_V(cs, dv, 0);
    // flip (U_j, vor_k, du_i)
    if (ex) {
        if constexpr (PY != 0) du = du + vf_flux_per<T>(a.v, cy, a.oz, cs, Ly, bex);
        else du = du + vf_flux<T, 0>(a.v, cy, a.oz, by, g.sy, j, g.ny, oex);
    }
    STAGE_SEQ_V(by, du, 1);
    STAGE_SEQ_V(cs, du, 1);
    // direc(U_j, vor_i, du_k)
    if (ez) {
        if constexpr (PY != 0) dw = dw - vf_flux_per<T>(a.v, cy, a.ox, cs, Ly, bez);
        else dw = dw - vf_flux<T, 0>(a.v, cy, a.ox, by, g.sy, j, g.ny, oez);
    }
    STAGE_SEQ_V(bz, dw, 0);
    STAGE_SEQ_V(cs, dw, 0);
    // flip (U_k, vor_i, du_j)
    if (ey) {
        if constexpr (PZ != 0) dv = dv + vf_flux_per<T>(a.w, cz, a.ox, cs, Lz, bey);
        else dv = dv + vf_flux<T, 0>(a.w, cz, a.ox, bz, g.sz, k, g.nz, oey);
    }
    STAGE_SEQ_V(bz, dv, 1);
    STAGE_SEQ_V(cs, dv, 1);
    // direc(U_k, vor_j, du_i)
    if (ex) {
        if constexpr (PZ != 0) du = du - vf_flux_per<T>(a.w, cz, a.oy, cs, Lz, bex);
        else du = du - vf_flux<T, 0>(a.w, cz, a.oy, bz, g.sz, k, g.nz, oex);
    }
    STAGE_SEQ_V(bx, du, 0);
    STAGE_SEQ_V(cs, du, 0);
    // flip (U_i, vor_j, du_k)
    if (ez) {
        if constexpr (PX != 0) dw = dw + vf_flux_per<T>(a.u, cx, a.oy, cs, Lx, bez);
        else dw = dw + vf_flux<T, 0>(a.u, cx, a.oy, bx, 1, i, g.nx, oez);
    }
    }

    // bernoulli (core/bernoulli.py:14-23, core/fortran_bernoulli.f90:2-58)
    const T kec = ldb(a.ke, cb);
    if (ex) du = du - (ldb(a.ke, cb + bex) - kec);
    if (ey) dv = dv - (ldb(a.ke, cb + bey) - kec);
    if (ez) {
        const T cff = T(0.5 * a.dz);
        if (with_b) dw = dw - (ldb(a.ke, cb + bez) - kec) + cff * (ldb(a.b, cb + bez) + ldb(a.b, cb));
        else dw = dw - (ldb(a.ke, cb + bez) - kec);
    }
    // viscosity at the last stage (core/model_les.py:140-141)
    if (a.visc) viscosity_cell<T, PX, PY, PZ>(g, a, cb, i, j, k, du, dv, dw);

    if (MODE == 0 || MODE == 1 || MODE == 2) {   // read at a later RK stage
        stb_stream(a.tb, cb, db); stb_stream(a.tu, cb, du);
        stb_stream(a.tv, cb, dv); stb_stream(a.tw, cb, dw);
    }
    const T cc = T(a.c);
    if (MODE == 1) {            // s += dt*ds0
        stb(a.bn, cb, ldb(a.b, cb) + cc * db);
        stb(a.un, cb, ldb(a.u, cb) + cc * du);
        stb(a.vn, cb, ldb(a.v, cb) + cc * dv);
        stb(a.wn, cb, ldb(a.w, cb) + cc * dw);
    } else if (MODE == 2) {     // s += (dt/4.)*(ds1-3*ds0)
        stb(a.bn, cb, ldb(a.b, cb) + cc * (db - T(3) * ldb(a.d0b, cb)));
        stb(a.un, cb, ldb(a.u, cb) + cc * (du - T(3) * ldb(a.d0u, cb)));
        stb(a.vn, cb, ldb(a.v, cb) + cc * (dv - T(3) * ldb(a.d0v, cb)));
        stb(a.wn, cb, ldb(a.w, cb) + cc * (dw - T(3) * ldb(a.d0w, cb)));
    } else if (MODE == 3) {     // s += (dt/12.)*(8*ds2-ds0-ds1)
        stb(a.bn, cb, ldb(a.b, cb) + cc * ((T(8) * db - ldb(a.d0b, cb)) - ldb(a.d1b, cb)));
        stb(a.un, cb, ldb(a.u, cb) + cc * ((T(8) * du - ldb(a.d0u, cb)) - ldb(a.d1u, cb)));
        stb(a.vn, cb, ldb(a.v, cb) + cc * ((T(8) * dv - ldb(a.d0v, cb)) - ldb(a.d1v, cb)));
        stb(a.wn, cb, ldb(a.w, cb) + cc * ((T(8) * dw - ldb(a.d0w, cb)) - ldb(a.d1w, cb)));
    } else if (MODE == 4) {     // LFAM3 predictor (core/timescheme.py:144-162)
        // s = sb + (2.*dt)*ds; s = (1./12.)*(5.*s+8.*sn-sb); sb = sn
        const T c2 = T(a.c2), c12 = T(1. / 12.);
        T sb, t1;
        sb = ldb(a.sbb, cb); t1 = sb + c2 * db; stb(a.bn, cb, c12 * (T(5) * t1 + T(8) * ldb(a.b, cb) - sb)); stb(a.sbb, cb, ldb(a.b, cb));
        sb = ldb(a.sbu, cb); t1 = sb + c2 * du; stb(a.un, cb, c12 * (T(5) * t1 + T(8) * ldb(a.u, cb) - sb)); stb(a.sbu, cb, ldb(a.u, cb));
        sb = ldb(a.sbv, cb); t1 = sb + c2 * dv; stb(a.vn, cb, c12 * (T(5) * t1 + T(8) * ldb(a.v, cb) - sb)); stb(a.sbv, cb, ldb(a.v, cb));
        sb = ldb(a.sbw, cb); t1 = sb + c2 * dw; stb(a.wn, cb, c12 * (T(5) * t1 + T(8) * ldb(a.w, cb) - sb)); stb(a.sbw, cb, ldb(a.w, cb));
    } else if (MODE == 5) {     // LFAM3 corrector: s = sn + dt*ds (core/timescheme.py:170-175)
        stb(a.bn, cb, ldb(a.bn, cb) + cc * db);
        stb(a.un, cb, ldb(a.un, cb) + cc * du);
        stb(a.vn, cb, ldb(a.vn, cb) + cc * dv);
        stb(a.wn, cb, ldb(a.wn, cb) + cc * dw);
    } else if (MODE == 6) {     // first LFAM3 step: sn = s; sb = s; s += dt*ds (:352-364)
        stb(a.sbb, cb, ldb(a.b, cb)); stb(a.sbu, cb, ldb(a.u, cb)); stb(a.sbv, cb, ldb(a.v, cb)); stb(a.sbw, cb, ldb(a.w, cb));
        stb(a.bn, cb, ldb(a.b, cb) + cc * db);
        stb(a.un, cb, ldb(a.u, cb) + cc * du);
        stb(a.vn, cb, ldb(a.v, cb) + cc * dv);
        stb(a.wn, cb, ldb(a.w, cb) + cc * dw);
    }
}

template <typename T, int MODE, int PX, int PY, int PZ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NYL_STAGE_WAVES)))
void k_stage(Geo g, StageArgs<T> a) {
#if NYL_STAGE_YSLAB
    const Blk3 bk = xcd_block_yslab();
#else
    const Blk3 bk = xcd_block();
#endif
    const int i = bk.x * NYL_BX + threadIdx.x;
    const int j = bk.y * NYL_BY + threadIdx.y;
    const int k = bk.z;
    if (i >= g.nx || j >= g.ny) return;
    stage_cell<T, MODE, PX, PY, PZ>(g, a, i, j, k);
}

template <typename T, int PX, int PY, int PZ>
static void launch_stage_p(const Geo &g, const StageArgs<T> &a, int mode, hipStream_t s) {
    dim3 blk(NYL_BX, NYL_BY, 1), grd = cell_grid(g);
    switch (mode) {
    case 0: NYL_LAUNCH((k_stage<T, 0, PX, PY, PZ>), grd, blk, 0, s, g, a); break;
    case 1: NYL_LAUNCH((k_stage<T, 1, PX, PY, PZ>), grd, blk, 0, s, g, a); break;
    case 2: NYL_LAUNCH((k_stage<T, 2, PX, PY, PZ>), grd, blk, 0, s, g, a); break;
    case 3: NYL_LAUNCH((k_stage<T, 3, PX, PY, PZ>), grd, blk, 0, s, g, a); break;
    case 4: NYL_LAUNCH((k_stage<T, 4, PX, PY, PZ>), grd, blk, 0, s, g, a); break;
    case 5: NYL_LAUNCH((k_stage<T, 5, PX, PY, PZ>), grd, blk, 0, s, g, a); break;
    case 6: NYL_LAUNCH((k_stage<T, 6, PX, PY, PZ>), grd, blk, 0, s, g, a); break;
    default: break;
    }
}


// ---------------------------------------------------------------------------
// The stage kernel as a 2.5-D march (k_stage25).  A block owns a tile of
// S25_TX x S25_TY columns and walks a range of z planes.  Per plane:
//  * the fields that the in-plane (x, y) stencils read -- b, the three
//    vorticity components, u, v, w and ke of plane k, u and v of plane
//    k+1 -- are staged in LDS, each element loaded once, with the halo each
//    stencil reaches (3 cells for the WENO stencils, 1 for the averages);
//  * the x and y tracer face fluxes of the tile are evaluated once each into
//    LDS (the per-cell kernel evaluates every face twice, once for each
//    cell beside it);
//  * the z stencils (the tracer and vortex-force terms 4 and 5) read
//    per-thread register windows of the column (b over k-3..k+3, vor_i and
//    vor_j over k-3..k+2), the z face flux of the tracer and the z-averaged
//    velocities of terms 4 and 5 are carried to the next plane.
// Per cell that is about 9.5 reconstructions instead of 12, and no stencil
// value is gathered from L1/L2.  Every value is the per-cell kernel's
// expression in the same order (stage_cell, vf_flux, tracer_axis), so the
// results are the same bits (checked on the CPU against k_stage).
#define S25_TX 64
#define S25_TY 4
#define S25_NT (S25_TX * S25_TY)
#ifndef NYL_STAGE25
#define NYL_STAGE25 0
#endif
// planes per block (the grid needs many blocks to fill the chip)
#ifndef NYL_S25_BLOCKS
#define NYL_S25_BLOCKS 4096
#endif

constexpr int S25_BW = S25_TX + 6, S25_BH = S25_TY + 6;
constexpr int S25_W1 = S25_TX + 1, S25_H1 = S25_TY + 1;

// Sequencing of the per-cell terms of k_stage25: the LDS column index of
// the next term passes through an empty asm that takes the previous term's
// result, so the next term's loads are not hoisted above it and each term
// holds only its own operands (no instruction is emitted).
#if NYL_STAGE_SEQ
#define S25_SEQ(idx, dep) asm volatile("" : "+v"(idx) : "v"(dep))
#else
#define S25_SEQ(idx, dep) ((void)0)
#endif

template <typename T>
struct S25Lds {
    T b[S25_BH * S25_BW];     // b(k):   x in [-3, TX+3), y in [-3, TY+3)
    T oz[S25_BH * S25_BW];    // vor_k:  same box
    T ox[S25_BH * S25_TX];    // vor_i:  x in [0, TX),    y in [-3, TY+3)
    T oy[S25_TY * S25_BW];    // vor_j:  x in [-3, TX+3), y in [0, TY)
    T u0[S25_H1 * S25_W1];    // u(k):   x in [-1, TX),   y in [0, TY+1)
    T u1[S25_TY * S25_W1];    // u(k+1): x in [-1, TX),   y in [0, TY)
    T v0[S25_H1 * S25_W1];    // v(k):   x in [0, TX+1),  y in [-1, TY)
    T v1[S25_H1 * S25_TX];    // v(k+1): x in [0, TX),    y in [-1, TY)
    T w0[S25_H1 * S25_W1];    // w(k):   x in [0, TX+1),  y in [0, TY+1)
    T ke[S25_H1 * S25_W1];    // ke(k):  x in [0, TX+1),  y in [0, TY+1)
    T fx[S25_TY * S25_W1];    // x-face tracer fluxes, faces -1 .. TX-1 of each row
    T fy[S25_H1 * S25_TX];    // y-face tracer fluxes, faces -1 .. TY-1 of each column
};

// Columns of the LDS boxes for one lane: x = i0 + X0 + lane (lo) and
// x = i0 + X0 + 64 + lane (hi, the lanes past 64 of a wide box), wrapped on
// a periodic axis, -1 off the grid; X0 = -3, -1, 0.  Plane-invariant.
struct S25Cols {
    int lo[3], hi[3];
};
template <int X0>
__device__ constexpr int s25_xi() { return X0 == -3 ? 0 : (X0 == -1 ? 1 : 2); }

// One LDS box: field f of plane offset pl (elements) over x in
// [i0+X0, i0+X0+W), y in [j0+Y0, j0+Y0+H), row-major: wave w fills rows w,
// w+TY, ... (the row's offset is wave-uniform), each lane one or two
// columns.  Periodic positions wrap; positions off a wall axis (or past a
// partial tile) store 0: the closures never read them.
template <typename T, int PY, int X0, int Y0, int W, int H>
__device__ __forceinline__ void s25_box(T *__restrict__ dst, const T *__restrict__ f,
                                        const Geo &g, int j0, int64_t pl, int lane, int wave,
                                        const S25Cols &cc) {
    constexpr uint32_t ES = sizeof(T);
    constexpr int XI = s25_xi<X0>();
    const int lo = cc.lo[XI], hi = cc.hi[XI];
    for (int r = wave; r < H; r += S25_TY) {
        int gj = j0 + Y0 + r;
        if (PY) gj = gj < 0 ? gj + g.ny : (gj >= g.ny ? gj - g.ny : gj);
        const bool rok = gj >= 0 && gj < g.ny;
        const T *__restrict__ row = f + (rok ? pl + (int64_t)gj * g.sy : 0);
        dst[r * W + lane] = (rok && lo >= 0) ? ldb(row, (uint32_t)lo * ES) : T(0);
        if (W > S25_TX && lane < W - S25_TX)
            dst[r * W + S25_TX + lane] = (rok && hi >= 0) ? ldb(row, (uint32_t)hi * ES) : T(0);
    }
}

// A window of n values of the thread's column along z, held in scalars (a
// register array indexed by a selected position would be placed in scratch
// memory); at(d) with a constant d.
template <typename T>
struct S25Win7 {
    T v0, v1, v2, v3, v4, v5, v6;
    __device__ __forceinline__ T at(int d) const {
        switch (d) {
        case 0: return v0; case 1: return v1; case 2: return v2; case 3: return v3;
        case 4: return v4; case 5: return v5; default: return v6;
        }
    }
    __device__ __forceinline__ void shift() { v0 = v1; v1 = v2; v2 = v3; v3 = v4; v4 = v5; v5 = v6; }
};

// vf_flux on staged values: UU(dm) is the averaged contravariant velocity at
// line position md + dm (dm = -1, 0), om(m) the vorticity at line position m
template <typename T, int PER, typename UF, typename OF>
__device__ __forceinline__ T s25_vf(int md, int nd, const UF &UU, const OF &om) {
    T uu_prev;
    if (!PER && md == 0) uu_prev = T(0);   // UU_0 = 0. at the line start
    else uu_prev = UU(-1);
    const T u1d = T(0.5) * (uu_prev + UU(0));
    auto q = [&](int m) -> T {             // q(1) = 0., q(k) = vort(k-1)
        if (!PER && m == 0) return T(0);
        return om(m - 1);
    };
    return face_flux<T, PER>(md, nd, u1d, q);
}

template <typename T, int MODE, int PX, int PY, int PZ>
__global__ __launch_bounds__(S25_NT) void k_stage25(Geo g, StageArgs<T> a, int kz) {
    __shared__ S25Lds<T> L;
    constexpr uint32_t ES = sizeof(T);
    const Blk3 bk = xcd_block();
    const int tid = threadIdx.x;
    const int tx = tid % S25_TX, ty = tid / S25_TX;
    const int i0 = bk.x * S25_TX, j0 = bk.y * S25_TY;
    const int i = i0 + tx, j = j0 + ty;
    const int nz = g.nz;
    const int k0 = bk.z * kz, k1 = k0 + kz < nz ? k0 + kz : nz;
    const bool live = i < g.nx && j < g.ny;
    // the thread's column (threads past a partial tile use column 0, and
    // compute nothing from it)
    const int ic = live ? i : 0, jc = live ? j : 0;
    const int64_t col = (int64_t)jc * g.sy + ic;
    const T cx = T(a.idx2), cy = T(a.idy2), cz = T(a.idz2);
    auto zw = [&](int kk) -> int { return PZ ? (kk < 0 ? kk + nz : (kk >= nz ? kk - nz : kk)) : kk; };
    auto zok = [&](int kk) -> bool { return PZ || (kk >= 0 && kk < nz); };
    // field f at the thread's column (offset dc within the plane), plane kk
    auto colv = [&](const T *__restrict__ f, int kk, int64_t dc) -> T {
        return zok(kk) ? ldb(f, (uint32_t)((int64_t)zw(kk) * g.sz + col + dc) * ES) : T(0);
    };
    // offsets of the x+1 and y+1 neighbours of the column (0 when absent)
    const int64_t dxp = (PX || ic + 1 < g.nx) ? (int64_t)((PX && ic + 1 >= g.nx) ? 1 - g.nx : 1)
                                              : 0;
    const int64_t dyp = (PY || jc + 1 < g.ny) ? ((PY && jc + 1 >= g.ny) ? (1 - g.ny) * g.sy : g.sy)
                                              : 0;
    // register windows along z: b over k-3 .. k+3, vor_i and vor_j over
    // k-3 .. k+2 (the last slot of those unused)
    S25Win7<T> bw, xw, yw;
    bw.v0 = colv(a.b, k0 - 3, 0); bw.v1 = colv(a.b, k0 - 2, 0); bw.v2 = colv(a.b, k0 - 1, 0);
    bw.v3 = colv(a.b, k0, 0); bw.v4 = colv(a.b, k0 + 1, 0); bw.v5 = colv(a.b, k0 + 2, 0);
    xw.v0 = colv(a.ox, k0 - 3, 0); xw.v1 = colv(a.ox, k0 - 2, 0); xw.v2 = colv(a.ox, k0 - 1, 0);
    xw.v3 = colv(a.ox, k0, 0); xw.v4 = colv(a.ox, k0 + 1, 0);
    yw.v0 = colv(a.oy, k0 - 3, 0); yw.v1 = colv(a.oy, k0 - 2, 0); yw.v2 = colv(a.oy, k0 - 1, 0);
    yw.v3 = colv(a.oy, k0, 0); yw.v4 = colv(a.oy, k0 + 1, 0);
    bw.v6 = xw.v5 = xw.v6 = yw.v5 = yw.v6 = T(0);
    // carried from the plane below: the tracer flux through face k-1 and the
    // z-averaged U_k of vortex-force terms 4 (averaged in y) and 5 (in x)
    T fzl = T(0), uu4p = T(0), uu5p = T(0);
    if (PZ || k0 > 0) {
        const T wm = colv(a.w, k0 - 1, 0);
        fzl = face_flux_rel<T, PZ>(zw(k0 - 1), nz, wm * cz,
                                   [&](int d) -> T { return bw.at(d + 2); });
        uu4p = T(0.5) * (wm * cz + colv(a.w, k0 - 1, dyp) * cz);
        uu5p = T(0.5) * (wm * cz + colv(a.w, k0 - 1, dxp) * cz);
    }
    // the lane's box columns
    S25Cols cc;
    {
        const int x0s[3] = {-3, -1, 0};
#pragma unroll
        for (int q = 0; q < 3; q++) {
            int lo = i0 + x0s[q] + tx, hi = lo + S25_TX;
            if (PX) {
                lo = lo < 0 ? lo + g.nx : (lo >= g.nx ? lo - g.nx : lo);
                hi = hi >= g.nx ? hi - g.nx : hi;
            }
            cc.lo[q] = (lo >= 0 && lo < g.nx) ? lo : -1;
            cc.hi[q] = (hi >= 0 && hi < g.nx) ? hi : -1;
        }
    }
    for (int k = k0; k < k1; k++) {
        const int64_t pl = (int64_t)k * g.sz;
        const bool has_kp = zok(k + 1);
        const int64_t plp = (int64_t)zw(k + 1) * g.sz;
        bw.v6 = colv(a.b, k + 3, 0);
        xw.v5 = colv(a.ox, k + 2, 0);
        yw.v5 = colv(a.oy, k + 2, 0);
        const T ke_up = has_kp ? ldb(a.ke, (uint32_t)(plp + col) * ES) : T(0);
        s25_box<T, PY, -3, -3, S25_BW, S25_BH>(L.b, a.b, g, j0, pl, tx, ty, cc);
        s25_box<T, PY, -3, -3, S25_BW, S25_BH>(L.oz, a.oz, g, j0, pl, tx, ty, cc);
        s25_box<T, PY, 0, -3, S25_TX, S25_BH>(L.ox, a.ox, g, j0, pl, tx, ty, cc);
        s25_box<T, PY, -3, 0, S25_BW, S25_TY>(L.oy, a.oy, g, j0, pl, tx, ty, cc);
        s25_box<T, PY, -1, 0, S25_W1, S25_H1>(L.u0, a.u, g, j0, pl, tx, ty, cc);
        s25_box<T, PY, 0, -1, S25_W1, S25_H1>(L.v0, a.v, g, j0, pl, tx, ty, cc);
        s25_box<T, PY, 0, 0, S25_W1, S25_H1>(L.w0, a.w, g, j0, pl, tx, ty, cc);
        s25_box<T, PY, 0, 0, S25_W1, S25_H1>(L.ke, a.ke, g, j0, pl, tx, ty, cc);
        if (has_kp) {
            s25_box<T, PY, -1, 0, S25_W1, S25_TY>(L.u1, a.u, g, j0, plp, tx, ty, cc);
            s25_box<T, PY, 0, -1, S25_TX, S25_H1>(L.v1, a.v, g, j0, plp, tx, ty, cc);
        }
        __syncthreads();
        // tracer face fluxes of the tile: x faces i0-1 .. i0+TX-1 of each
        // row, y faces j0-1 .. j0+TY-1 of each column.  Wave w: the x faces
        // i0-1 .. i0+TX-2 of row w and the y faces of row j0+w-1; then wave 0
        // the y faces of row j0+TY-1 and wave 1 the last x face of each row.
        {
            auto xface = [&](int r, int fr) -> T {
                const int f = i0 + fr;
                if (!((PX || f >= 0) && f < g.nx && j0 + r < g.ny)) return T(0);
                auto q = [&](int m) -> T { return L.b[(r + 3) * S25_BW + (m - i0 + 3)]; };
                return face_flux<T, PX>(f, g.nx, L.u0[r * S25_W1 + fr + 1] * cx, q);
            };
            auto yface = [&](int c, int fr) -> T {
                const int f = j0 + fr;
                if (!((PY || f >= 0) && f < g.ny && i0 + c < g.nx)) return T(0);
                auto q = [&](int m) -> T { return L.b[(m - j0 + 3) * S25_BW + (c + 3)]; };
                return face_flux<T, PY>(f, g.ny, L.v0[(fr + 1) * S25_W1 + c] * cy, q);
            };
            const T fxv = xface(ty, tx - 1);
            L.fx[ty * S25_W1 + tx] = fxv;
            int cq = tx;
            S25_SEQ(cq, fxv);
            const T fyv = yface(cq, ty - 1);
            L.fy[ty * S25_TX + tx] = fyv;
            S25_SEQ(cq, fyv);
            if (ty == 0) L.fy[S25_TY * S25_TX + tx] = yface(cq, S25_TY - 1);
            else if (ty == 1 && tx < S25_TY) L.fx[tx * S25_W1 + S25_TX] = xface(cq, S25_TX - 1);
        }
        __syncthreads();
        if (live) {
            const uint32_t cb = (uint32_t)(pl + col) * ES;
            // tracer b: rhstrac, directions i, j, k (core/tracer.py:63-72)
            T db = T(0);
            {
                const T fl = L.fx[ty * S25_W1 + tx], fr = L.fx[ty * S25_W1 + tx + 1];
                if (!PX && i == 0) db = db - fr;
                else db = (db + fl) - fr;
            }
            {
                const T fl = L.fy[ty * S25_TX + tx], fr = L.fy[(ty + 1) * S25_TX + tx];
                if (!PY && j == 0) db = db - fr;
                else db = (db + fl) - fr;
            }
            const T wc = L.w0[ty * S25_W1 + tx];
            {
                const T fr = face_flux_rel<T, PZ>(k, nz, wc * cz,
                                                  [&](int d) -> T { return bw.at(d + 3); });
                if (!PZ && k == 0) db = db - fr;
                else db = (db + fzl) - fr;
                fzl = fr;
            }
            // vortex force, triplets 'ikj', 'jik', 'kji' (core/vortex_force.py:69-81)
            const bool ex = PX || i < g.nx - 1;
            const bool ey = PY || j < g.ny - 1;
            const bool ez = PZ || k < nz - 1;
            int lx = tx;             // the LDS column, sequenced between the terms
            S25_SEQ(lx, db);
            const T uu4 = T(0.5) * (wc * cz + L.w0[(ty + 1) * S25_W1 + lx] * cz);
            const T uu5 = T(0.5) * (wc * cz + L.w0[ty * S25_W1 + lx + 1] * cz);
            T du = T(0), dv = T(0), dw = T(0);
            // direc(U_i, vor_k, du_j): along x, U_i averaged in y
            if (ey)
                dv = dv - s25_vf<T, PX>(i, g.nx,
                    [&](int dm) -> T {
                        return T(0.5) * (L.u0[ty * S25_W1 + lx + 1 + dm] * cx +
                                         L.u0[(ty + 1) * S25_W1 + lx + 1 + dm] * cx);
                    },
                    [&](int m) -> T { return L.oz[(ty + 3) * S25_BW + (m - i + lx + 3)]; });
            S25_SEQ(lx, dv);
            // flip(U_j, vor_k, du_i): along y, U_j averaged in x
            if (ex)
                du = du + s25_vf<T, PY>(j, g.ny,
                    [&](int dm) -> T {
                        return T(0.5) * (L.v0[(ty + 1 + dm) * S25_W1 + lx] * cy +
                                         L.v0[(ty + 1 + dm) * S25_W1 + lx + 1] * cy);
                    },
                    [&](int m) -> T { return L.oz[(m - j + ty + 3) * S25_BW + (lx + 3)]; });
            S25_SEQ(lx, du);
            // direc(U_j, vor_i, du_k): along y, U_j averaged in z
            if (ez)
                dw = dw - s25_vf<T, PY>(j, g.ny,
                    [&](int dm) -> T {
                        return T(0.5) * (L.v0[(ty + 1 + dm) * S25_W1 + lx] * cy +
                                         L.v1[(ty + 1 + dm) * S25_TX + lx] * cy);
                    },
                    [&](int m) -> T { return L.ox[(m - j + ty + 3) * S25_TX + lx]; });
            S25_SEQ(lx, dw);
            // flip(U_k, vor_i, du_j): along z, U_k averaged in y
            if (ey) {
                const T uu_prev = (!PZ && k == 0) ? T(0) : uu4p;
                const T u1d = T(0.5) * (uu_prev + uu4);
                dv = dv + face_flux_rel<T, PZ>(k, nz, u1d, [&](int d) -> T {
                    if (!PZ && k + d == 0) return T(0);
                    return xw.at(d + 2);
                });
            }
            // direc(U_k, vor_j, du_i): along z, U_k averaged in x
            if (ex) {
                const T uu_prev = (!PZ && k == 0) ? T(0) : uu5p;
                const T u1d = T(0.5) * (uu_prev + uu5);
                du = du - face_flux_rel<T, PZ>(k, nz, u1d, [&](int d) -> T {
                    if (!PZ && k + d == 0) return T(0);
                    return yw.at(d + 2);
                });
            }
            uu4p = uu4;
            uu5p = uu5;
            S25_SEQ(lx, du);
            // flip(U_i, vor_j, du_k): along x, U_i averaged in z
            if (ez)
                dw = dw + s25_vf<T, PX>(i, g.nx,
                    [&](int dm) -> T {
                        return T(0.5) * (L.u0[ty * S25_W1 + lx + 1 + dm] * cx +
                                         L.u1[ty * S25_W1 + lx + 1 + dm] * cx);
                    },
                    [&](int m) -> T { return L.oy[ty * S25_BW + (m - i + lx + 3)]; });

            // bernoulli (core/bernoulli.py:14-23, core/fortran_bernoulli.f90:2-58)
            const T kec = L.ke[ty * S25_W1 + tx];
            if (ex) du = du - (L.ke[ty * S25_W1 + tx + 1] - kec);
            if (ey) dv = dv - (L.ke[(ty + 1) * S25_W1 + tx] - kec);
            if (ez) {
                const T cff = T(0.5 * a.dz);
                dw = dw - (ke_up - kec) + cff * (bw.v4 + bw.v3);
            }
            // viscosity at the last stage (core/model_les.py:140-141)
            if (a.visc) viscosity_cell<T, PX, PY, PZ>(g, a, cb, i, j, k, du, dv, dw);

            const T b0 = bw.v3, u0 = L.u0[ty * S25_W1 + tx + 1];
            const T v0 = L.v0[(ty + 1) * S25_W1 + tx], w0 = wc;
            if (MODE == 0 || MODE == 1 || MODE == 2) {
                stb_stream(a.tb, cb, db); stb_stream(a.tu, cb, du); stb_stream(a.tv, cb, dv); stb_stream(a.tw, cb, dw);
            }
            const T cc = T(a.c);
            if (MODE == 1) {            // s += dt*ds0
                stb(a.bn, cb, b0 + cc * db);
                stb(a.un, cb, u0 + cc * du);
                stb(a.vn, cb, v0 + cc * dv);
                stb(a.wn, cb, w0 + cc * dw);
            } else if (MODE == 2) {     // s += (dt/4.)*(ds1-3*ds0)
                stb(a.bn, cb, b0 + cc * (db - T(3) * ldb(a.d0b, cb)));
                stb(a.un, cb, u0 + cc * (du - T(3) * ldb(a.d0u, cb)));
                stb(a.vn, cb, v0 + cc * (dv - T(3) * ldb(a.d0v, cb)));
                stb(a.wn, cb, w0 + cc * (dw - T(3) * ldb(a.d0w, cb)));
            } else if (MODE == 3) {     // s += (dt/12.)*(8*ds2-ds0-ds1)
                stb(a.bn, cb, b0 + cc * ((T(8) * db - ldb(a.d0b, cb)) - ldb(a.d1b, cb)));
                stb(a.un, cb, u0 + cc * ((T(8) * du - ldb(a.d0u, cb)) - ldb(a.d1u, cb)));
                stb(a.vn, cb, v0 + cc * ((T(8) * dv - ldb(a.d0v, cb)) - ldb(a.d1v, cb)));
                stb(a.wn, cb, w0 + cc * ((T(8) * dw - ldb(a.d0w, cb)) - ldb(a.d1w, cb)));
            } else if (MODE == 4) {     // LFAM3 predictor (core/timescheme.py:144-162)
                const T c2 = T(a.c2), c12 = T(1. / 12.);
                T sb, t1;
                sb = ldb(a.sbb, cb); t1 = sb + c2 * db; stb(a.bn, cb, c12 * (T(5) * t1 + T(8) * b0 - sb)); stb(a.sbb, cb, b0);
                sb = ldb(a.sbu, cb); t1 = sb + c2 * du; stb(a.un, cb, c12 * (T(5) * t1 + T(8) * u0 - sb)); stb(a.sbu, cb, u0);
                sb = ldb(a.sbv, cb); t1 = sb + c2 * dv; stb(a.vn, cb, c12 * (T(5) * t1 + T(8) * v0 - sb)); stb(a.sbv, cb, v0);
                sb = ldb(a.sbw, cb); t1 = sb + c2 * dw; stb(a.wn, cb, c12 * (T(5) * t1 + T(8) * w0 - sb)); stb(a.sbw, cb, w0);
            } else if (MODE == 5) {     // LFAM3 corrector: s = sn + dt*ds
                stb(a.bn, cb, ldb(a.bn, cb) + cc * db);
                stb(a.un, cb, ldb(a.un, cb) + cc * du);
                stb(a.vn, cb, ldb(a.vn, cb) + cc * dv);
                stb(a.wn, cb, ldb(a.wn, cb) + cc * dw);
            } else if (MODE == 6) {     // first LFAM3 step: sn = s; sb = s; s += dt*ds
                stb(a.sbb, cb, b0); stb(a.sbu, cb, u0); stb(a.sbv, cb, v0); stb(a.sbw, cb, w0);
                stb(a.bn, cb, b0 + cc * db);
                stb(a.un, cb, u0 + cc * du);
                stb(a.vn, cb, v0 + cc * dv);
                stb(a.wn, cb, w0 + cc * dw);
            }
        }
        bw.shift();
        xw.shift();
        yw.shift();
        __syncthreads();
    }
}

static inline int s25_kz(const Geo &g) {
    const int64_t tiles = (int64_t)((g.nx + S25_TX - 1) / S25_TX) * ((g.ny + S25_TY - 1) / S25_TY);
    int64_t chunks = (NYL_S25_BLOCKS + tiles - 1) / tiles;
    if (chunks > g.nz) chunks = g.nz;
    if (chunks < 1) chunks = 1;
    int kz = (int)((g.nz + chunks - 1) / chunks);
    return kz < 1 ? 1 : kz;
}

// the 2.5-D form needs a full-width tile row (narrower grids keep the
// per-cell kernel)
static inline bool s25_fits(const Geo &g) { return NYL_STAGE25 && g.nx >= S25_TX; }

template <typename T, int PX, int PY, int PZ>
static void launch_stage25_p(const Geo &g, const StageArgs<T> &a, int mode, hipStream_t s) {
    const int kz = s25_kz(g);
    dim3 blk(S25_NT), grd((g.nx + S25_TX - 1) / S25_TX, (g.ny + S25_TY - 1) / S25_TY,
                          (g.nz + kz - 1) / kz);
    switch (mode) {
    case 0: NYL_LAUNCH((k_stage25<T, 0, PX, PY, PZ>), grd, blk, 0, s, g, a, kz); break;
    case 1: NYL_LAUNCH((k_stage25<T, 1, PX, PY, PZ>), grd, blk, 0, s, g, a, kz); break;
    case 2: NYL_LAUNCH((k_stage25<T, 2, PX, PY, PZ>), grd, blk, 0, s, g, a, kz); break;
    case 3: NYL_LAUNCH((k_stage25<T, 3, PX, PY, PZ>), grd, blk, 0, s, g, a, kz); break;
    case 4: NYL_LAUNCH((k_stage25<T, 4, PX, PY, PZ>), grd, blk, 0, s, g, a, kz); break;
    case 5: NYL_LAUNCH((k_stage25<T, 5, PX, PY, PZ>), grd, blk, 0, s, g, a, kz); break;
    case 6: NYL_LAUNCH((k_stage25<T, 6, PX, PY, PZ>), grd, blk, 0, s, g, a, kz); break;
    default: break;
    }
}

template <typename T>
hipError_t launch_stage(const Geo &g, const StageArgs<T> &a, int mode, hipStream_t s) {
    if (mode < 0 || mode > 6) return hipErrorInvalidValue;
    const bool t25 = s25_fits(g) && a.flags == 0;   // the 2.5-D form runs the LES only
    switch (g.px + 2 * g.py + 4 * g.pz) {
#define NYL_STAGE_CASE(c, X, Y, Z)                                                   \
    case c:                                                                          \
        if (t25) launch_stage25_p<T, X, Y, Z>(g, a, mode, s);                        \
        else launch_stage_p<T, X, Y, Z>(g, a, mode, s);                              \
        break;
    NYL_STAGE_CASE(0, 0, 0, 0)
    NYL_STAGE_CASE(1, 1, 0, 0)
    NYL_STAGE_CASE(2, 0, 1, 0)
    NYL_STAGE_CASE(3, 1, 1, 0)
    NYL_STAGE_CASE(4, 0, 0, 1)
    NYL_STAGE_CASE(5, 1, 0, 1)
    NYL_STAGE_CASE(6, 0, 1, 1)
    default:
        if (t25) launch_stage25_p<T, 1, 1, 1>(g, a, mode, s);
        else launch_stage_p<T, 1, 1, 1>(g, a, mode, s);
        break;
#undef NYL_STAGE_CASE
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Passive tracers: the b part of stage_cell for another scalar, with the
// stage update of the mode (same expressions as b's).
template <typename T, int MODE, int PX, int PY, int PZ>
__device__ __forceinline__ void tracer_cell(const Geo &g, const TracerArgs<T> &a, int i, int j,
                                            int k) {
    const int64_t c = (int64_t)k * g.sz + (int64_t)j * g.sy + i;
    constexpr int ES = (int)sizeof(T);
    const uint32_t cb = (uint32_t)c * (uint32_t)ES;
    const T cx = T(a.idx2), cy = T(a.idy2), cz = T(a.idz2);
    T dq = T(0);
    const int64_t bx = c - i, by = c - (int64_t)j * g.sy, bz = c - (int64_t)k * g.sz;
    if (MODE == 0 && a.upw != 0) {   // the linear upwind branch (tendencies-only mode)
        dq = lin_tracer_axis<T, PX>(dq, a.q, a.u, cx, bx, 1, i, g.nx, a.upw);
        dq = lin_tracer_axis<T, PY>(dq, a.q, a.v, cy, by, g.sy, j, g.ny, a.upw);
        dq = lin_tracer_axis<T, PZ>(dq, a.q, a.w, cz, bz, g.sz, k, g.nz, a.upw);
    } else {
        if constexpr (PX != 0)
            dq = tracer_axis_per<T>(dq, a.q, a.u, cx, cb, per_line(i, g.nx, ES));
        else dq = tracer_axis<T, 0>(dq, a.q, a.u, cx, bx, 1, i, g.nx);
        if constexpr (PY != 0)
            dq = tracer_axis_per<T>(dq, a.q, a.v, cy, cb, per_line(j, g.ny, (int)g.sy * ES));
        else dq = tracer_axis<T, 0>(dq, a.q, a.v, cy, by, g.sy, j, g.ny);
        if constexpr (PZ != 0)
            dq = tracer_axis_per<T>(dq, a.q, a.w, cz, cb, per_line(k, g.nz, (int)g.sz * ES));
        else dq = tracer_axis<T, 0>(dq, a.q, a.w, cz, bz, g.sz, k, g.nz);
    }

    if (MODE == 0 || MODE == 1 || MODE == 2) stb(a.tout, cb, dq);
    const T cc = T(a.c);
    if (MODE == 1) {
        stb(a.qn, cb, ldb(a.q, cb) + cc * dq);
    } else if (MODE == 2) {
        stb(a.qn, cb, ldb(a.q, cb) + cc * (dq - T(3) * ldb(a.d0, cb)));
    } else if (MODE == 3) {
        stb(a.qn, cb, ldb(a.q, cb) + cc * ((T(8) * dq - ldb(a.d0, cb)) - ldb(a.d1, cb)));
    } else if (MODE == 4) {
        const T c2 = T(a.c2), c12 = T(1. / 12.);
        const T sb = ldb(a.sb, cb), t1 = sb + c2 * dq;
        stb(a.qn, cb, c12 * (T(5) * t1 + T(8) * ldb(a.q, cb) - sb));
        stb(a.sb, cb, ldb(a.q, cb));
    } else if (MODE == 5) {
        stb(a.qn, cb, ldb(a.qn, cb) + cc * dq);
    } else if (MODE == 6) {
        stb(a.sb, cb, ldb(a.q, cb));
        stb(a.qn, cb, ldb(a.q, cb) + cc * dq);
    }
}

template <typename T, int MODE, int PX, int PY, int PZ>
__global__ __launch_bounds__(256) void k_tracer_stage(Geo g, TracerArgs<T> a) {
    const Blk3 bk = xcd_block();
    const int i = bk.x * NYL_BX + threadIdx.x;
    const int j = bk.y * NYL_BY + threadIdx.y;
    if (i >= g.nx || j >= g.ny) return;
    tracer_cell<T, MODE, PX, PY, PZ>(g, a, i, j, bk.z);
}

template <typename T, int PX, int PY, int PZ>
static void launch_tracer_p(const Geo &g, const TracerArgs<T> &a, int mode, hipStream_t s) {
    dim3 blk(NYL_BX, NYL_BY, 1), grd = cell_grid(g);
    switch (mode) {
    case 0: NYL_LAUNCH((k_tracer_stage<T, 0, PX, PY, PZ>), grd, blk, 0, s, g, a); break;
    case 1: NYL_LAUNCH((k_tracer_stage<T, 1, PX, PY, PZ>), grd, blk, 0, s, g, a); break;
    case 2: NYL_LAUNCH((k_tracer_stage<T, 2, PX, PY, PZ>), grd, blk, 0, s, g, a); break;
    case 3: NYL_LAUNCH((k_tracer_stage<T, 3, PX, PY, PZ>), grd, blk, 0, s, g, a); break;
    case 4: NYL_LAUNCH((k_tracer_stage<T, 4, PX, PY, PZ>), grd, blk, 0, s, g, a); break;
    case 5: NYL_LAUNCH((k_tracer_stage<T, 5, PX, PY, PZ>), grd, blk, 0, s, g, a); break;
    case 6: NYL_LAUNCH((k_tracer_stage<T, 6, PX, PY, PZ>), grd, blk, 0, s, g, a); break;
    default: break;
    }
}

template <typename T>
hipError_t launch_tracer_stage(const Geo &g, const TracerArgs<T> &a, int mode, hipStream_t s) {
    if (mode < 0 || mode > 6) return hipErrorInvalidValue;
    switch (g.px + 2 * g.py + 4 * g.pz) {
    case 0: launch_tracer_p<T, 0, 0, 0>(g, a, mode, s); break;
    case 1: launch_tracer_p<T, 1, 0, 0>(g, a, mode, s); break;
    case 2: launch_tracer_p<T, 0, 1, 0>(g, a, mode, s); break;
    case 3: launch_tracer_p<T, 1, 1, 0>(g, a, mode, s); break;
    case 4: launch_tracer_p<T, 0, 0, 1>(g, a, mode, s); break;
    case 5: launch_tracer_p<T, 1, 0, 1>(g, a, mode, s); break;
    case 6: launch_tracer_p<T, 0, 1, 1>(g, a, mode, s); break;
    default: launch_tracer_p<T, 1, 1, 1>(g, a, mode, s); break;
    }
    return hipGetLastError();
}

// The stage combinations of stage_cell with the tendency read from memory
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_update(int64_t n, UpdateArgs<T> a) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    const T cc = T(a.c);
    if (MODE == 1) {
        a.sn[q] = a.s[q] + cc * a.ds0[q];
    } else if (MODE == 2) {
        a.sn[q] = a.s[q] + cc * (a.ds1[q] - T(3) * a.ds0[q]);
    } else if (MODE == 3) {
        a.sn[q] = a.s[q] + cc * ((T(8) * a.ds2[q] - a.ds0[q]) - a.ds1[q]);
    } else if (MODE == 4) {
        const T c2 = T(a.c2), c12 = T(1. / 12.);
        const T sb = a.sb[q], t1 = sb + c2 * a.ds0[q];
        a.sn[q] = c12 * (T(5) * t1 + T(8) * a.s[q] - sb);
        a.sb[q] = a.s[q];
    } else if (MODE == 5) {
        a.sn[q] = a.sn[q] + cc * a.ds0[q];
    } else if (MODE == 6) {
        a.sb[q] = a.s[q];
        a.sn[q] = a.s[q] + cc * a.ds0[q];
    }
}

template <typename T>
hipError_t launch_update(int64_t n, const UpdateArgs<T> &a, int mode, hipStream_t s) {
    const dim3 grd((unsigned)((n + 255) / 256));
    switch (mode) {
    case 1: NYL_LAUNCH((k_update<T, 1>), grd, dim3(256), 0, s, n, a); break;
    case 2: NYL_LAUNCH((k_update<T, 2>), grd, dim3(256), 0, s, n, a); break;
    case 3: NYL_LAUNCH((k_update<T, 3>), grd, dim3(256), 0, s, n, a); break;
    case 4: NYL_LAUNCH((k_update<T, 4>), grd, dim3(256), 0, s, n, a); break;
    case 5: NYL_LAUNCH((k_update<T, 5>), grd, dim3(256), 0, s, n, a); break;
    case 6: NYL_LAUNCH((k_update<T, 6>), grd, dim3(256), 0, s, n, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Four cells along x per thread: the per-cell kernels here are latency-bound
// at one cell per thread (one 4-byte load per field and cell); the x4 forms
// move the same bytes with 16-byte loads and stores, a quarter of the memory
// instructions.  Each cell's arithmetic is that of div_cell, correct_cell or
// diag_cell, term for term, so the results are the same bits.  They need
// nx % 4 == 0 and 16-byte aligned fields (the workspace is 256-byte aligned).

template <typename T>
struct Q4 {
    T v[4];
};
__device__ __forceinline__ Q4<float> ld4(const float *p) {
    const float4 q = *reinterpret_cast<const float4 *>(p);
    return {{q.x, q.y, q.z, q.w}};
}
__device__ __forceinline__ Q4<double> ld4(const double *p) {
    const double2 a = *reinterpret_cast<const double2 *>(p);
    const double2 b = *reinterpret_cast<const double2 *>(p + 2);
    return {{a.x, a.y, b.x, b.y}};
}
__device__ __forceinline__ void st4(float *p, const Q4<float> &q) {
    *reinterpret_cast<float4 *>(p) = make_float4(q.v[0], q.v[1], q.v[2], q.v[3]);
}
__device__ __forceinline__ void st4(double *p, const Q4<double> &q) {
    *reinterpret_cast<double2 *>(p) = make_double2(q.v[0], q.v[1]);
    *reinterpret_cast<double2 *>(p + 2) = make_double2(q.v[2], q.v[3]);
}
// st4 with the nontemporal hint (the four stores merge into wide ones)
template <typename T>
__device__ __forceinline__ void st4_stream(T *p, const Q4<T> &q) {
#pragma unroll
    for (int m = 0; m < 4; m++) store_stream(p + m, q.v[m]);
}

static inline dim3 cell_grid4(const Geo &g) {
    return dim3((g.nx / 4 + NYL_BX - 1) / NYL_BX, (g.ny + NYL_BY - 1) / NYL_BY, g.nz / cell_kz(g));
}

static inline bool quad_ok(const Geo &g, std::initializer_list<const void *> ptrs) {
    if (g.nx % 4 != 0) return false;
    for (const void *p : ptrs)
        if (p && (reinterpret_cast<uintptr_t>(p) & 15u) != 0) return false;
    return true;
}

// diag_cell for cells i0 .. i0+3; WHAT is the diag_cell selection as a
// compile-time constant.  FP32: every load the selection needs is issued at
// the top of a plane, together (one memory latency per plane; a neighbour
// outside a wall axis is loaded from the own cell instead and never used),
// and the results are stored with the nontemporal hint (the next stage
// kernel reads them): 115 -> 95 us at 256^3.  FP64: the neighbour loads of
// each part just before it and plain stores (the register cost of hoisting
// the FP64 loads lowers occupancy; 206 vs 232 us, and the hint gains nothing).
template <typename T, int WHAT>
__global__ __launch_bounds__(256) void k_diag4(Geo g, DiagArgs<T> a) {
    constexpr bool HOIST = sizeof(T) == 4;
    const Blk3 bk = CELL_BLOCK();
    const int i0 = (bk.x * NYL_BX + threadIdx.x) * 4;
    const int j = bk.y * NYL_BY + threadIdx.y;
    if (i0 >= g.nx || j >= g.ny) return;
    const int kz = g.nz / (int)gridDim.z;
    // x neighbours of the quad's end cells
    const bool hxm = g.px || i0 > 0;                  // cell i0 has an x-1 neighbour
    const int64_t oxm = g.px ? (int64_t)(i0 - wrapi(i0 - 1, g.nx)) : 1;
    const bool ex3 = g.px || i0 + 3 < g.nx - 1;       // cell i0+3 has an x+1 neighbour
    const int64_t oxp3 = g.px ? (int64_t)(wrapi(i0 + 4, g.nx) - (i0 + 3)) : 1;
    const bool ey = g.py || j < g.ny - 1, hym = g.py || j > 0;
    const int64_t oyp = g.py ? (int64_t)(wrapi(j + 1, g.ny) - j) * g.sy : g.sy;
    const int64_t oym = g.py ? (int64_t)(j - wrapi(j - 1, g.ny)) * g.sy : g.sy;
    auto out4 = [&](T *p, const Q4<T> &q) {
        if constexpr (HOIST) st4_stream(p, q);
        else st4(p, q);
    };
    for (int k = bk.z * kz; k < (bk.z + 1) * kz; k++) {
        const int64_t c = (int64_t)k * g.sz + (int64_t)j * g.sy + i0;
        const bool ez = g.pz || k < g.nz - 1, hzm = g.pz || k > 0;
        const int64_t ozp = g.pz ? (int64_t)(wrapi(k + 1, g.nz) - k) * g.sz : g.sz;
        const int64_t ozm = g.pz ? (int64_t)(k - wrapi(k - 1, g.nz)) * g.sz : g.sz;
        const Q4<T> u = ld4(a.u + c), v = ld4(a.v + c), w = ld4(a.w + c);
        T v4 = T(0), w4 = T(0), um0 = T(0);
        Q4<T> uyp{}, wyp{}, vzp{}, uzp{}, vm{}, wm{};
        if (HOIST && (WHAT & 2)) {
            const int64_t cx = c + 3 + (ex3 ? oxp3 : 0);
            const int64_t cy = c + (ey ? oyp : 0), cz = c + (ez ? ozp : 0);
            v4 = a.v[cx];
            w4 = a.w[cx];
            uyp = ld4(a.u + cy); wyp = ld4(a.w + cy);
            vzp = ld4(a.v + cz); uzp = ld4(a.u + cz);
        }
        if (HOIST && (WHAT & 4)) {
            um0 = a.u[c - (hxm ? oxm : 0)];
            vm = ld4(a.v + c - (hym ? oym : 0));
            wm = ld4(a.w + c - (hzm ? ozm : 0));
        }
        if (WHAT & 1) {
            Q4<T> X, Y, Z;
#pragma unroll
            for (int m = 0; m < 4; m++) {
                X.v[m] = u.v[m] * T(a.idx2);
                Y.v[m] = v.v[m] * T(a.idy2);
                Z.v[m] = w.v[m] * T(a.idz2);
            }
            out4(a.Ux + c, X); out4(a.Uy + c, Y); out4(a.Uz + c, Z);
        }
        if (WHAT & 2) {
            if (!HOIST) {
                v4 = ex3 ? a.v[c + 3 + oxp3] : T(0);
                w4 = ex3 ? a.w[c + 3 + oxp3] : T(0);
                if (ey) { uyp = ld4(a.u + c + oyp); wyp = ld4(a.w + c + oyp); }
                if (ez) { vzp = ld4(a.v + c + ozp); uzp = ld4(a.u + c + ozp); }
            }
            Q4<T> OX, OY, OZ;
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const bool ex = m < 3 || ex3;
                const T vxp = m < 3 ? v.v[m + 1] : v4, wxp = m < 3 ? w.v[m + 1] : w4;
                const T uc = u.v[m], vc = v.v[m], wc = w.v[m];
                T oz = T(0);
                if (ex && ey) {
                    oz = vxp - vc - uyp.v[m] + uc;
                    if (a.fpar > 0.0) oz = oz + T(a.fpar);
                }
                T ox = T(0);
                if (ey && ez) ox = wyp.v[m] - wc - vzp.v[m] + vc;
                T oy = T(0);
                if (ez && ex) oy = uzp.v[m] - uc - wxp + wc;
                OX.v[m] = ox; OY.v[m] = oy; OZ.v[m] = oz;
            }
            out4(a.ox + c, OX); out4(a.oy + c, OY); out4(a.oz + c, OZ);
        }
        if (WHAT & 4) {
            if (!HOIST) {
                um0 = hxm ? a.u[c - oxm] : T(0);
                if (hym) vm = ld4(a.v + c - oym);
                if (hzm) wm = ld4(a.w + c - ozm);
            }
            Q4<T> KE;
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const T uc = u.v[m], vc = v.v[m], wc = w.v[m];
                T ke = T(0);
                if (m > 0 || hxm) {
                    const T um = m > 0 ? u.v[m - 1] : um0;
                    const T cff2 = T(0.5 * a.idx2);
                    ke = ke + cff2 * T(0.5) * (uc * uc + um * um);
                }
                if (hym) {
                    const T cff2 = T(0.5 * a.idy2);
                    ke = ke + cff2 * T(0.5) * (vc * vc + vm.v[m] * vm.v[m]);
                }
                if (hzm) {
                    const T cff2 = T(0.5 * a.idz2);
                    ke = ke + cff2 * T(0.5) * (wc * wc + wm.v[m] * wm.v[m]);
                }
                KE.v[m] = ke;
            }
            out4(a.ke + c, KE);
        }
    }
}

template <typename T>
static void launch_diag4(const Geo &g, const DiagArgs<T> &a, int what, hipStream_t s) {
    const dim3 grd = cell_grid4(g), blk(NYL_BX, NYL_BY, 1);
    switch (what & 7) {
    case 1: NYL_LAUNCH((k_diag4<T, 1>), grd, blk, 0, s, g, a); break;
    case 2: NYL_LAUNCH((k_diag4<T, 2>), grd, blk, 0, s, g, a); break;
    case 3: NYL_LAUNCH((k_diag4<T, 3>), grd, blk, 0, s, g, a); break;
    case 4: NYL_LAUNCH((k_diag4<T, 4>), grd, blk, 0, s, g, a); break;
    case 5: NYL_LAUNCH((k_diag4<T, 5>), grd, blk, 0, s, g, a); break;
    case 6: NYL_LAUNCH((k_diag4<T, 6>), grd, blk, 0, s, g, a); break;
    case 7: NYL_LAUNCH((k_diag4<T, 7>), grd, blk, 0, s, g, a); break;
    default: break;
    }
}

// div_cell for cells i0 .. i0+3
template <typename T, typename TO>
__global__ __launch_bounds__(256) void k_div4(Geo g, DivArgs<T, TO> a) {
    const Blk3 bk = CELL_BLOCK();
    const int i0 = (bk.x * NYL_BX + threadIdx.x) * 4;
    const int j = bk.y * NYL_BY + threadIdx.y;
    if (i0 >= g.nx || j >= g.ny) return;
    const int kz = g.nz / (int)gridDim.z;
    const TO cx = TO(a.idx2), cy = TO(a.idy2), cz = TO(a.idz2);
    const bool hxm = g.px || i0 > 0, hym = g.py || j > 0;
    const int64_t oxm = g.px ? (int64_t)(i0 - wrapi(i0 - 1, g.nx)) : 1;
    const int64_t oym = g.py ? (int64_t)(j - wrapi(j - 1, g.ny)) * g.sy : g.sy;
    for (int k = bk.z * kz; k < (bk.z + 1) * kz; k++) {
        const int64_t c = (int64_t)k * g.sz + (int64_t)j * g.sy + i0;
        const bool hzm = g.pz || k > 0;
        const int64_t ozm = g.pz ? (int64_t)(k - wrapi(k - 1, g.nz)) * g.sz : g.sz;
        // every load at once; a neighbour outside a wall axis is read from
        // the own cell and never used
        const Q4<T> u = ld4(a.u + c), v = ld4(a.v + c), w = ld4(a.w + c);
        const T um0 = a.u[c - (hxm ? oxm : 0)];
        const Q4<T> vm = ld4(a.v + c - (hym ? oym : 0)), wm = ld4(a.w + c - (hzm ? ozm : 0));
        Q4<TO> D;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            TO d;
            const TO Uc = TO(u.v[m]) * cx;
            if (m == 0 && !hxm) d = Uc;
            else d = Uc - TO(m > 0 ? u.v[m - 1] : um0) * cx;
            const TO Vc = TO(v.v[m]) * cy;
            if (!hym) d = d + Vc;
            else d = d + (Vc - TO(vm.v[m]) * cy);
            const TO Wc = TO(w.v[m]) * cz;
            if (!hzm) d = d + Wc;
            else d = d + (Wc - TO(wm.v[m]) * cz);
            D.v[m] = d;
        }
        st4(a.out + c, D);
    }
}

// correct_cell for cells i0 .. i0+3 (a cell without the face keeps its value)
template <typename T, typename TX>
__global__ __launch_bounds__(256) void k_correct4(Geo g, CorrectArgs<T, TX> a) {
    const Blk3 bk = CELL_BLOCK();
    const int i0 = (bk.x * NYL_BX + threadIdx.x) * 4;
    const int j = bk.y * NYL_BY + threadIdx.y;
    if (i0 >= g.nx || j >= g.ny) return;
    const int kz = g.nz / (int)gridDim.z;
    const bool ex3 = g.px || i0 + 3 < g.nx - 1;
    const int64_t oxp3 = g.px ? (int64_t)(wrapi(i0 + 4, g.nx) - (i0 + 3)) : 1;
    const bool ey = g.py || j < g.ny - 1;
    const int64_t oyp = g.py ? (int64_t)(wrapi(j + 1, g.ny) - j) * g.sy : g.sy;
    const TX dx2 = TX(a.dx2);
    for (int k = bk.z * kz; k < (bk.z + 1) * kz; k++) {
        const int64_t c = (int64_t)k * g.sz + (int64_t)j * g.sy + i0;
        const bool ez = g.pz || k < g.nz - 1;
        const int64_t ozp = g.pz ? (int64_t)(wrapi(k + 1, g.nz) - k) * g.sz : g.sz;
        const Q4<TX> x = ld4(a.x + c);
        Q4<T> u = ld4(a.u + c), v = ld4(a.v + c), w = ld4(a.w + c);
        const TX x4 = ex3 ? a.x[c + 3 + oxp3] : TX(0);
        Q4<TX> xy{}, xz{};
        if (ey) xy = ld4(a.x + c + oyp);
        if (ez) xz = ld4(a.x + c + ozp);
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const TX pc = x.v[m] * dx2;
            if (m < 3 || ex3) u.v[m] = T(TX(u.v[m]) - ((m < 3 ? x.v[m + 1] : x4) * dx2 - pc));
            if (ey) v.v[m] = T(TX(v.v[m]) - (xy.v[m] * dx2 - pc));
            if (ez) w.v[m] = T(TX(w.v[m]) - (xz.v[m] * dx2 - pc));
        }
        st4(a.u + c, u); st4(a.v + c, v); st4(a.w + c, w);
    }
}

// U_from_u (core/cov_to_contra.py:15-20), vorticity (core/vorticity.py:25-34,
// core/fortran_vorticity.f90:2-28) and kinenergy (core/kinenergy.py:13-24,
// core/fortran_kinenergy.f90:47-55) from the covariant velocity.
template <typename T>
__device__ __forceinline__ void diag_cell(const Geo &g, const DiagArgs<T> &a, int what, int i,
                                          int j, int k) {
    const int64_t c = (int64_t)k * g.sz + (int64_t)j * g.sy + i;
    const T uc = a.u[c], vc = a.v[c], wc = a.w[c];
    if (what & 1) {
        a.Ux[c] = uc * T(a.idx2);
        a.Uy[c] = vc * T(a.idy2);
        a.Uz[c] = wc * T(a.idz2);
    }
    const int64_t oxp = g.px ? (int64_t)(wrapi(i + 1, g.nx) - i) : 1;
    const int64_t oyp = g.py ? (int64_t)(wrapi(j + 1, g.ny) - j) * g.sy : g.sy;
    const int64_t ozp = g.pz ? (int64_t)(wrapi(k + 1, g.nz) - k) * g.sz : g.sz;
    const bool ex = g.px || i < g.nx - 1, ey = g.py || j < g.ny - 1, ez = g.pz || k < g.nz - 1;
    if (what & 2) {
        // vor_k = delta_i u_j - delta_j u_i (flipview('k'): ui=u, uj=v)
        T oz = T(0);
        if (ex && ey) {
            oz = a.v[c + oxp] - vc - a.u[c + oyp] + uc;
            if (a.fpar > 0.0) oz = oz + T(a.fpar);
        }
        // vor_i (flipview('i') = [x][z][y]: ui=v, uj=w)
        T ox = T(0);
        if (ey && ez) ox = a.w[c + oyp] - wc - a.v[c + ozp] + vc;
        // vor_j (flipview('j') = [y][x][z]: ui=w, uj=u)
        T oy = T(0);
        if (ez && ex) oy = a.u[c + ozp] - uc - a.w[c + oxp] + wc;
        a.ox[c] = ox; a.oy[c] = oy; a.oz[c] = oz;
    }
    if (what & 4) {
        // ke = 0; ke(i) += (cff2*0.5)*(u(i)**2+u(i-1)**2) for i >= 2, per direction
        T ke = T(0);
        if (g.px || i > 0) {
            const T um = a.u[c - (g.px ? (int64_t)(i - wrapi(i - 1, g.nx)) : 1)];
            const T cff2 = T(0.5 * a.idx2);
            ke = ke + cff2 * T(0.5) * (uc * uc + um * um);
        }
        if (g.py || j > 0) {
            const T vm = a.v[c - (g.py ? (int64_t)(j - wrapi(j - 1, g.ny)) * g.sy : g.sy)];
            const T cff2 = T(0.5 * a.idy2);
            ke = ke + cff2 * T(0.5) * (vc * vc + vm * vm);
        }
        if (g.pz || k > 0) {
            const T wm = a.w[c - (g.pz ? (int64_t)(k - wrapi(k - 1, g.nz)) * g.sz : g.sz)];
            const T cff2 = T(0.5 * a.idz2);
            ke = ke + cff2 * T(0.5) * (wc * wc + wm * wm);
        }
        a.ke[c] = ke;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_diag(Geo g, DiagArgs<T> a, int what) {
    const Blk3 bk = CELL_BLOCK();
    const int i = bk.x * NYL_BX + threadIdx.x;
    const int j = bk.y * NYL_BY + threadIdx.y;
    if (i >= g.nx || j >= g.ny) return;
    const int kz = g.nz / (int)gridDim.z;
    for (int k = bk.z * kz; k < (bk.z + 1) * kz; k++) diag_cell<T>(g, a, what, i, j, k);
}

template <typename T>
hipError_t launch_diag(const Geo &g, const DiagArgs<T> &a, int what, hipStream_t s) {
    if (quad_ok(g, {a.u, a.v, a.w, a.Ux, a.Uy, a.Uz, a.ox, a.oy, a.oz, a.ke}))
        launch_diag4<T>(g, a, what, s);
    else
        NYL_LAUNCH((k_diag<T>), cell_grid_kz(g), dim3(NYL_BX, NYL_BY, 1), 0, s, g, a, what);
    return hipGetLastError();
}

// compute_div (core/projection.py:15-35 over core/fortran_bernoulli.f90:61-97)
// with U = u*idx2 evaluated in place (core/cov_to_contra.py:18-20).
template <typename T, typename TO>
__device__ __forceinline__ void div_cell(const Geo &g, const DivArgs<T, TO> &a, int i, int j,
                                         int k) {
    const int64_t c = (int64_t)k * g.sz + (int64_t)j * g.sy + i;
    // evaluated in the output precision (FP64 right-hand side of the mixed path)
    const TO cx = TO(a.idx2), cy = TO(a.idy2), cz = TO(a.idz2);
    TO d;
    const TO Uc = TO(a.u[c]) * cx;
    if (!g.px && i == 0) d = Uc;
    else d = Uc - TO(a.u[c - (g.px ? (int64_t)(i - wrapi(i - 1, g.nx)) : 1)]) * cx;
    const TO Vc = TO(a.v[c]) * cy;
    if (!g.py && j == 0) d = d + Vc;
    else d = d + (Vc - TO(a.v[c - (g.py ? (int64_t)(j - wrapi(j - 1, g.ny)) * g.sy : g.sy)]) * cy);
    const TO Wc = TO(a.w[c]) * cz;
    if (!g.pz && k == 0) d = d + Wc;
    else d = d + (Wc - TO(a.w[c - (g.pz ? (int64_t)(k - wrapi(k - 1, g.nz)) * g.sz : g.sz)]) * cz);
    a.out[c] = d;
}

template <typename T, typename TO>
__global__ __launch_bounds__(256) void k_div(Geo g, DivArgs<T, TO> a) {
    const Blk3 bk = CELL_BLOCK();
    const int i = bk.x * NYL_BX + threadIdx.x;
    const int j = bk.y * NYL_BY + threadIdx.y;
    if (i >= g.nx || j >= g.ny) return;
    const int kz = g.nz / (int)gridDim.z;
    for (int k = bk.z * kz; k < (bk.z + 1) * kz; k++) div_cell<T, TO>(g, a, i, j, k);
}

template <typename T, typename TO>
hipError_t launch_div(const Geo &g, const T *u, const T *v, const T *w, double idx2,
                      double idy2, double idz2, TO *out, hipStream_t s) {
    DivArgs<T, TO> a;
    a.u = u; a.v = v; a.w = w; a.idx2 = idx2; a.idy2 = idy2; a.idz2 = idz2; a.out = out;
    if (quad_ok(g, {u, v, w, out}))
        NYL_LAUNCH((k_div4<T, TO>), cell_grid4(g), dim3(NYL_BX, NYL_BY, 1), 0, s, g, a);
    else
        NYL_LAUNCH((k_div<T, TO>), cell_grid_kz(g), dim3(NYL_BX, NYL_BY, 1), 0, s, g, a);
    return hipGetLastError();
}

// u -= delta[p] with p = x*dx**2 (core/projection.py:84-87 over
// core/fortran_bernoulli.f90:2-26; p from core/mgfordriver.py:78)
template <typename T, typename TX>
__device__ __forceinline__ void correct_cell(const Geo &g, const CorrectArgs<T, TX> &a, int i,
                                             int j, int k) {
    const int64_t c = (int64_t)k * g.sz + (int64_t)j * g.sy + i;
    const TX *x = a.x;
    const TX pc = x[c] * TX(a.dx2);
    if (g.px || i < g.nx - 1) {
        const int64_t o = g.px ? (int64_t)(wrapi(i + 1, g.nx) - i) : 1;
        a.u[c] = T(TX(a.u[c]) - (x[c + o] * TX(a.dx2) - pc));
    }
    if (g.py || j < g.ny - 1) {
        const int64_t o = g.py ? (int64_t)(wrapi(j + 1, g.ny) - j) * g.sy : g.sy;
        a.v[c] = T(TX(a.v[c]) - (x[c + o] * TX(a.dx2) - pc));
    }
    if (g.pz || k < g.nz - 1) {
        const int64_t o = g.pz ? (int64_t)(wrapi(k + 1, g.nz) - k) * g.sz : g.sz;
        a.w[c] = T(TX(a.w[c]) - (x[c + o] * TX(a.dx2) - pc));
    }
}

template <typename T, typename TX>
__global__ __launch_bounds__(256) void k_correct(Geo g, CorrectArgs<T, TX> a) {
    const Blk3 bk = CELL_BLOCK();
    const int i = bk.x * NYL_BX + threadIdx.x;
    const int j = bk.y * NYL_BY + threadIdx.y;
    if (i >= g.nx || j >= g.ny) return;
    const int kz = g.nz / (int)gridDim.z;
    for (int k = bk.z * kz; k < (bk.z + 1) * kz; k++) correct_cell<T, TX>(g, a, i, j, k);
}

template <typename T, typename TX>
hipError_t launch_correct(const Geo &g, T *u, T *v, T *w, const TX *x, double dx2,
                          hipStream_t s) {
    CorrectArgs<T, TX> a;
    a.u = u; a.v = v; a.w = w; a.x = x; a.dx2 = dx2;
    if (quad_ok(g, {u, v, w, x}))
        NYL_LAUNCH((k_correct4<T, TX>), cell_grid4(g), dim3(NYL_BX, NYL_BY, 1), 0, s, g, a);
    else
        NYL_LAUNCH((k_correct<T, TX>), cell_grid_kz(g), dim3(NYL_BX, NYL_BY, 1), 0, s, g, a);
    return hipGetLastError();
}

template <typename TX, typename T>
__global__ void k_pressure(int64_t n, const TX *__restrict__ x, double dx2, T *__restrict__ p) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = T(x[q] * TX(dx2));
}

template <typename TX, typename T>
hipError_t launch_pressure(const Geo &g, const TX *x, double dx2, T *p, hipStream_t s) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    NYL_LAUNCH((k_pressure<TX, T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       n, x, dx2, p);
    return hipGetLastError();
}

__global__ void k_fill(double *p, int64_t n, double v) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = v;
}

template <typename T>
__global__ void k_zero(T *p, int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = T(0);
}

template <typename T>
hipError_t launch_zero(T *p, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    NYL_LAUNCH((k_zero<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}
template hipError_t launch_zero<float>(float *, int64_t, hipStream_t);
template hipError_t launch_zero<double>(double *, int64_t, hipStream_t);

hipError_t launch_fill(double *p, int64_t n, double v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    NYL_LAUNCH(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Ensemble batches: one launch covers every member (BHead), each block a
// tile of TPB cells of one member (rows of consecutive cells, or a brick:
// StageTile), running the single-member cell functions.  The stage kernel
// is persistent: a fixed grid strides over all tiles; its RK coefficient is
// the host's dt, dt/4., dt/12. of the member's dt (same IEEE operations).

// the autotuner's candidates: rows, then bricks from flat to cubic
const StageTile kStageTiles[NYL_N_TILES] = {
    {128, -1, 0}, {128, 5, 2}, {128, 4, 2}, {128, 3, 2},
    {256, -1, 0}, {256, 6, 2}, {256, 5, 2}, {256, 4, 2}, {256, 3, 3},
    {512, -1, 0}, {512, 6, 3}, {512, 5, 3}, {512, 5, 2}, {512, 4, 3}, {512, 3, 3}};

int64_t stage_tiles(const Geo &g, const StageTile &t) {
    if (t.lx < 0) return ((int64_t)g.nx * g.ny * g.nz + t.tpb - 1) / t.tpb;
    const int tz = t.tpb >> (t.lx + t.ly);
    return (int64_t)((g.nx + (1 << t.lx) - 1) >> t.lx) * ((g.ny + (1 << t.ly) - 1) >> t.ly)
         * ((g.nz + tz - 1) / tz);
}

// cell (i, j, k) of thread th in tile `tile` of a member; false for a thread
// past the member's edge
template <int TPB>
__device__ __forceinline__ bool tile_cell(const Geo &g, int64_t tile, int th, int lx, int ly,
                                          int &i, int &j, int &k) {
    if (lx < 0) return cell_of(g, tile * TPB + th, i, j, k);
    const uint32_t bx = (uint32_t)(g.nx + (1 << lx) - 1) >> lx;
    const uint32_t by = (uint32_t)(g.ny + (1 << ly) - 1) >> ly;
    const uint32_t t = (uint32_t)tile, r = t / bx, c = r / by;
    i = (int)(((t - r * bx) << lx) + (th & ((1 << lx) - 1)));
    j = (int)(((r - c * by) << ly) + ((th >> lx) & ((1 << ly) - 1)));
    k = (int)(c * (uint32_t)(TPB >> (lx + ly)) + (th >> (lx + ly)));
    return i < g.nx && j < g.ny && k < g.nz;
}

template <typename T, int MODE, int PX, int PY, int PZ, int TPB>
__global__ __launch_bounds__(TPB) void k_b_stage(const StageItem<T> *__restrict__ items, int n,
                                                 int64_t nblk, const double *__restrict__ dt,
                                                 int lx, int ly) {
    for (int64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
        const StageItem<T> &it = find_item(items, n, blk);
        int i, j, k;
        if (!tile_cell<TPB>(it.h.g, blk - it.h.start, threadIdx.x, lx, ly, i, j, k)) continue;
        StageArgs<T> a = it.a;
        const double d = dt[it.h.member];
        a.c = MODE == 1 ? d : (MODE == 2 ? d / 4. : d / 12.);
        stage_cell<T, MODE, PX, PY, PZ>(it.h.g, a, i, j, k);
    }
}

int batch_blocks(int64_t ncell) { return (int)((ncell + NYL_BATCH_TPB - 1) / NYL_BATCH_TPB); }

template <typename T, int PX, int PY, int PZ, int TPB>
static void b_stage_p(const StageItem<T> *items, int n, int64_t nblk, int mode, const double *dt,
                      int lx, int ly, hipStream_t s) {
    const unsigned grid = (unsigned)(nblk < 2048 ? nblk : 2048);
    switch (mode) {
    case 1: NYL_LAUNCH((k_b_stage<T, 1, PX, PY, PZ, TPB>), grid, TPB, 0, s, items, n, nblk, dt, lx, ly); break;
    case 2: NYL_LAUNCH((k_b_stage<T, 2, PX, PY, PZ, TPB>), grid, TPB, 0, s, items, n, nblk, dt, lx, ly); break;
    default: NYL_LAUNCH((k_b_stage<T, 3, PX, PY, PZ, TPB>), grid, TPB, 0, s, items, n, nblk, dt, lx, ly); break;
    }
}

template <typename T, int TPB>
static void b_stage_t(const StageItem<T> *items, int n, int64_t nblk, int mode, int perio,
                      const double *dt, int lx, int ly, hipStream_t s) {
    switch (perio) {
    case 0: b_stage_p<T, 0, 0, 0, TPB>(items, n, nblk, mode, dt, lx, ly, s); break;
    case 1: b_stage_p<T, 1, 0, 0, TPB>(items, n, nblk, mode, dt, lx, ly, s); break;
    case 2: b_stage_p<T, 0, 1, 0, TPB>(items, n, nblk, mode, dt, lx, ly, s); break;
    case 3: b_stage_p<T, 1, 1, 0, TPB>(items, n, nblk, mode, dt, lx, ly, s); break;
    case 4: b_stage_p<T, 0, 0, 1, TPB>(items, n, nblk, mode, dt, lx, ly, s); break;
    case 5: b_stage_p<T, 1, 0, 1, TPB>(items, n, nblk, mode, dt, lx, ly, s); break;
    case 6: b_stage_p<T, 0, 1, 1, TPB>(items, n, nblk, mode, dt, lx, ly, s); break;
    default: b_stage_p<T, 1, 1, 1, TPB>(items, n, nblk, mode, dt, lx, ly, s); break;
    }
}

template <typename T>
hipError_t launch_b_stage(const StageItem<T> *items, int n, int64_t nblk, int mode, int perio,
                          const StageTile &tile, const double *dt, hipStream_t s) {
    if (mode < 1 || mode > 3 || nblk <= 0) return hipErrorInvalidValue;
    if (tile.lx >= 0 && (tile.ly < 0 || tile.lx + tile.ly > 9 || (tile.tpb >> (tile.lx + tile.ly)) < 1))
        return hipErrorInvalidValue;
    switch (tile.tpb) {
    case 128: b_stage_t<T, 128>(items, n, nblk, mode, perio, dt, tile.lx, tile.ly, s); break;
    case 256: b_stage_t<T, 256>(items, n, nblk, mode, perio, dt, tile.lx, tile.ly, s); break;
    case 512: b_stage_t<T, 512>(items, n, nblk, mode, perio, dt, tile.lx, tile.ly, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T>
__global__ __launch_bounds__(NYL_BATCH_TPB) void k_b_div(const DivItem<T> *__restrict__ items,
                                                         int n) {
    const DivItem<T> &it = find_item(items, n, blockIdx.x);
    int i, j, k;
    if (!cell_of(it.h.g, ((int64_t)blockIdx.x - it.h.start) * NYL_BATCH_TPB + threadIdx.x, i, j, k))
        return;
    div_cell<T, double>(it.h.g, it.a, i, j, k);
}

template <typename T>
hipError_t launch_b_div(const DivItem<T> *items, int n, int64_t nblk, hipStream_t s) {
    NYL_LAUNCH((k_b_div<T>), dim3((unsigned)nblk), NYL_BATCH_TPB, 0, s, items, n);
    return hipGetLastError();
}

template <typename T>
__global__ __launch_bounds__(NYL_BATCH_TPB) void k_b_correct(const CorrItem<T> *__restrict__ items,
                                                             int n) {
    const CorrItem<T> &it = find_item(items, n, blockIdx.x);
    int i, j, k;
    if (!cell_of(it.h.g, ((int64_t)blockIdx.x - it.h.start) * NYL_BATCH_TPB + threadIdx.x, i, j, k))
        return;
    correct_cell<T, double>(it.h.g, it.a, i, j, k);
}

template <typename T>
hipError_t launch_b_correct(const CorrItem<T> *items, int n, int64_t nblk, hipStream_t s) {
    NYL_LAUNCH((k_b_correct<T>), dim3((unsigned)nblk), NYL_BATCH_TPB, 0, s, items, n);
    return hipGetLastError();
}

template <typename T>
__global__ __launch_bounds__(NYL_BATCH_TPB) void k_b_diag(const DiagItem<T> *__restrict__ items,
                                                          int n, int what) {
    const DiagItem<T> &it = find_item(items, n, blockIdx.x);
    int i, j, k;
    if (!cell_of(it.h.g, ((int64_t)blockIdx.x - it.h.start) * NYL_BATCH_TPB + threadIdx.x, i, j, k))
        return;
    diag_cell<T>(it.h.g, it.a, what, i, j, k);
}

template <typename T>
hipError_t launch_b_diag(const DiagItem<T> *items, int n, int64_t nblk, int what, hipStream_t s) {
    NYL_LAUNCH((k_b_diag<T>), dim3((unsigned)nblk), NYL_BATCH_TPB, 0, s, items, n, what);
    return hipGetLastError();
}

#define NYL_B_STENCIL_INST(T)                                                                    \
    template hipError_t launch_b_stage<T>(const StageItem<T> *, int, int64_t, int, int,         \
                                          const StageTile &, const double *, hipStream_t);      \
    template hipError_t launch_b_div<T>(const DivItem<T> *, int, int64_t, hipStream_t);         \
    template hipError_t launch_b_correct<T>(const CorrItem<T> *, int, int64_t, hipStream_t);    \
    template hipError_t launch_b_diag<T>(const DiagItem<T> *, int, int64_t, int, hipStream_t);
NYL_B_STENCIL_INST(double)
NYL_B_STENCIL_INST(float)
#undef NYL_B_STENCIL_INST

template hipError_t launch_stage<double>(const Geo &, const StageArgs<double> &, int, hipStream_t);
template hipError_t launch_stage<float>(const Geo &, const StageArgs<float> &, int, hipStream_t);
template hipError_t launch_tracer_stage<double>(const Geo &, const TracerArgs<double> &, int,
                                                hipStream_t);
template hipError_t launch_tracer_stage<float>(const Geo &, const TracerArgs<float> &, int,
                                               hipStream_t);
template hipError_t launch_update<double>(int64_t, const UpdateArgs<double> &, int, hipStream_t);
template hipError_t launch_update<float>(int64_t, const UpdateArgs<float> &, int, hipStream_t);
template hipError_t launch_diag<double>(const Geo &, const DiagArgs<double> &, int, hipStream_t);
template hipError_t launch_diag<float>(const Geo &, const DiagArgs<float> &, int, hipStream_t);
template hipError_t launch_div<double, double>(const Geo &, const double *, const double *,
                                               const double *, double, double, double, double *,
                                               hipStream_t);
template hipError_t launch_div<float, double>(const Geo &, const float *, const float *,
                                              const float *, double, double, double, double *,
                                              hipStream_t);
template hipError_t launch_correct<double, double>(const Geo &, double *, double *, double *,
                                                   const double *, double, hipStream_t);
template hipError_t launch_correct<float, double>(const Geo &, float *, float *, float *,
                                                  const double *, double, hipStream_t);
template hipError_t launch_pressure<double, double>(const Geo &, const double *, double,
                                                    double *, hipStream_t);

}  // namespace nyl
