// nyl_device.h -- device helpers shared by the Nyles HIP kernels (gfx950).
//
// Fields are interior-only, C order [z][y][x].  Along a periodic axis a
// neighbour index is wrapped; along a wall axis the reference's array-edge
// closures are applied (the reference stores no halo on wall axes,
// core/mpi/topology.py:253-304).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nyl {

struct Geo {
    int nx, ny, nz;      // interior cells
    int px, py, pz;      // 1 = periodic axis
    int64_t sy, sz;      // strides of y and z (x stride is 1)
};

__device__ __forceinline__ int wrapi(int m, int n) {
    // |m - [0,n)| < n by construction (offsets are at most 3, n >= 4)
    return m < 0 ? m + n : (m >= n ? m - n : m);
}

// Block coordinates remapped so that each of the 8 XCDs (which receive
// workgroups round-robin in dispatch order) owns a contiguous range of the
// grid: with x fastest, then y, then z, an XCD sweeps a slab of z planes and
// the z-neighbours of a plane stay in its own L2.
struct Blk3 {
    int x, y, z;
};

// Same, with the blocks of an XCD ordered x fastest, then z, then y: each
// XCD owns a slab of y rows and sweeps it plane after plane, so a stencil
// reaching several planes up and down (the stage kernel's +-3) finds them in
// that XCD's L2 (the working set is a y slab, not whole planes).
__device__ __forceinline__ Blk3 xcd_block_yslab() {
    const uint32_t gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
    const uint32_t nb = gx * gy * gz;
    uint32_t b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const uint32_t per = nb >> 3;
    if (b < (per << 3)) b = (b & 7u) * per + (b >> 3);
    Blk3 r;
    r.x = (int)(b % gx);
    const uint32_t t = b / gx;
    r.z = (int)(t % gz);
    r.y = (int)(t / gz);
    return r;
}

__device__ __forceinline__ Blk3 xcd_block() {
    const uint32_t gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
    const uint32_t nb = gx * gy * gz;
    uint32_t b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const uint32_t per = nb >> 3;
    if (b < (per << 3)) b = (b & 7u) * per + (b >> 3);
    Blk3 r;
    r.x = (int)(b % gx);
    const uint32_t t = b / gx;
    r.y = (int)(t % gy);
    r.z = (int)(t / gy);
    return r;
}

// ---------------------------------------------------------------------------
// WENO reconstructions of core/weno.f90.  Real literals there are REAL(4)
// constants widened to the working precision (the f2py build has no
// -fdefault-real-8); tau5 is an implicitly typed REAL(4) (core/weno.f90:45).
// Each operation is written in the reference's evaluation order; FP64
// instantiations are compiled without contraction.

// Division of the reconstruction weights: correctly rounded in FP64 (the
// reference's arithmetic); the FP32 path (mixed precision, no bitwise
// target) multiplies by the hardware reciprocal (v_rcp_f32, about 1 ulp).
// Every divisor is >= eps (1e-16 or 1e-14, normal floats).
template <typename T>
__device__ __forceinline__ T wdiv(T a, T b) {
    if constexpr (sizeof(T) == 4) return a * __builtin_amdgcn_rcpf(b);
    else return a / b;
}

// core/weno.f90:1-22
template <typename T>
__device__ __forceinline__ T weno3(T qm, T q0, T qp) {
    const T eps = T(1e-14f);
    T qi1 = (-qm + T(3) * q0) * T(0.5);
    T qi2 = (q0 + qp) * T(0.5);
    T d1 = q0 - qm, d2 = qp - q0;
    T beta1 = d1 * d1;
    T beta2 = d2 * d2;
    T tau = fabs(beta2 - beta1);
    T w1 = (T(1) + wdiv(tau, beta1 + eps));
    T w2 = (T(1) + wdiv(tau, beta2 + eps)) * T(2);
    return wdiv(w1 * qi1 + w2 * qi2, w1 + w2);
}

// The FP32 reconstruction of the mixed-precision path (no bitwise target):
// core/weno.f90:25-54 with the products and sums contracted into fused
// multiply-adds and the weights' divisions by the hardware reciprocal, about
// 40 % fewer instructions than the operation-by-operation form (the FP32
// stage kernel issues VALU instructions at the ceiling).  Same literals and
// weights.
__device__ __forceinline__ float weno5_f32(float qmm, float qm, float q0, float qp, float qpp) {
    const float eps = 1e-16f;
    const float c13 = 1.0f / 3.0f, c76 = 7.0f / 6.0f, c116 = 11.0f / 6.0f;
    const float c16 = 1.0f / 6.0f, c56 = 5.0f / 6.0f;
    const float k1 = 13.0f / 12.0f, k2 = 0.25f;
    const float qi1 = fmaf(c13, qmm, fmaf(-c76, qm, c116 * q0));
    const float qi2 = fmaf(c56, q0, fmaf(c13, qp, -(c16 * qm)));
    const float qi3 = fmaf(c13, q0, fmaf(c56, qp, -(c16 * qpp)));
    const float t1 = fmaf(-2.0f, qm, qmm + q0), t2 = fmaf(-4.0f, qm, fmaf(3.0f, q0, qmm));
    const float t3 = fmaf(-2.0f, q0, qm + qp), t4 = qm - qp;
    const float t5 = fmaf(-2.0f, qp, q0 + qpp), t6 = fmaf(-4.0f, qp, fmaf(3.0f, q0, qpp));
    const float beta1 = fmaf(k1 * t1, t1, (k2 * t2) * t2);
    const float beta2 = fmaf(k1 * t3, t3, (k2 * t4) * t4);
    const float beta3 = fmaf(k1 * t5, t5, (k2 * t6) * t6);
    const float tau5 = fabsf(beta1 - beta3);
    const float w1 = fmaf(tau5, __builtin_amdgcn_rcpf(beta1 + eps), 1.0f);
    const float w2 = fmaf(6.0f * tau5, __builtin_amdgcn_rcpf(beta2 + eps), 6.0f);
    const float w3 = fmaf(3.0f * tau5, __builtin_amdgcn_rcpf(beta3 + eps), 3.0f);
    return fmaf(w1, qi1, fmaf(w2, qi2, w3 * qi3)) * __builtin_amdgcn_rcpf(w1 + w2 + w3);
}

// core/weno.f90:25-54
template <typename T>
__device__ __forceinline__ T weno5(T qmm, T qm, T q0, T qp, T qpp) {
    if constexpr (sizeof(T) == 4) return weno5_f32(qmm, qm, q0, qp, qpp);
    const T eps = T(1e-16f);
    const T c13 = T(1.0f / 3.0f), c76 = T(7.0f / 6.0f), c116 = T(11.0f / 6.0f);
    const T c16 = T(1.0f / 6.0f), c56 = T(5.0f / 6.0f);
    const T k1 = T(13.0f / 12.0f), k2 = T(0.25);

    T qi1 = c13 * qmm - c76 * qm + c116 * q0;
    T qi2 = -(c16 * qm) + c56 * q0 + c13 * qp;
    T qi3 = c13 * q0 + c56 * qp - c16 * qpp;

    T t1 = qmm - T(2) * qm + q0, t2 = qmm - T(4) * qm + T(3) * q0;
    T beta1 = k1 * (t1 * t1) + k2 * (t2 * t2);
    T t3 = qm - T(2) * q0 + qp, t4 = qm - qp;
    T beta2 = k1 * (t3 * t3) + k2 * (t4 * t4);
    T t5 = q0 - T(2) * qp + qpp, t6 = T(3) * q0 - T(4) * qp + qpp;
    T beta3 = k1 * (t5 * t5) + k2 * (t6 * t6);

    T tau5 = T((float)fabs(beta1 - beta3));

    T w1 = (T(1) + wdiv(tau5, beta1 + eps));
    T w2 = T(6) * (T(1) + wdiv(tau5, beta2 + eps));
    T w3 = T(3) * (T(1) + wdiv(tau5, beta3 + eps));
    return wdiv(w1 * qi1 + w2 * qi2 + w3 * qi3, w1 + w2 + w3);
}

// Flux at line position f (0-based) of core/weno.f90:106-153 flux1d, for a
// line of n cells; u is the advecting velocity at f, q(m) the advected
// quantity at line position m.  On a periodic axis every position uses the
// interior WENO5 stencil (the reference's halo makes the closures fall in
// the halo).  On a wall axis the closures are applied with the reference's
// precedence (later assignments win): n-1 > n-2 > n-3 > {0, 1} > interior.
// The stencil values are selected first and one reconstruction evaluated,
// so a wavefront with mixed upwind directions runs WENO5 once.  PER is a
// compile-time flag so that periodic axes compile to the interior path only.
template <typename T, int PER, typename QF>
__device__ __forceinline__ T face_flux(int f, int n, T u, const QF &q) {
    const bool pos = u > T(0);
    if (PER || (f >= 2 && f <= n - 4)) {
        T a0, a1, a2, a3, a4;
        if (pos) { a0 = q(f - 2); a1 = q(f - 1); a2 = q(f); a3 = q(f + 1); a4 = q(f + 2); }
        else     { a0 = q(f + 3); a1 = q(f + 2); a2 = q(f + 1); a3 = q(f); a4 = q(f - 1); }
        return u * weno5<T>(a0, a1, a2, a3, a4);
    }
    int kind;
    T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0), a4 = T(0);
    if (f == n - 1) {
        kind = 0;
    } else if (f == n - 2) {
        if (pos) { kind = 3; a0 = q(f - 1); a1 = q(f); a2 = q(f + 1); }
        else     { kind = 1; a0 = q(f + 1); }
    } else if (f == n - 3) {
        if (pos) { kind = 5; a0 = q(f - 2); a1 = q(f - 1); a2 = q(f); a3 = q(f + 1); a4 = q(f + 2); }
        else     { kind = 3; a0 = q(f + 2); a1 = q(f + 1); a2 = q(f); }
    } else if (f == 0) {
        if (pos) { kind = 1; a0 = q(0); }
        else     { kind = 3; a0 = q(2); a1 = q(1); a2 = q(0); }
    } else {  // f == 1
        if (pos) { kind = 3; a0 = q(0); a1 = q(1); a2 = q(2); }
        else     { kind = 5; a0 = q(4); a1 = q(3); a2 = q(2); a3 = q(1); a4 = q(0); }
    }
    T w;
    if (kind == 5) w = weno5<T>(a0, a1, a2, a3, a4);
    else if (kind == 3) w = weno3<T>(a0, a1, a2);
    else if (kind == 1) w = a0;
    else return T(0);
    return u * w;
}

// face_flux with the stencil given relative to the face: qr(d) = q(f + d),
// d in [-2, 3] (the positions every branch of face_flux reads).  For values
// held in registers: each branch reads constant offsets.  Same values,
// same operations, same results as face_flux.
template <typename T, int PER, typename QR>
__device__ __forceinline__ T face_flux_rel(int f, int n, T u, const QR &qr) {
    const bool pos = u > T(0);
    // every position read up front with a constant offset: the selections
    // below then pick among values, never among positions (a position picked
    // at run time would index the register window dynamically)
    const T m2 = qr(-2), m1 = qr(-1), z0 = qr(0), p1 = qr(1), p2 = qr(2), p3 = qr(3);
    if (PER || (f >= 2 && f <= n - 4)) {
        T a0, a1, a2, a3, a4;
        if (pos) { a0 = m2; a1 = m1; a2 = z0; a3 = p1; a4 = p2; }
        else     { a0 = p3; a1 = p2; a2 = p1; a3 = z0; a4 = m1; }
        return u * weno5<T>(a0, a1, a2, a3, a4);
    }
    int kind;
    T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0), a4 = T(0);
    if (f == n - 1) {
        kind = 0;
    } else if (f == n - 2) {
        if (pos) { kind = 3; a0 = m1; a1 = z0; a2 = p1; }
        else     { kind = 1; a0 = p1; }
    } else if (f == n - 3) {
        if (pos) { kind = 5; a0 = m2; a1 = m1; a2 = z0; a3 = p1; a4 = p2; }
        else     { kind = 3; a0 = p2; a1 = p1; a2 = z0; }
    } else if (f == 0) {
        if (pos) { kind = 1; a0 = z0; }
        else     { kind = 3; a0 = p2; a1 = p1; a2 = z0; }
    } else {  // f == 1
        if (pos) { kind = 3; a0 = m1; a1 = z0; a2 = p1; }
        else     { kind = 5; a0 = p3; a1 = p2; a2 = p1; a3 = z0; a4 = m1; }
    }
    T w;
    if (kind == 5) w = weno5<T>(a0, a1, a2, a3, a4);
    else if (kind == 3) w = weno3<T>(a0, a1, a2);
    else if (kind == 1) w = a0;
    else return T(0);
    return u * w;
}

// Store with the nontemporal hint, for results that no kernel reads soon
// after (the next RK stage's tendencies, the diagnostics read by the next
// stage kernel): a stream of such stores runs at close to the HBM write rate,
// where write-allocating them in the caches costs far more.
template <typename T>
__device__ __forceinline__ void store_stream(T *p, T v) {
    __builtin_nontemporal_store(v, p);
}

// ---------------------------------------------------------------------------
// Ensemble batches: the item (one per member) owning flattened block `blk`
// (items sorted by h.start, h.start of the first item is 0).
template <typename I>
__device__ __forceinline__ const I &find_item(const I *items, int n, int64_t blk) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].h.start <= blk) lo = mid;
        else hi = mid - 1;
    }
    return items[lo];
}

// cell (i, j, k) of flattened index q of a level, false past its last cell
// (a level has fewer than 2^31 cells: the divisions run in 32 bits, far
// cheaper than 64-bit ones on the GPU)
__device__ __forceinline__ bool cell_of(const Geo &g, int64_t q, int &i, int &j, int &k) {
    if (q >= (int64_t)g.nx * g.ny * g.nz) return false;
    const uint32_t q32 = (uint32_t)q, nx = (uint32_t)g.nx, ny = (uint32_t)g.ny;
    const uint32_t t = q32 / nx;
    i = (int)(q32 - t * nx);
    j = (int)(t % ny);
    k = (int)(t / ny);
    return true;
}

}  // namespace nyl
