// nyl_noalloc.hip -- nyl_reduce on caller-owned fields with no context, no
// workspace and no device allocation, for callers that cannot allocate
// (code under hipGraph capture, tight serving loops).
//
// The reductions are two-stage deterministic sums (or maxima): one partial
// per block of the first kernel, then a one-block kernel over the partials.
// nyl_reduce keeps the partials in the context's workspace; here they live in
// a file-scope device array, `scratch`, of NYL_RED_BLOCKS doubles (128 KiB per
// GPU, held until the process ends).  The first-stage kernel writes every
// partial the second stage reads, so nothing needs clearing between calls.
// Calls that overlap on one GPU share the array and race.
#include <mutex>

#include "nyl_ctx.h"

// first-stage partials of the reductions (one double per block)
static __device__ __align__(256) double scratch[NYL_RED_BLOCKS];

namespace nyl {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_scratch[kMaxDevices];   // device address of `scratch` on each GPU

// The device address of this GPU's copy of `scratch`.  The array's name in
// host code is only a host-side stand-in, so the address is looked up once per
// device (with that device current) and kept.
int scratch_ptr(double **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return set_error(NYL_E_HIP, "hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices)
        return set_error(NYL_E_HIP, "device %d: at most %d devices are supported", dev,
                         kMaxDevices);
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(NYL_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                             hipGetErrorString(e));
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    return NYL_OK;
}

template <typename T>
int reduce_fields(const Geo &g, const nyl_desc &d, const void *b, const void *u, const void *v,
                  const void *w, int what, double *part, double *out, hipStream_t s) {
    const T *tb = static_cast<const T *>(b), *tu = static_cast<const T *>(u),
            *tv = static_cast<const T *>(v), *tw = static_cast<const T *>(w);
    hipError_t e;
    switch (what) {
    case NYL_RED_MAXU2:
        e = launch_maxu2<T>(g, tu, tv, tw, d.idx2, d.idy2, d.idz2, part, out, s);
        break;
    case NYL_RED_KE:
        e = launch_kesum<T>(g, tu, tv, tw, d.idx2, d.idy2, d.idz2, part, out, s);
        break;
    case NYL_RED_BSUM:
        e = launch_sum<T>(g, tb, part, out, s);
        break;
    default:
        e = launch_nonfinite<T>(g, tb, tu, tv, tw, part, out, s);
        break;
    }
    if (e != hipSuccess) return set_error(NYL_E_HIP, "reduction launch: %s", hipGetErrorString(e));
    return NYL_OK;
}

}  // namespace
}  // namespace nyl

using namespace nyl;

extern "C" int nyl_reduce_noalloc(const nyl_desc *d, int32_t dtype, const void *b,
                                  const void *u, const void *v, const void *w, int32_t what,
                                  double *dev_out, void *stream) {
    if (!d || !dev_out) return set_error(NYL_E_ARG, "null argument");
    if (dtype != NYL_DT_F64 && dtype != NYL_DT_F32)
        return set_error(NYL_E_ARG, "unknown dtype %d", dtype);
    if (what < NYL_RED_MAXU2 || what > NYL_RED_NONFINITE)
        return set_error(NYL_E_ARG, "unknown reduction %d", what);
    const bool need_b = what == NYL_RED_BSUM || what == NYL_RED_NONFINITE;
    const bool need_u = what != NYL_RED_BSUM;
    if ((need_b && !b) || (need_u && (!u || !v || !w)))
        return set_error(NYL_E_ARG, "null field for reduction %d", what);
    if (d->nx < 1 || d->ny < 1 || d->nz < 1)
        return set_error(NYL_E_SIZE, "grid %dx%dx%d: every axis needs at least one cell", d->nx,
                         d->ny, d->nz);
    // the same field limit as nyl_ctx_create: below 4 GiB per field
    const int64_t esz = dtype == NYL_DT_F64 ? 8 : 4;
    const int64_t cells = (int64_t)d->nx * d->ny * d->nz;
    if (cells * esz >= ((int64_t)1 << 32))
        return set_error(NYL_E_SIZE, "grid of %lld cells: a field must stay below 4 GiB",
                         (long long)cells);
    double *part = nullptr;
    int rc = scratch_ptr(&part);
    if (rc) return rc;
    Geo g;
    g.nx = d->nx; g.ny = d->ny; g.nz = d->nz;
    g.px = d->periodic_x; g.py = d->periodic_y; g.pz = d->periodic_z;
    g.sy = d->nx;
    g.sz = (int64_t)d->nx * d->ny;
    hipStream_t s = static_cast<hipStream_t>(stream);
    return dtype == NYL_DT_F64 ? reduce_fields<double>(g, *d, b, u, v, w, what, part, dev_out, s)
                               : reduce_fields<float>(g, *d, b, u, v, w, what, part, dev_out, s);
}
