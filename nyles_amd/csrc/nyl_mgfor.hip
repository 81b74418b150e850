// nyl_mgfor.hip -- the ctypes ABI of the reference's mgfor multigrid library
// over the device multigrid, so that the reference's MG driver
// (core/mgfordriver.py:5-105 with the generated mgmod.py,
// core/build.py:262-314) solves on the GPU unchanged.
//
// The driver finds each function by the substring "MOD_" + name among the
// library's symbols (core/build.py:107-121), calls it with every argument by
// reference (Python int and bool -> int32*, float -> double*, ndarray ->
// double*, core/build.py:44-82) and reads a double result (a subroutine's
// is ignored).  get_ptrmg returns an opaque handle
// (core/mgfor/mg_setup.f90:318-437).  Arrays cross as numpy arrays of shape
// (nz+2nh, ny+2nh, nx+2nh) passed with dims (n3, n2, n1)
// (core/mgfor/pytools.f90:17-60): the interior sits at offset nh on every
// axis (the z halo is inside the array, core/mgfor/mg_setup.f90:239).
//
// Each handle owns a model context in FP64 (the mgfor parameters: omega 0.9,
// tol 1e-6 on the squared ratio, 20 V-cycles, nh = 3,
// core/mgfor/mg_types.f90:15-26), its device workspace (allocated here, as
// mgfor allocates its own arrays) and a stream.  Supported: level 1,
// ivar 1 (x), 2 (b), 4 (y).  Unlike the reference there is no `stop` on a
// bad argument: the call prints the reason and returns.
#include <stdio.h>
#include <string.h>

#include <vector>

#include "nyl_ctx.h"

namespace {

constexpr int kNh = 3;   // MG_Param%nh (core/mgfor/mg_types.f90:19)

struct MgHandle {
    int64_t magic;           // the driver holds a POINTER(c_int64) to the handle
    nyl_ctx *ctx;
    void *ws;
    hipStream_t stream;
    int nx, ny, nz, px, py, pz, topology;
    int calls;
};
constexpr int64_t kMagic = 0x6d67666f72L;   // "mgfor"

MgHandle *handle(void *mg) {
    MgHandle *h = static_cast<MgHandle *>(mg);
    if (!h || h->magic != kMagic) {
        fprintf(stderr, "mgfor shim: invalid handle\n");
        return nullptr;
    }
    return h;
}

bool level_one(const int32_t *lev) {
    if (!lev || *lev != 1) {
        fprintf(stderr, "mgfor shim: only level 1 is exchanged (got %d)\n", lev ? *lev : -1);
        return false;
    }
    return true;
}

int field_of(int ivar) {
    switch (ivar) {
    case 1: return NYL_F_MGX;
    case 2: return NYL_F_MGB;
    case 4: return NYL_F_MGY;
    default: return -1;
    }
}

}  // namespace

extern "C" {

// core/mgfor/mg_setup.f90:318 get_ptrmg(npx, npy, nx, ny, nz, vertices, short,
// is3d, topology); topology: 1 closed, 2 xperio, 3 yperio, 4 zperio,
// 5 xyperio, 6 xyzperio (core/mgfor/mg_enums.f90:6)
void *__mg_setup_MOD_get_ptrmg(const int32_t *npx, const int32_t *npy, const int32_t *nx,
                               const int32_t *ny, const int32_t *nz, const int32_t *vertices,
                               const int32_t *shortp, const int32_t *is3d,
                               const int32_t *topology) {
    (void)vertices; (void)shortp;
    if (!npx || !npy || !nx || !ny || !nz) return nullptr;
    if (*npx * *npy != 1) {
        fprintf(stderr, "wrong number of mpi cores, should be : %d (one GPU: 1)\n", *npx * *npy);
        return nullptr;
    }
    if (is3d && !*is3d) {
        fprintf(stderr, "mgfor shim: only the 3-D multigrid is available\n");
        return nullptr;
    }
    const int topo = topology ? *topology : 1;
    static const int per[7][3] = {{0, 0, 0}, {0, 0, 0}, {1, 0, 0}, {0, 1, 0},
                                  {0, 0, 1}, {1, 1, 0}, {1, 1, 1}};
    if (topo < 1 || topo > 6) {
        fprintf(stderr, "mgfor shim: unknown topology %d\n", topo);
        return nullptr;
    }
    nyl_desc d;
    memset(&d, 0, sizeof(d));
    d.nx = *nx; d.ny = *ny; d.nz = *nz;
    d.periodic_x = per[topo][0]; d.periodic_y = per[topo][1]; d.periodic_z = per[topo][2];
    d.dx = d.dy = d.dz = 1.0;
    d.idx2 = d.idy2 = d.idz2 = 1.0;
    d.dx2 = 1.0;
    d.precision = NYL_PREC_FP64;
    d.mg_maxite = 20;
    d.mg_tol = 1e-6;
    d.mg_omega = 0.9;
    nyl_ctx *ctx = nullptr;
    if (nyl_ctx_create(&d, &ctx) != NYL_OK) {
        fprintf(stderr, "mgfor shim: %s\n", nyl_last_error());
        return nullptr;
    }
    MgHandle *h = new MgHandle();
    memset(h, 0, sizeof(*h));
    h->magic = kMagic;
    h->ctx = ctx;
    h->nx = *nx; h->ny = *ny; h->nz = *nz;
    h->px = d.periodic_x; h->py = d.periodic_y; h->pz = d.periodic_z;
    h->topology = topo;
    uint64_t bytes = 0;
    hipError_t e = hipSuccess;
    if (nyl_workspace_bytes(ctx, &bytes) != NYL_OK || (e = hipMalloc(&h->ws, bytes)) != hipSuccess ||
        (e = hipStreamCreate(&h->stream)) != hipSuccess ||
        nyl_set_stream(ctx, h->stream) != NYL_OK || nyl_bind_workspace(ctx, h->ws, bytes) != NYL_OK) {
        fprintf(stderr, "mgfor shim: %s %s\n", e != hipSuccess ? hipGetErrorString(e) : "",
                nyl_last_error());
        if (h->ws) (void)hipFree(h->ws);
        if (h->stream) (void)hipStreamDestroy(h->stream);
        nyl_ctx_destroy(ctx);
        delete h;
        return nullptr;
    }
    return h;
}

// core/mgfor/mg_setup.f90:65-80
double __mg_setup_MOD_print_mginfos(void *mg) {
    MgHandle *h = handle(mg);
    if (!h) return 0.0;
    int32_t n = 0;
    nyl_mg_levels(h->ctx, &n);
    printf(" MG has %2d levels\n", n);
    printf(" parameters= %d %d %d %d %d %g %g (device multigrid, topology %d)\n", 1, 1, 20,
           kNh, h->ctx->d.mg_maxite, h->ctx->d.mg_tol, h->ctx->d.mg_omega, h->topology);
    for (int l = 0; l < n; l++) {
        int32_t a, b, c;
        nyl_mg_level_shape(h->ctx, l, &a, &b, &c);
        printf(" %3d x%3d x%3d |   1 x  1\n", a, b, c + 2 * kNh);
    }
    fflush(stdout);
    return 0.0;
}

// core/mgfor/pytools.f90:8-15: (nz+2nh, ny+2nh, nx+2nh) of a level
double __pytools_MOD_get_pyshape(void *mg, const int32_t *lev, int32_t *gshape) {
    MgHandle *h = handle(mg);
    if (!h || !lev || !gshape) return 0.0;
    int32_t a, b, c;
    if (nyl_mg_level_shape(h->ctx, *lev - 1, &a, &b, &c) != NYL_OK) {
        fprintf(stderr, "mgfor shim: %s\n", nyl_last_error());
        return 0.0;
    }
    gshape[0] = c + 2 * kNh;
    gshape[1] = b + 2 * kNh;
    gshape[2] = a + 2 * kNh;
    return 0.0;
}

// core/mgfor/pytools.f90:46-60: the level array from x(n1, n2, n3)
double __pytools_MOD_set_pyarray(void *mg, const int32_t *lev, const int32_t *ivar,
                                 const int32_t *n1, const int32_t *n2, const int32_t *n3,
                                 const double *x) {
    MgHandle *h = handle(mg);
    if (!h || !level_one(lev) || !ivar || !x) return 0.0;
    const int fid = field_of(*ivar);
    if (fid < 0 || *n1 != h->nx + 2 * kNh || *n2 != h->ny + 2 * kNh || *n3 != h->nz + 2 * kNh) {
        fprintf(stderr, "mgfor shim: set_pyarray ivar %d of %dx%dx%d not supported\n", *ivar, *n1,
                *n2, *n3);
        return 0.0;
    }
    uint64_t off = 0;
    int64_t nelem = 0;
    int32_t dt = 0;
    nyl_field_offset(h->ctx, fid, &off, &nelem, &dt);
    std::vector<double> in((size_t)nelem);
    const int sx = *n1, sy = *n1 * *n2;
    for (int k = 0; k < h->nz; k++)
        for (int j = 0; j < h->ny; j++)
            memcpy(&in[((size_t)k * h->ny + j) * h->nx],
                   x + (size_t)(k + kNh) * sy + (size_t)(j + kNh) * sx + kNh,
                   sizeof(double) * h->nx);
    hipError_t e = hipMemcpyAsync(h->ctx->ws + off, in.data(), in.size() * sizeof(double),
                                  hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) fprintf(stderr, "mgfor shim: %s\n", hipGetErrorString(e));
    return 0.0;
}

// core/mgfor/pytools.f90:17-44: x(n1, n2, n3) from the level array; the
// halo holds the periodic images on periodic axes and 0 on wall axes
double __pytools_MOD_get_pyarray(void *mg, const int32_t *lev, const int32_t *ivar,
                                 const int32_t *n1, const int32_t *n2, const int32_t *n3,
                                 double *x) {
    MgHandle *h = handle(mg);
    if (!h || !level_one(lev) || !ivar || !x) return 0.0;
    const int fid = field_of(*ivar);
    if (fid < 0 || *n1 != h->nx + 2 * kNh || *n2 != h->ny + 2 * kNh || *n3 != h->nz + 2 * kNh) {
        fprintf(stderr, "mgfor shim: get_pyarray ivar %d of %dx%dx%d not supported\n", *ivar, *n1,
                *n2, *n3);
        return 0.0;
    }
    uint64_t off = 0;
    int64_t nelem = 0;
    int32_t dt = 0;
    nyl_field_offset(h->ctx, fid, &off, &nelem, &dt);
    std::vector<double> out((size_t)nelem);
    hipError_t e = hipMemcpyAsync(out.data(), h->ctx->ws + off, out.size() * sizeof(double),
                                  hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        fprintf(stderr, "mgfor shim: %s\n", hipGetErrorString(e));
        return 0.0;
    }
    const int N1 = *n1, N2 = *n2, N3 = *n3;
    auto src = [&](int a, int n, int p) -> int {   // array index -> interior index or -1
        int m = a - kNh;
        if (m >= 0 && m < n) return m;
        if (!p) return -1;
        return m < 0 ? m + n : m - n;
    };
    for (int k = 0; k < N3; k++) {
        const int kk = src(k, h->nz, h->pz);
        for (int j = 0; j < N2; j++) {
            const int jj = src(j, h->ny, h->py);
            double *row = x + (size_t)k * N1 * N2 + (size_t)j * N1;
            for (int i = 0; i < N1; i++) {
                const int ii = src(i, h->nx, h->px);
                row[i] = (kk < 0 || jj < 0 || ii < 0)
                             ? 0.0
                             : out[((size_t)kk * h->ny + jj) * h->nx + ii];
            }
        }
    }
    return 0.0;
}

// core/mgfor/solvers.f90:8-33: solve with b, warm-started from x
double __solvers_MOD_solve(void *mg) {
    MgHandle *h = handle(mg);
    if (!h) return 0.0;
    nyl_mg_stats st;
    if (nyl_mg_solve(h->ctx, &st) != NYL_OK) fprintf(stderr, "mgfor shim: %s\n", nyl_last_error());
    h->calls++;
    return 0.0;
}

// core/mgfor/mg_setup.f90:309-315: the MG_Param record
// (core/mgfor/mg_types.f90:15-26, bind(C), REAL = double)
struct MgParam {
    int32_t npre, npost, nexact, nh, maxite;
    double tol, omega;
    int32_t debug, vertices, is2d;
};
double __mg_setup_MOD_get_param(void *mg, MgParam *param) {
    MgHandle *h = handle(mg);
    if (!h || !param) return 0.0;
    param->npre = 1;
    param->npost = 1;
    param->nexact = 20;
    param->nh = kNh;
    param->maxite = h->ctx->d.mg_maxite;
    param->tol = h->ctx->d.mg_tol;
    param->omega = h->ctx->d.mg_omega;
    param->debug = 0;
    param->vertices = 0;
    param->is2d = 0;
    return 0.0;
}

}  // extern "C"
