// nyl_coarse.hip -- direct solve of a coarsest multigrid level that cannot be
// halved further.
//
// The reference's hierarchy halves x and y until one of them is 2
// (core/mgfor/mg_setup.f90:262-303) and its restriction pairs cells
// (core/mgfor/basicoperators.f90:32-60), so a level with an odd size loses
// cells (and on one rank a size 3 never reaches 2).  Here the hierarchy stops
// at the first level with an odd size and that level is solved exactly.
//
// The level's operator is mgfor's 7-point operator on a box without land:
// A x = sum of the in-domain neighbours - diag * x, diag = number of in-domain
// neighbours.  It is the negated Kronecker sum of one 1-D operator per axis:
// the Neumann second difference on a wall axis (eigenvectors of the DCT-II,
// mu_k = 2 - 2 cos(pi k / n)) and the periodic one on a periodic axis (cos /
// sin Fourier modes, mu = 2 - 2 cos(2 pi m / n)).  With E_a the orthonormal
// eigenvector matrix of axis a,
//     x = (Ez^T Ey^T Ex^T) diag(-1 / (mu_x + mu_y + mu_z)) (Ez Ey Ex) b,
// and the constant mode (the null space of the all-Neumann / periodic
// operator) is set to zero.  Three launches: the x and y transforms of each
// z plane, the z transform + scaling + inverse z transform of each y row,
// the inverse y and x transforms of each z plane; every pass holds its plane
// and its y or z matrix in LDS (the x matrix is read from the table in
// memory).  Sizes are limited to CS_MAXX cells along x and CS_MAX along y
// and z (a lock-exchange box n x n/2 x n/2 stops at 2m x m x m, m odd).
#include <math.h>

#include "nyl_internal.h"

namespace nyl {

#define CS_T 256

// one pass over plane / row `p` of a level (PASS 0: forward x, y; 1: z with
// scaling; 2: inverse y, x)
template <int PASS>
__device__ __forceinline__ void cs_pass(const CoarseArgs &a, int p) {
    __shared__ double P[CS_MAXX * CS_MAX];
    __shared__ double T[CS_MAXX * CS_MAX];
    __shared__ double E1[CS_MAX * CS_MAX];
    const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
    const int tid = threadIdx.x;
    const double *ex = a.tab, *ey = a.tab + CS_MAXX * CS_MAXX;
    const double *ez = ey + CS_MAX * CS_MAX;
    const double *mux = ez + CS_MAX * CS_MAX, *muy = mux + CS_MAXX, *muz = muy + CS_MAX;
    double *E2 = E1;   // the y matrix in the plane passes
    if (PASS != 1) {
        // plane z = p: P[y*nx + i]
        const int64_t base = (int64_t)p * a.g.sz;
        const double *src = PASS == 0 ? a.b : a.s;
        for (int t = tid; t < nx * ny; t += CS_T) P[t] = src[base + t];
        for (int t = tid; t < ny * ny; t += CS_T) E2[t] = ey[t];
        __syncthreads();
        if (PASS == 0) {
            // T[y*nx + k] = sum_i Ex[k][i] P[y][i]
            for (int t = tid; t < nx * ny; t += CS_T) {
                const int y = t / nx, k = t - y * nx;
                double acc = 0.0;
                for (int i = 0; i < nx; i++) acc = fma(ex[k * nx + i], P[y * nx + i], acc);
                T[t] = acc;
            }
            __syncthreads();
            // s[kb*nx + k] = sum_y Ey[kb][y] T[y][k]
            for (int t = tid; t < nx * ny; t += CS_T) {
                const int kb = t / nx, k = t - kb * nx;
                double acc = 0.0;
                for (int y = 0; y < ny; y++) acc = fma(E2[kb * ny + y], T[y * nx + k], acc);
                a.s[base + t] = acc;
            }
        } else {
            // T[y*nx + k] = sum_kb Ey[kb][y] P[kb][k]
            for (int t = tid; t < nx * ny; t += CS_T) {
                const int y = t / nx, k = t - y * nx;
                double acc = 0.0;
                for (int kb = 0; kb < ny; kb++) acc = fma(E2[kb * ny + y], P[kb * nx + k], acc);
                T[t] = acc;
            }
            __syncthreads();
            // x[y*nx + i] = sum_k Ex[k][i] T[y][k]
            for (int t = tid; t < nx * ny; t += CS_T) {
                const int y = t / nx, i = t - y * nx;
                double acc = 0.0;
                for (int k = 0; k < nx; k++) acc = fma(ex[k * nx + i], T[y * nx + k], acc);
                a.x[base + t] = acc;
            }
        }
        return;
    }
    // row kb = p of every z plane: P[z*nx + k] = s[z][kb][k]
    const int kb = p;
    for (int t = tid; t < nx * nz; t += CS_T) {
        const int z = t / nx, k = t - z * nx;
        P[t] = a.s[(int64_t)z * a.g.sz + (int64_t)kb * nx + k];
    }
    for (int t = tid; t < nz * nz; t += CS_T) E1[t] = ez[t];
    __syncthreads();
    // T[c*nx + k] = -sum_z Ez[c][z] P[z][k] / (mu_x[k] + mu_y[kb] + mu_z[c])
    for (int t = tid; t < nx * nz; t += CS_T) {
        const int cz = t / nx, k = t - cz * nx;
        double acc = 0.0;
        for (int z = 0; z < nz; z++) acc = fma(E1[cz * nz + z], P[z * nx + k], acc);
        const double mu = mux[k] + muy[kb] + muz[cz];
        T[t] = mu > 0.0 ? -acc / mu : 0.0;
    }
    __syncthreads();
    // s[z][kb][k] = sum_c Ez[c][z] T[c][k]
    for (int t = tid; t < nx * nz; t += CS_T) {
        const int z = t / nx, k = t - z * nx;
        double acc = 0.0;
        for (int cz = 0; cz < nz; cz++) acc = fma(E1[cz * nz + z], T[cz * nx + k], acc);
        a.s[(int64_t)z * a.g.sz + (int64_t)kb * nx + k] = acc;
    }
}

template <int PASS>
__global__ __launch_bounds__(CS_T) void k_cs(CoarseArgs a) {
    cs_pass<PASS>(a, blockIdx.x);
}

bool coarse_fits(const Geo &g) {
    return g.nx >= 1 && g.ny >= 1 && g.nz >= 1 && g.nx <= CS_MAXX && g.ny <= CS_MAX &&
           g.nz <= CS_MAX;
}

hipError_t launch_coarse_solve(const CoarseArgs &a, hipStream_t s) {
    NYL_LAUNCH((k_cs<0>), dim3(a.g.nz), dim3(CS_T), 0, s, a);
    NYL_LAUNCH((k_cs<1>), dim3(a.g.ny), dim3(CS_T), 0, s, a);
    NYL_LAUNCH((k_cs<2>), dim3(a.g.nz), dim3(CS_T), 0, s, a);
    return hipGetLastError();
}

// batched: items own h.nblk = nz (passes 0, 2) or ny (pass 1) blocks
template <int PASS>
__global__ __launch_bounds__(CS_T) void k_b_cs(const CoarseItem *__restrict__ items, int n,
                                               const int *__restrict__ active) {
    const CoarseItem &it = find_item(items, n, blockIdx.x);
    if (active && !active[it.h.member]) return;
    cs_pass<PASS>(it.a, (int)((int64_t)blockIdx.x - it.h.start));
}

hipError_t launch_b_coarse(const CoarseItem *items, int n, int64_t nblk, int pass,
                           const int *active, hipStream_t s) {
    switch (pass) {
    case 0: NYL_LAUNCH((k_b_cs<0>), dim3((unsigned)nblk), dim3(CS_T), 0, s, items, n, active); break;
    case 1: NYL_LAUNCH((k_b_cs<1>), dim3((unsigned)nblk), dim3(CS_T), 0, s, items, n, active); break;
    default: NYL_LAUNCH((k_b_cs<2>), dim3((unsigned)nblk), dim3(CS_T), 0, s, items, n, active); break;
    }
    return hipGetLastError();
}

// orthonormal eigenvectors (rows of E, E[k*n + i]) and eigenvalues mu of the
// 1-D operator of one axis
static void axis_tables(int n, int periodic, double *E, double *mu) {
    const double pi = 3.14159265358979323846;
    for (int k = 0; k < n; k++) {
        if (!periodic) {
            mu[k] = 2.0 - 2.0 * cos(pi * k / n);
            const double c = k == 0 ? sqrt(1.0 / n) : sqrt(2.0 / n);
            for (int i = 0; i < n; i++) E[k * n + i] = c * cos(pi * k * (i + 0.5) / n);
            continue;
        }
        // periodic: k = 0 constant; k = 2m-1, 2m: cos, sin of wave number m;
        // n even: k = n-1 the alternating mode (m = n/2)
        const int m = (k + 1) / 2;
        mu[k] = 2.0 - 2.0 * cos(2.0 * pi * m / n);
        for (int i = 0; i < n; i++) {
            double v;
            if (k == 0) v = sqrt(1.0 / n);
            else if (2 * m == n) v = ((i & 1) ? -1.0 : 1.0) * sqrt(1.0 / n);
            else if (k & 1) v = sqrt(2.0 / n) * cos(2.0 * pi * m * i / n);
            else v = sqrt(2.0 / n) * sin(2.0 * pi * m * i / n);
            E[k * n + i] = v;
        }
    }
}

// table layout: Ex (CS_MAXX^2), Ey, Ez (CS_MAX^2 each), mu_x (CS_MAXX),
// mu_y, mu_z (CS_MAX each)
void coarse_tables(const Geo &g, double *tab) {
    for (int q = 0; q < coarse_table_doubles(); q++) tab[q] = 0.0;
    double *ey = tab + CS_MAXX * CS_MAXX, *ez = ey + CS_MAX * CS_MAX;
    double *mux = ez + CS_MAX * CS_MAX, *muy = mux + CS_MAXX, *muz = muy + CS_MAX;
    axis_tables(g.nx, g.px, tab, mux);
    axis_tables(g.ny, g.py, ey, muy);
    axis_tables(g.nz, g.pz, ez, muz);
}

int coarse_table_doubles() {
    return CS_MAXX * CS_MAXX + 2 * CS_MAX * CS_MAX + CS_MAXX + 2 * CS_MAX;
}

}  // namespace nyl
