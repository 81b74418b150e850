// Stage kernel: the 2.5-D march (k_stage25) against the per-cell kernel
// (k_stage), bit for bit, for every mode, every periodicity, with and
// without viscosity, on grids with partial tiles in x and y, z ranges split
// over several blocks and the smallest wall axes (8 cells).
#define NYL_STAGE_SEQ 0
#include "../../nyles_amd/csrc/nyl_stencil.hip"
#include "emu.h"
using namespace nyl;

template <typename T, int PX, int PY, int PZ>
void k_stage_dispatch(const Geo &g, const StageArgs<T> &a, int mode) {
    switch (mode) {
    case 0: k_stage<T, 0, PX, PY, PZ>(g, a); break;
    case 1: k_stage<T, 1, PX, PY, PZ>(g, a); break;
    case 2: k_stage<T, 2, PX, PY, PZ>(g, a); break;
    case 3: k_stage<T, 3, PX, PY, PZ>(g, a); break;
    case 4: k_stage<T, 4, PX, PY, PZ>(g, a); break;
    case 5: k_stage<T, 5, PX, PY, PZ>(g, a); break;
    default: k_stage<T, 6, PX, PY, PZ>(g, a); break;
    }
}

template <typename T, int PX, int PY, int PZ>
void k_stage25_dispatch(const Geo &g, const StageArgs<T> &a, int mode, int kz) {
    switch (mode) {
    case 0: k_stage25<T, 0, PX, PY, PZ>(g, a, kz); break;
    case 1: k_stage25<T, 1, PX, PY, PZ>(g, a, kz); break;
    case 2: k_stage25<T, 2, PX, PY, PZ>(g, a, kz); break;
    case 3: k_stage25<T, 3, PX, PY, PZ>(g, a, kz); break;
    case 4: k_stage25<T, 4, PX, PY, PZ>(g, a, kz); break;
    case 5: k_stage25<T, 5, PX, PY, PZ>(g, a, kz); break;
    default: k_stage25<T, 6, PX, PY, PZ>(g, a, kz); break;
    }
}

template <typename T>
int check(int nx, int ny, int nz, int px, int py, int pz, int visc) {
    Geo g = mk(nx, ny, nz, px, py, pz);
    const size_t n = (size_t)nx * ny * nz;
    std::mt19937_64 rng(3 * nx + 5 * ny + 7 * nz + px + 2 * py + 4 * pz + 8 * visc);
    // state, vorticity, ke, tendency slots, LFAM3 'before': inputs
    std::vector<std::vector<T>> in(19);
    for (auto &v : in) v = rnd<T>(n, rng);
    // exact zeros in u, v, w: the u > 0 / u <= 0 split of the upwinding
    for (int f = 1; f <= 3; f++)
        for (size_t q = 0; q < n; q += 7) in[f][q] = T(0);
    // the wall faces carry zero velocity (core/fortran_vortex_force.f90:42)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const size_t c = ((size_t)z * ny + y) * nx + x;
                if (!px && x == nx - 1) in[1][c] = T(0);
                if (!py && y == ny - 1) in[2][c] = T(0);
                if (!pz && z == nz - 1) in[3][c] = T(0);
            }
    int bad = 0;
    for (int mode = 0; mode <= 6; mode++) {
        // outputs: next state (4), tendencies (4), 'before' state (4), each
        // run of the kernel on its own copies
        std::vector<std::vector<T>> o1(12), o2;
        for (int f = 0; f < 12; f++) o1[f] = f < 4 ? in[f] : (f < 8 ? std::vector<T>(n, T(5)) : in[15 + (f - 8)]);
        if (mode == 5)
            for (int f = 0; f < 4; f++) o1[f] = in[11 + f];   // the saved 'now' state
        o2 = o1;
        auto args = [&](std::vector<std::vector<T>> &o) {
            StageArgs<T> a;
            memset(&a, 0, sizeof(a));
            a.b = in[0].data(); a.u = in[1].data(); a.v = in[2].data(); a.w = in[3].data();
            a.bn = o[0].data(); a.un = o[1].data(); a.vn = o[2].data(); a.wn = o[3].data();
            a.ox = in[4].data(); a.oy = in[5].data(); a.oz = in[6].data(); a.ke = in[7].data();
            a.d0b = in[8].data(); a.d0u = in[9].data(); a.d0v = in[10].data(); a.d0w = in[14].data();
            a.d1b = in[11].data(); a.d1u = in[12].data(); a.d1v = in[13].data(); a.d1w = in[18].data();
            a.tb = o[4].data(); a.tu = o[5].data(); a.tv = o[6].data(); a.tw = o[7].data();
            a.sbb = o[8].data(); a.sbu = o[9].data(); a.sbv = o[10].data(); a.sbw = o[11].data();
            a.idx2 = 1.7; a.idy2 = 2.3; a.idz2 = 0.9; a.dz = 0.37;
            a.c = 0.013; a.c2 = 0.026;
            a.visc = visc;
            a.vcx = 0.011; a.vcy = 0.013; a.vcz = 0.007;
            return a;
        };
        StageArgs<T> a1 = args(o1), a2 = args(o2);
        run_blocks(cell_grid(g), dim3(NYL_BX, NYL_BY, 1), [&]() {
            switch (px + 2 * py + 4 * pz) {
            case 0: k_stage_dispatch<T, 0, 0, 0>(g, a1, mode); break;
            case 1: k_stage_dispatch<T, 1, 0, 0>(g, a1, mode); break;
            case 2: k_stage_dispatch<T, 0, 1, 0>(g, a1, mode); break;
            case 3: k_stage_dispatch<T, 1, 1, 0>(g, a1, mode); break;
            case 4: k_stage_dispatch<T, 0, 0, 1>(g, a1, mode); break;
            case 5: k_stage_dispatch<T, 1, 0, 1>(g, a1, mode); break;
            case 6: k_stage_dispatch<T, 0, 1, 1>(g, a1, mode); break;
            default: k_stage_dispatch<T, 1, 1, 1>(g, a1, mode); break;
            }
        }, false);
        // block sizes along z: planes split with a remainder, and (for the
        // fully periodic and the closed box) one block per column
        for (int kz : {3, nz}) {
            if (kz == nz && px + py + pz != 0 && px + py + pz != 3) continue;
            std::vector<std::vector<T>> o3 = o2;
            StageArgs<T> a3 = args(o3);
            dim3 grd((nx + S25_TX - 1) / S25_TX, (ny + S25_TY - 1) / S25_TY, (nz + kz - 1) / kz);
            run_blocks(grd, dim3(S25_NT, 1, 1), [&]() {
                switch (px + 2 * py + 4 * pz) {
                case 0: k_stage25_dispatch<T, 0, 0, 0>(g, a3, mode, kz); break;
                case 1: k_stage25_dispatch<T, 1, 0, 0>(g, a3, mode, kz); break;
                case 2: k_stage25_dispatch<T, 0, 1, 0>(g, a3, mode, kz); break;
                case 3: k_stage25_dispatch<T, 1, 1, 0>(g, a3, mode, kz); break;
                case 4: k_stage25_dispatch<T, 0, 0, 1>(g, a3, mode, kz); break;
                case 5: k_stage25_dispatch<T, 1, 0, 1>(g, a3, mode, kz); break;
                case 6: k_stage25_dispatch<T, 0, 1, 1>(g, a3, mode, kz); break;
                default: k_stage25_dispatch<T, 1, 1, 1>(g, a3, mode, kz); break;
                }
            }, true);
            int b = 0;
            for (int f = 0; f < 12; f++) b += ndiff(o1[f], o3[f]);
            if (b) printf("  mode %d kz %d: %d mismatches\n", mode, kz, b);
            bad += b;
        }
    }
    printf("stage25<%s> %dx%dx%d p=%d%d%d visc=%d: %d mismatches\n", sizeof(T) == 8 ? "f64" : "f32",
           nx, ny, nz, px, py, pz, visc, bad);
    return bad != 0;
}

int main() {
    int fails = 0;
    // 72 x 10: partial tiles in x and y; 8 planes: the smallest wall axis
    for (int p = 0; p < 8; p++) {
        fails += check<double>(72, 10, 8, p & 1, (p >> 1) & 1, (p >> 2) & 1, p == 0 || p == 5);
        fails += check<float>(72, 10, 8, p & 1, (p >> 1) & 1, (p >> 2) & 1, 0);
    }
    // one full tile in x, two in y, 9 planes (remainder of the z split)
    fails += check<double>(64, 8, 9, 0, 0, 0, 1);
    fails += check<double>(64, 8, 9, 1, 1, 1, 0);
    printf(fails ? "FAIL %d\n" : "ALL OK\n", fails);
    return fails != 0;
}
