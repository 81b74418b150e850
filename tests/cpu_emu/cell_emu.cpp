// Per-cell diagnostics: the four-cells-per-thread kernels against the
// one-cell ones, bit for bit, on small grids of every periodicity.
#define NYL_STAGE_SEQ 0
#include "../../nyles_amd/csrc/nyl_stencil.hip"
#include "emu.h"
using namespace nyl;

template <typename T, typename TO>
int check(int nx, int ny, int nz, int px, int py, int pz) {
    Geo g = mk(nx, ny, nz, px, py, pz);
    const size_t n = (size_t)nx * ny * nz;
    std::mt19937_64 rng(nx + 7 * ny + 13 * nz + px + 2 * py + 4 * pz);
    auto u = rnd<T>(n, rng), v = rnd<T>(n, rng), w = rnd<T>(n, rng);
    int bad = 0;
    {
        std::vector<TO> o1(n, TO(7)), o2(n, TO(7));
        DivArgs<T, TO> a;
        a.u = u.data(); a.v = v.data(); a.w = w.data();
        a.idx2 = 1.7; a.idy2 = 2.3; a.idz2 = 0.9;
        a.out = o1.data();
        run_blocks(cell_grid_kz(g), dim3(NYL_BX, NYL_BY, 1), [&]() { k_div<T, TO>(g, a); }, false);
        a.out = o2.data();
        run_blocks(cell_grid4(g), dim3(NYL_BX, NYL_BY, 1), [&]() { k_div4<T, TO>(g, a); }, false);
        bad += ndiff(o1, o2);
    }
    {
        auto x = rnd<TO>(n, rng);
        auto u1 = u, v1 = v, w1 = w, u2 = u, v2 = v, w2 = w;
        CorrectArgs<T, TO> a;
        a.x = x.data(); a.dx2 = 0.37;
        a.u = u1.data(); a.v = v1.data(); a.w = w1.data();
        run_blocks(cell_grid_kz(g), dim3(NYL_BX, NYL_BY, 1), [&]() { k_correct<T, TO>(g, a); }, false);
        a.u = u2.data(); a.v = v2.data(); a.w = w2.data();
        run_blocks(cell_grid4(g), dim3(NYL_BX, NYL_BY, 1), [&]() { k_correct4<T, TO>(g, a); }, false);
        bad += ndiff(u1, u2) + ndiff(v1, v2) + ndiff(w1, w2);
    }
    for (int what : {7, 6, 4, 2, 1})
        for (double fp : {0.0, 0.25}) {
            std::vector<std::vector<T>> o1(7, std::vector<T>(n, T(3))), o2 = o1;
            DiagArgs<T> a;
            a.u = u.data(); a.v = v.data(); a.w = w.data();
            a.idx2 = 1.7; a.idy2 = 2.3; a.idz2 = 0.9; a.fpar = fp;
            auto set = [&](std::vector<std::vector<T>> &o) {
                a.Ux = o[0].data(); a.Uy = o[1].data(); a.Uz = o[2].data();
                a.ox = o[3].data(); a.oy = o[4].data(); a.oz = o[5].data(); a.ke = o[6].data();
            };
            set(o1);
            run_blocks(cell_grid_kz(g), dim3(NYL_BX, NYL_BY, 1), [&]() { k_diag<T>(g, a, what); }, false);
            set(o2);
            run_blocks(cell_grid4(g), dim3(NYL_BX, NYL_BY, 1), [&]() {
                switch (what) {
                case 7: k_diag4<T, 7>(g, a); break;
                case 6: k_diag4<T, 6>(g, a); break;
                case 4: k_diag4<T, 4>(g, a); break;
                case 2: k_diag4<T, 2>(g, a); break;
                default: k_diag4<T, 1>(g, a); break;
                } }, false);
            for (int f = 0; f < 7; f++) bad += ndiff(o1[f], o2[f]);
        }
    printf("cell4<%s,%s> %dx%dx%d p=%d%d%d: %d mismatches\n", sizeof(T) == 8 ? "f64" : "f32",
           sizeof(TO) == 8 ? "f64" : "f32", nx, ny, nz, px, py, pz, bad);
    return bad != 0;
}

int main() {
    int fails = 0;
    const int sizes[][3] = {{4, 4, 4}, {12, 6, 5}, {260, 5, 4}};
    for (auto &s : sizes)
        for (int p = 0; p < 8; p++) {
            fails += check<double, double>(s[0], s[1], s[2], p & 1, (p >> 1) & 1, (p >> 2) & 1);
            fails += check<float, double>(s[0], s[1], s[2], p & 1, (p >> 1) & 1, (p >> 2) & 1);
        }
    printf(fails ? "FAIL %d\n" : "ALL OK\n", fails);
    return fails != 0;
}
