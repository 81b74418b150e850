// Multigrid kernels: the tiled and fused forms against the per-cell ones,
// bit for bit, on small grids of every periodicity.
#include "../../nyles_amd/csrc/nyl_mg.hip"
#include "emu.h"
using namespace nyl;

// tiled sweep, restriction (with the coarse first sweep), prolongation (in
// place and out of place) and two-sweep pass against the per-cell kernels
template <typename T>
int check_tiled(int nx, int ny, int nz, int px, int py, int pz) {
    Geo gf = mk(nx, ny, nz, px, py, pz), gc = mk(nx / 2, ny / 2, nz / 2, px, py, pz);
    MGTabs tab = mktab();
    const size_t nf = (size_t)nx * ny * nz, nc = nf / 8;
    std::mt19937_64 rng(nx + 3 * ny + 7 * nz + px + 2 * py + 4 * pz);
    auto x = rnd<T>(nf, rng), b = rnd<T>(nf, rng), xc = rnd<T>(nc, rng);
    const double om = 0.9;
    int bad = 0;
    // sweep
    std::vector<T> y1(nf), y2(nf);
    run_blocks(mg_grid(gf), dim3(MG_BX, MG_BY, 1), [&]() {
        k_mg_sweep<T>(gf, x.data(), y1.data(), b.data(), om, tab); }, false);
    const int kzs = sweep25_kz(gf, 1);
    const dim3 gs(nx / SW_TX, ny / SW_TY, nz / kzs);
    run_blocks(gs, dim3(SW_TX, 4, 1), [&]() {
        k_mg_sweep25<T, false, false, false>(gf, x.data(), y2.data(), b.data(), om, tab, kzs,
                                             nullptr, gf, nullptr, nullptr); }, true);
    bad += ndiff(y1, y2);
    // the sweep variants: residual norm of the input (same output), the
    // prolongation of xc folded into the input, the result added to an FP64 x
    {
        std::vector<T> y3(nf);
        std::vector<double> part(nf, 0.0), rn(NYL_RED_BLOCKS, 0.0);
        run_blocks(gs, dim3(SW_TX, 4, 1), [&]() {
            k_mg_sweep25<T, true, false, false>(gf, x.data(), y3.data(), b.data(), om, tab, kzs,
                                                part.data(), gf, nullptr, nullptr); }, true);
        const int nbr = red_blocks((int64_t)nf);
        run_blocks(dim3(nbr), dim3(RED_T), [&]() {
            k_resnorm2<T>(gf, x.data(), b.data(), tab, rn.data()); }, true);
        double s1 = 0, s2 = 0;
        for (double v : part) s1 += v;
        for (double v : rn) s2 += v;
        bad += ndiff(y1, y3) + (std::fabs(s1 - s2) > 1e-9 * s2);
        std::vector<T> xp = x, yp1(nf), yp2(nf);
        run_blocks(mg_grid(gf), dim3(MG_BX, MG_BY, 1), [&]() {
            k_mg_prolong<T, T>(gf, gc, xp.data(), xc.data(), tab); }, false);
        run_blocks(mg_grid(gf), dim3(MG_BX, MG_BY, 1), [&]() {
            k_mg_sweep<T>(gf, xp.data(), yp1.data(), b.data(), om, tab); }, false);
        run_blocks(gs, dim3(SW_TX, 4, 1), [&]() {
            k_mg_sweep25<T, false, true, false>(gf, x.data(), yp2.data(), b.data(), om, tab, kzs,
                                                nullptr, gc, xc.data(), nullptr); }, true);
        bad += ndiff(yp1, yp2);
        std::vector<double> xo1(nf), xo2(nf);
        for (size_t q = 0; q < nf; q++) xo1[q] = xo2[q] = 0.25 * (double)q;
        for (size_t q = 0; q < nf; q++) xo1[q] = xo1[q] + (double)y1[q];
        run_blocks(gs, dim3(SW_TX, 4, 1), [&]() {
            k_mg_sweep25<T, false, false, true>(gf, x.data(), nullptr, b.data(), om, tab, kzs,
                                                nullptr, gf, nullptr, xo2.data()); }, true);
        bad += ndiff(xo1, xo2);
    }
    // restriction with the coarse level's first sweep
    std::vector<T> bc1(nc), bc2(nc), yc1(nc), yc2(nc), xd(nc);
    run_blocks(mg_grid(gc), dim3(MG_BX, MG_BY, 1), [&]() {
        k_mg_resid_restrict<T, T>(gf, gc, x.data(), b.data(), bc1.data(), xd.data(), tab,
                                  yc1.data(), om); }, false);
    const int kzr = sweep25_kz(gf, 2);
    run_blocks(dim3(nx / SW_TX, ny / SW_TY, nz / kzr), dim3(SW_TX, 4, 1), [&]() {
        k_mg_restrict25<T, T>(gf, gc, x.data(), b.data(), bc2.data(), xd.data(), tab, kzr,
                              yc2.data(), om); }, true);
    bad += ndiff(bc1, bc2) + ndiff(yc1, yc2);
    for (int kzq : {2, 4, 6, 8}) {   // every even march length
        if (nz % kzq) continue;
        std::vector<T> bc3(nc), yc3(nc);
        run_blocks(dim3(nx / SW_TX, ny / SW_TY, nz / kzq), dim3(SW_TX, 4, 1), [&]() {
            k_mg_restrict25<T, T>(gf, gc, x.data(), b.data(), bc3.data(), xd.data(), tab, kzq,
                                  yc3.data(), om); }, true);
        bad += ndiff(bc1, bc3) + ndiff(yc1, yc3);
    }
    // prolongation, in place and into another array
    std::vector<T> p1 = x, p2 = x, p3(nf), x0 = x;
    run_blocks(mg_grid(gf), dim3(MG_BX, MG_BY, 1), [&]() {
        k_mg_prolong<T, T>(gf, gc, p1.data(), xc.data(), tab); }, false);
    run_blocks(dim3(nx / SW_TX, ny / SW_TY, nz / kzr), dim3(SW_TX, 4, 1), [&]() {
        k_mg_prolong25<T, T>(gf, gc, p2.data(), xc.data(), tab, kzr, nullptr); }, true);
    std::vector<T> p4 = x;
    run_blocks(dim3(nx / SW_TX, ny / SW_TY, nz / kzr), dim3(SW_TX, 4, 1), [&]() {
        k_mg_prolong25<T, T>(gf, gc, p4.data(), xc.data(), tab, kzr, p3.data()); }, true);
    bad += ndiff(p1, p2) + ndiff(p1, p3) + ndiff(p4, x0);
    for (int kzp : {2, 4, 6, 8}) {   // every march length of the coarse-plane ring
        if (nz % kzp) continue;
        std::vector<T> p5 = x, p6(nf);
        run_blocks(dim3(nx / SW_TX, ny / SW_TY, nz / kzp), dim3(SW_TX, 4, 1), [&]() {
            k_mg_prolong25<T, T>(gf, gc, p5.data(), xc.data(), tab, kzp, nullptr); }, true);
        std::vector<T> x1 = x;
        run_blocks(dim3(nx / SW_TX, ny / SW_TY, nz / kzp), dim3(SW_TX, 4, 1), [&]() {
            k_mg_prolong25<T, T>(gf, gc, x1.data(), xc.data(), tab, kzp, p6.data()); }, true);
        bad += ndiff(p1, p5) + ndiff(p1, p6);
    }
    // two sweeps in one pass
    std::vector<T> z1(nf), z2(nf);
    run_blocks(mg_grid(gf), dim3(MG_BX, MG_BY, 1), [&]() {
        k_mg_sweep<T>(gf, y1.data(), z1.data(), b.data(), om, tab); }, false);
    const int kz2 = nz % 4 == 0 ? 4 : 1;
    run_blocks(dim3(nx / SW_TX, ny / SW_TY, nz / kz2), dim3(SW_TX, 4, 1), [&]() {
        k_mg_sweep2x25<T>(gf, x.data(), z2.data(), b.data(), om, tab, kz2, nullptr); }, true);
    bad += ndiff(z1, z2);
    // ... with the result added to an FP64 x instead of stored
    {
        std::vector<double> xo1(nf), xo2(nf);
        for (size_t q = 0; q < nf; q++) xo1[q] = xo2[q] = 0.5 - 0.125 * (double)q;
        for (size_t q = 0; q < nf; q++) xo1[q] = xo1[q] + (double)z1[q];
        run_blocks(dim3(nx / SW_TX, ny / SW_TY, nz / kz2), dim3(SW_TX, 4, 1), [&]() {
            k_mg_sweep2x25<T>(gf, x.data(), nullptr, b.data(), om, tab, kz2, xo2.data()); },
            true);
        bad += ndiff(xo1, xo2);
    }
    // ... and the pass with register z-windows, for several march lengths
    for (int kzt : {1, 2, 3, 4, 8}) {
        if (nz % kzt) continue;
        std::vector<T> z4(nf);
        run_blocks(dim3(nx / SW_TX, ny / SW_TY, nz / kzt), dim3(SW_TX, 4, 1), [&]() {
            (pz ? k_mg_sweep2t<T, false, 1, false, false> : k_mg_sweep2t<T, false, 0, false, false>)(
                gf, x.data(), z4.data(), b.data(), om, tab, kzt, nullptr, nullptr); }, true);
        bad += ndiff(z1, z4);
        // with the residual norm of the input: same output, same sum
        std::vector<T> z5(nf);
        std::vector<double> pr(NYL_RED_BLOCKS, 0.0), rn2(NYL_RED_BLOCKS, 0.0);
        run_blocks(dim3(nx / SW_TX, ny / SW_TY, nz / kzt), dim3(SW_TX, 4, 1), [&]() {
            (pz ? k_mg_sweep2t<T, false, 1, true, false> : k_mg_sweep2t<T, false, 0, true, false>)(
                gf, x.data(), z5.data(), b.data(), om, tab, kzt, nullptr, pr.data()); }, true);
        run_blocks(dim3(red_blocks((int64_t)nf)), dim3(RED_T), [&]() {
            k_resnorm2<T>(gf, x.data(), b.data(), tab, rn2.data()); }, true);
        double t1 = 0, t2 = 0;
        for (double v : pr) t1 += v;
        for (double v : rn2) t2 += v;
        bad += ndiff(z1, z5) + (std::fabs(t1 - t2) > 1e-9 * t2);
        // from a zero input (src not read) against two sweeps of zeros
        std::vector<T> zero(nf, T(0)), s1(nf), s2(nf), z6(nf);
        run_blocks(mg_grid(gf), dim3(MG_BX, MG_BY, 1), [&]() {
            k_mg_sweep<T>(gf, zero.data(), s1.data(), b.data(), om, tab); }, false);
        run_blocks(mg_grid(gf), dim3(MG_BX, MG_BY, 1), [&]() {
            k_mg_sweep<T>(gf, s1.data(), s2.data(), b.data(), om, tab); }, false);
        run_blocks(dim3(nx / SW_TX, ny / SW_TY, nz / kzt), dim3(SW_TX, 4, 1), [&]() {
            (pz ? k_mg_sweep2t<T, false, 1, false, true> : k_mg_sweep2t<T, false, 0, false, true>)(
                gf, nullptr, z6.data(), b.data(), om, tab, kzt, nullptr, nullptr); }, true);
        bad += ndiff(s2, z6);
        std::vector<double> xa(nf), xb(nf);
        for (size_t q = 0; q < nf; q++) xa[q] = xb[q] = 0.5 - 0.125 * (double)q;
        for (size_t q = 0; q < nf; q++) xa[q] = xa[q] + (double)z1[q];
        run_blocks(dim3(nx / SW_TX, ny / SW_TY, nz / kzt), dim3(SW_TX, 4, 1), [&]() {
            (pz ? k_mg_sweep2t<T, true, 1, false, false> : k_mg_sweep2t<T, true, 0, false, false>)(
                gf, x.data(), nullptr, b.data(), om, tab, kzt, xb.data(), nullptr); },
            true);
        bad += ndiff(xa, xb);
    }
    printf("tiled<%s> %dx%dx%d p=%d%d%d: %d mismatches\n", sizeof(T) == 8 ? "f64" : "f32",
           nx, ny, nz, px, py, pz, bad);
    return bad != 0;
}

// fused level kernels against sweep + restriction and prolongation + two sweeps
template <typename Tf, typename Tc>
int check_fused(int nx, int ny, int nz, int px, int py, int pz) {
    Geo gf = mk(nx, ny, nz, px, py, pz), gc = mk(nx / 2, ny / 2, nz / 2, px, py, pz);
    if (!fused_fits(gf, gc)) return 1;
    MGTabs tab = mktab();
    const size_t nf = (size_t)nx * ny * nz, nc = nf / 8;
    std::mt19937_64 rng(nx * 131 + ny * 17 + nz + px * 5 + py * 7 + pz * 11);
    auto y = rnd<Tf>(nf, rng), b = rnd<Tf>(nf, rng), x0 = rnd<Tf>(nf, rng);
    auto xc = rnd<Tc>(nc, rng);
    std::vector<Tf> xr(nf, Tf(3)), xf(nf, Tf(3));
    std::vector<Tc> bcr(nc, Tc(5)), bcf(nc, Tc(5)), ycr(nc, Tc(6)), ycf(nc, Tc(6)), xd(nc);
    const double om = 0.9;
    run_blocks(mg_grid(gf), dim3(MG_BX, MG_BY, 1), [&]() {
        k_mg_sweep<Tf>(gf, y.data(), xr.data(), b.data(), om, tab); }, false);
    run_blocks(mg_grid(gc), dim3(MG_BX, MG_BY, 1), [&]() {
        k_mg_resid_restrict<Tf, Tc>(gf, gc, xr.data(), b.data(), bcr.data(), xd.data(), tab,
                                    ycr.data(), om); }, false);
    run_blocks(fused_grid(gf), dim3(FUSED_T, 1, 1), [&]() {
        k_mg_down<Tf, Tc>(gf, gc, y.data(), b.data(), xf.data(), bcf.data(), ycf.data(), om,
                          tab); }, true);
    int bad = ndiff(xr, xf) + ndiff(bcr, bcf) + ndiff(ycr, ycf);
    std::vector<Tf> xp = x0, t(nf), ur(nf), uf(nf, Tf(9)), x0c = x0;
    run_blocks(mg_grid(gf), dim3(MG_BX, MG_BY, 1), [&]() {
        k_mg_prolong<Tf, Tc>(gf, gc, xp.data(), xc.data(), tab); }, false);
    run_blocks(mg_grid(gf), dim3(MG_BX, MG_BY, 1), [&]() {
        k_mg_sweep<Tf>(gf, xp.data(), t.data(), b.data(), om, tab); }, false);
    run_blocks(mg_grid(gf), dim3(MG_BX, MG_BY, 1), [&]() {
        k_mg_sweep<Tf>(gf, t.data(), ur.data(), b.data(), om, tab); }, false);
    run_blocks(fused_grid(gf), dim3(FUSED_T, 1, 1), [&]() {
        k_mg_up<Tf, Tc>(gf, gc, x0.data(), xc.data(), b.data(), uf.data(), om, tab); }, true);
    bad += ndiff(ur, uf) + ndiff(x0, x0c);
    printf("fused<%s,%s> %dx%dx%d p=%d%d%d: %d mismatches\n", sizeof(Tf) == 8 ? "f64" : "f32",
           sizeof(Tc) == 8 ? "f64" : "f32", nx, ny, nz, px, py, pz, bad);
    return bad != 0;
}

// The single-workgroup tail (down legs of levels l0 .. L-2, the coarsest
// level's second sweep, the up legs) against the fused launches it replaces:
// the same correction y of level l0, and nothing else written to memory (the
// levels below l0 live in the kernel's LDS).
template <typename T, typename TC>
int check_tail(int nx, int ny, int nz, int px, int py, int pz, int L, int l0) {
    MGTabs tab = mktab();
    const double om = 0.9;
    std::vector<Geo> g(L);
    std::vector<uint64_t> off(3 * L);
    uint64_t bytes = 0;
    for (int l = 0; l < L; l++) {
        g[l] = mk(nx >> l, ny >> l, nz >> l, px, py, pz);
        const uint64_t es = l == L - 1 ? sizeof(TC) : sizeof(T);
        for (int f = 0; f < 3; f++) {
            off[3 * l + f] = bytes;
            bytes += ((uint64_t)(nx >> l) * (ny >> l) * (nz >> l) * es + 255) / 256 * 256;
        }
    }
    std::mt19937_64 rng(nx + 5 * L + 11 * l0 + px + 2 * py + 4 * pz);
    std::vector<char> w1(bytes), w2;
    for (int l = 0; l < L; l++) {
        const size_t n = (size_t)(nx >> l) * (ny >> l) * (nz >> l);
        for (int f = 0; f < 3; f++) {
            if (l == L - 1) {
                auto v = rnd<TC>(n, rng);
                memcpy(w1.data() + off[3 * l + f], v.data(), n * sizeof(TC));
            } else {
                auto v = rnd<T>(n, rng);
                memcpy(w1.data() + off[3 * l + f], v.data(), n * sizeof(T));
            }
        }
    }
    w2 = w1;
    const std::vector<char> w0 = w1;
    char *ws = w1.data();
    auto P = [&](int l, int f) { return reinterpret_cast<T *>(ws + off[3 * l + f]); };
    auto PC = [&](int l, int f) { return reinterpret_cast<TC *>(ws + off[3 * l + f]); };
    // fields: 0 = x, 1 = b, 2 = y
    for (int l = l0; l <= L - 2; l++) {
        if (l + 1 == L - 1)
            run_blocks(fused_grid(g[l]), dim3(FUSED_T, 1, 1), [&]() {
                k_mg_down<T, TC>(g[l], g[l + 1], P(l, 2), P(l, 1), P(l, 0), PC(l + 1, 1),
                                 PC(l + 1, 2), om, tab); }, true);
        else
            run_blocks(fused_grid(g[l]), dim3(FUSED_T, 1, 1), [&]() {
                k_mg_down<T, T>(g[l], g[l + 1], P(l, 2), P(l, 1), P(l, 0), P(l + 1, 1),
                                P(l + 1, 2), om, tab); }, true);
    }
    run_blocks(mg_grid(g[L - 1]), dim3(MG_BX, MG_BY, 1), [&]() {
        k_mg_sweep<TC>(g[L - 1], PC(L - 1, 2), PC(L - 1, 0), PC(L - 1, 1), om, tab); }, false);
    for (int l = L - 2; l >= l0; l--) {
        if (l + 1 == L - 1)
            run_blocks(fused_grid(g[l]), dim3(FUSED_T, 1, 1), [&]() {
                k_mg_up<T, TC>(g[l], g[l + 1], P(l, 0), PC(l + 1, 0), P(l, 1), P(l, 2), om,
                               tab); }, true);
        else
            run_blocks(fused_grid(g[l]), dim3(FUSED_T, 1, 1), [&]() {
                k_mg_up<T, T>(g[l], g[l + 1], P(l, 0), P(l + 1, 2), P(l, 1), P(l, 2), om,
                              tab); }, true);
    }
    TailArgs a;
    memset(&a, 0, sizeof(a));
    a.ws = w2.data();
    for (int l = 0; l < L; l++) {
        a.lev[l].g = g[l];
        a.lev[l].x = off[3 * l];
        a.lev[l].b = off[3 * l + 1];
        a.lev[l].y = off[3 * l + 2];
    }
    a.l0 = l0;
    a.nlev = L;
    a.omega = om;
    a.tab = tab;
    if (tail_layout(a, sizeof(T), sizeof(TC)) > NYL_TAIL_LDS) return 1;
    run_blocks(dim3(1), dim3(TAIL_T, 1, 1), [&]() { k_mg_tail<T, TC>(a); }, true);
    int bad = 0;
    auto cmp = [&](int l, int f) {
        const size_t n = (size_t)(nx >> l) * (ny >> l) * (nz >> l);
        const size_t es = l == L - 1 ? sizeof(TC) : sizeof(T);
        bad += memcmp(w1.data() + off[3 * l + f], w2.data() + off[3 * l + f], n * es) != 0;
    };
    cmp(l0, 2);
    for (int l = 0; l < L; l++) {   // nothing else in memory is written
        for (int f = 0; f < 3; f++) {
            if (l == l0 && f == 2) continue;
            const size_t n = (size_t)(nx >> l) * (ny >> l) * (nz >> l);
            const size_t es = l == L - 1 ? sizeof(TC) : sizeof(T);
            bad += memcmp(w2.data() + off[3 * l + f], w0.data() + off[3 * l + f], n * es) != 0;
        }
    }
    printf("tail<%s,%s> %dx%dx%d L=%d l0=%d p=%d%d%d: %d mismatches\n",
           sizeof(T) == 8 ? "f64" : "f32", sizeof(TC) == 8 ? "f64" : "f32", nx, ny, nz, L, l0,
           px, py, pz, bad);
    return bad != 0;
}

// FP64 outer residual of the mixed solve: tiled (with and without the sum
// of b**2) against per-cell; the partial sums add up to the same totals
int check_outer(int nx, int ny, int nz, int px, int py, int pz) {
    Geo g = mk(nx, ny, nz, px, py, pz);
    MGTabs tab = mktab();
    const size_t n = (size_t)nx * ny * nz;
    std::mt19937_64 rng(7 + px + 2 * py + 4 * pz);
    auto x = rnd<double>(n, rng), b = rnd<double>(n, rng);
    std::vector<double> p1(NYL_RED_BLOCKS, 0.0), p2(NYL_RED_BLOCKS, 0.0), p3(NYL_RED_BLOCKS, 0.0),
        pb(NYL_RED_BLOCKS, 0.0);
    std::vector<float> r1(n), y1(n), r2(n), y2(n), r3(n), y3(n);
    run_blocks(outer_grid(g), dim3(MG_BX, MG_BY, 1), [&]() {
        k_mg_outer_resid<float>(g, x.data(), b.data(), 0.0, r1.data(), y1.data(), 0.9, tab,
                                p1.data()); }, true);
    const int kz = outer25_kz(g);
    const dim3 grd(nx / SW_TX, ny / SW_TY, nz / kz);
    run_blocks(grd, dim3(SW_TX, 4, 1), [&]() {
        k_mg_outer25<float>(g, x.data(), b.data(), 0.0, r2.data(), y2.data(), 0.9, tab, kz,
                            p2.data(), nullptr); }, true);
    run_blocks(grd, dim3(SW_TX, 4, 1), [&]() {
        k_mg_outer25<float>(g, x.data(), b.data(), 0.0, r3.data(), y3.data(), 0.9, tab, kz,
                            p3.data(), pb.data()); }, true);
    int bad = ndiff(r1, r2) + ndiff(y1, y2) + ndiff(r2, r3) + ndiff(y2, y3) + ndiff(p2, p3);
    double s1 = 0, s2 = 0, sb = 0, sb2 = 0;
    for (double v : p1) s1 += v;
    for (double v : p2) s2 += v;
    for (double v : pb) sb += v;
    for (double v : b) sb2 += v * v;
    bad += std::fabs(s1 - s2) > 1e-12 * s1 || std::fabs(sb - sb2) > 1e-12 * sb2;
    printf("outer %dx%dx%d p=%d%d%d: %d mismatches\n", nx, ny, nz, px, py, pz, bad);
    return bad != 0;
}

int main() {
    int fails = 0;
    for (int p = 0; p < 8; p++) {
        const int px = p & 1, py = (p >> 1) & 1, pz = (p >> 2) & 1;
        if (p == 0 || p == 3 || p == 4 || p == 7) {   // 000, 110, 001, 111
            fails += check_tiled<double>(64, 8, 8, px, py, pz);
            fails += check_tiled<float>(128, 16, 4, px, py, pz);
            fails += check_tiled<float>(64, 16, 6, px, py, pz);
            fails += check_tiled<double>(64, 8, 2, px, py, pz);
            fails += check_outer(64, 8, 16, px, py, pz);
        }
        const int sizes[][3] = {{8, 8, 8}, {2, 2, 4}, {4, 2, 2}, {20, 12, 6}};
        for (auto &s : sizes) {
            fails += check_fused<double, double>(s[0], s[1], s[2], px, py, pz);
            fails += check_fused<float, float>(s[0], s[1], s[2], px, py, pz);
            fails += check_fused<float, double>(s[0], s[1], s[2], px, py, pz);
        }
        if (p == 0 || p == 3 || p == 4 || p == 7) {
            fails += check_tail<double, double>(32, 32, 32, px, py, pz, 5, 2);
            fails += check_tail<double, double>(16, 8, 8, px, py, pz, 4, 1);
            fails += check_tail<float, double>(32, 32, 16, px, py, pz, 5, 2);
            fails += check_tail<float, float>(32, 32, 32, px, py, pz, 5, 1);
        }
    }
    printf(fails ? "FAIL %d\n" : "ALL OK\n", fails);
    return fails != 0;
}
