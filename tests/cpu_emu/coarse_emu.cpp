// Direct coarse-level solve (nyl_coarse.hip) on the CPU: the three passes
// solve A x = b - mean(b) to rounding for every periodicity and odd / even /
// one-cell axes, x has zero mean, and the batched form gives the same bits as
// the single-level launches.
#include "../../nyles_amd/csrc/nyl_coarse.hip"
#include "emu.h"
using namespace nyl;

// mgfor's operator on a box: sum of in-domain neighbours - diag * x
static void apply_A(const Geo &g, const std::vector<double> &x, std::vector<double> &ax) {
    for (int k = 0; k < g.nz; k++)
        for (int j = 0; j < g.ny; j++)
            for (int i = 0; i < g.nx; i++) {
                const int n3[3] = {g.nx, g.ny, g.nz}, p3[3] = {g.px, g.py, g.pz};
                int c3[3] = {i, j, k};
                double s = 0.0;
                int nw = 0;
                for (int a = 0; a < 3; a++)
                    for (int d = -1; d <= 1; d += 2) {
                        int q[3] = {c3[0], c3[1], c3[2]};
                        q[a] += d;
                        if (q[a] < 0 || q[a] >= n3[a]) {
                            if (!p3[a]) continue;
                            q[a] = (q[a] + n3[a]) % n3[a];
                        }
                        s += x[((size_t)q[2] * g.ny + q[1]) * g.nx + q[0]];
                        nw++;
                    }
                const size_t c = ((size_t)k * g.ny + j) * g.nx + i;
                ax[c] = s - nw * x[c];
            }
}

static void solve_single(const CoarseArgs &a) {
    run_blocks(dim3(a.g.nz), dim3(CS_T), [&]() { k_cs<0>(a); }, true);
    run_blocks(dim3(a.g.ny), dim3(CS_T), [&]() { k_cs<1>(a); }, true);
    run_blocks(dim3(a.g.nz), dim3(CS_T), [&]() { k_cs<2>(a); }, true);
}

struct Case {
    Geo g;
    std::vector<double> tab, b, s, x;
};

static Case make_case(int nx, int ny, int nz, int px, int py, int pz, std::mt19937_64 &rng) {
    Case c;
    c.g = mk(nx, ny, nz, px, py, pz);
    const size_t n = (size_t)nx * ny * nz;
    c.tab.assign(coarse_table_doubles(), 0.0);
    coarse_tables(c.g, c.tab.data());
    c.b = rnd<double>(n, rng);
    double m = 0.0;
    for (double v : c.b) m += v;
    m /= n;
    for (double &v : c.b) v -= m;
    c.s.assign(n, 0.0);
    c.x.assign(n, 0.0);
    return c;
}

static CoarseArgs args(Case &c) {
    CoarseArgs a;
    a.g = c.g;
    a.b = c.b.data();
    a.s = c.s.data();
    a.x = c.x.data();
    a.tab = c.tab.data();
    return a;
}

int main() {
    std::mt19937_64 rng(11);
    const int sizes[][3] = {{5, 5, 5}, {3, 3, 3}, {7, 6, 5}, {1, 4, 3}, {2, 2, 2}, {9, 1, 11},
                            {31, 31, 31}, {32, 8, 3}, {1, 1, 1}, {34, 17, 17}, {62, 31, 31},
                            {64, 32, 32}};
    int fails = 0;
    for (auto &sz : sizes) {
        for (int p = 0; p < 8; p++) {
            if (sz[0] * sz[1] * sz[2] > 4000 && p != 3) continue;   // the large case: perio_xy
            Case c = make_case(sz[0], sz[1], sz[2], p & 1, (p >> 1) & 1, (p >> 2) & 1, rng);
            solve_single(args(c));
            const size_t n = c.b.size();
            std::vector<double> ax(n);
            apply_A(c.g, c.x, ax);
            double rmax = 0.0, bmax = 0.0, xs = 0.0, xmax = 0.0;
            for (size_t q = 0; q < n; q++) {
                rmax = std::max(rmax, std::fabs(c.b[q] - ax[q]));
                bmax = std::max(bmax, std::fabs(c.b[q]));
                xmax = std::max(xmax, std::fabs(c.x[q]));
                xs += c.x[q];
            }
            const bool ok = rmax <= 1e-11 * std::max(bmax, 1e-300) + 1e-300 * 0 &&
                            std::fabs(xs) <= 1e-10 * (xmax * n + 1e-300);
            const bool all_zero_ok = n == 1 ? (c.x[0] == 0.0) : true;
            printf("coarse %dx%dx%d p=%d%d%d: |b-Ax| %.2e (|b| %.2e) sum x %.2e %s\n", sz[0],
                   sz[1], sz[2], p & 1, (p >> 1) & 1, (p >> 2) & 1, rmax, bmax, xs,
                   ok && all_zero_ok ? "ok" : "FAIL");
            fails += !(ok && all_zero_ok);
        }
    }
    // batched: two members of different sizes give the single results bitwise
    {
        Case c1 = make_case(5, 7, 3, 1, 0, 0, rng), c2 = make_case(6, 3, 9, 0, 1, 1, rng);
        Case d1 = c1, d2 = c2;
        solve_single(args(c1));
        solve_single(args(c2));
        std::vector<CoarseItem> items(2);
        Case *cs[2] = {&d1, &d2};
        for (int pass = 0; pass < 3; pass++) {
            int64_t start = 0;
            for (int q = 0; q < 2; q++) {
                items[q].h.g = cs[q]->g;
                items[q].h.member = q;
                items[q].h.nblk = pass == 1 ? cs[q]->g.ny : cs[q]->g.nz;
                items[q].h.start = start;
                start += items[q].h.nblk;
                items[q].a = args(*cs[q]);
            }
            const int active[2] = {1, 1};
            run_blocks(dim3((unsigned)start), dim3(CS_T), [&]() {
                if (pass == 0) k_b_cs<0>(items.data(), 2, active);
                else if (pass == 1) k_b_cs<1>(items.data(), 2, active);
                else k_b_cs<2>(items.data(), 2, active);
            }, true);
        }
        const int bad = ndiff(c1.x, d1.x) + ndiff(c2.x, d2.x);
        printf("batched coarse solve: %d mismatches\n", bad);
        fails += bad != 0;
    }
    if (fails) {
        printf("FAILED %d\n", fails);
        return 1;
    }
    printf("ALL OK\n");
    return 0;
}
