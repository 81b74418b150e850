// Harness of the kernel-equivalence tests: runs a kernel body over a grid on
// the CPU, one host thread per work-item of a block when the kernel uses
// barriers or lane exchanges, sequentially otherwise.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

thread_local dim3 threadIdx, blockIdx;
dim3 gridDim, blockDim;
pthread_barrier_t emu_barrier;
double emu_shfl[1024];

template <typename Fn>
void run_blocks(dim3 grid, dim3 blk, Fn body, bool threaded) {
    gridDim = grid;
    blockDim = blk;
    const int nt = blk.x * blk.y * blk.z;
    if (!threaded) {
        for (unsigned bz = 0; bz < grid.z; bz++)
            for (unsigned by = 0; by < grid.y; by++)
                for (unsigned bx = 0; bx < grid.x; bx++)
                    for (unsigned tz = 0; tz < blk.z; tz++)
                        for (unsigned ty = 0; ty < blk.y; ty++)
                            for (unsigned tx = 0; tx < blk.x; tx++) {
                                blockIdx = dim3(bx, by, bz);
                                threadIdx = dim3(tx, ty, tz);
                                body();
                            }
        return;
    }
    // one host thread per work-item, all blocks in turn; a barrier after each
    // block (the LDS arrays are shared by the blocks of a launch)
    pthread_barrier_init(&emu_barrier, nullptr, nt);
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
        th.emplace_back([&, t]() {
            threadIdx = dim3(t % blk.x, (t / blk.x) % blk.y, t / (blk.x * blk.y));
            for (unsigned bz = 0; bz < grid.z; bz++)
                for (unsigned by = 0; by < grid.y; by++)
                    for (unsigned bx = 0; bx < grid.x; bx++) {
                        blockIdx = dim3(bx, by, bz);
                        body();
                        pthread_barrier_wait(&emu_barrier);
                    }
        });
    for (auto &x : th) x.join();
    pthread_barrier_destroy(&emu_barrier);
}

static nyl::Geo mk(int nx, int ny, int nz, int px, int py, int pz) {
    nyl::Geo g;
    g.nx = nx; g.ny = ny; g.nz = nz; g.px = px; g.py = py; g.pz = pz;
    g.sy = nx; g.sz = (int64_t)nx * ny;
    return g;
}

// mgfor's coefficient tables in closed form (as the library builds them)
static nyl::MGTabs mktab() {
    nyl::MGTabs tab;
    for (int d = 0; d <= 6; d++) { tab.diag[d] = d; tab.idiag[d] = d ? 1.0 / d : 0.0; }
    for (int m = 0; m < 8; m++)
        tab.pcoef[m] = 1.0 / ((3 + (m & 1)) * (3 + ((m >> 1) & 1)) * (3 + ((m >> 2) & 1)));
    return tab;
}

template <typename T>
static std::vector<T> rnd(size_t n, std::mt19937_64 &rng) {
    std::normal_distribution<double> nd;
    std::vector<T> v(n);
    for (auto &x : v) x = T(nd(rng));
    return v;
}

template <typename T>
static int ndiff(const std::vector<T> &a, const std::vector<T> &b) {
    int bad = 0;
    for (size_t q = 0; q < a.size(); q++) bad += memcmp(&a[q], &b[q], sizeof(T)) != 0;
    return bad;
}
