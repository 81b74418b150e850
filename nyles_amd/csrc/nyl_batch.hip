// nyl_batch.hip -- ensembles: many model contexts (members, of uneven grid
// sizes and any geometry) advanced together by RK3_SSP steps, with one
// launch per kernel class and RK stage for the whole batch.
//
// Every batched kernel runs the single-member cell functions on a table of
// per-member items (BHead + arguments) held in the batch workspace, so a
// member's arithmetic does not depend on the batch it is in.  The multigrid
// stop rule (core/mgfor/solvers.f90:17-30) is evaluated per member on the
// device (k_b_final), members that have stopped skip the V-cycle kernels,
// and the host only reads one counter per V-cycle.  Each plan (stage kernel
// + divergence + initial norms, one V-cycle, correction + diagnostics) is a
// fixed launch sequence and is captured once into a hipGraph.
#include <string.h>

#include <algorithm>
#include <vector>

#include "nyl_ctx.h"

using namespace nyl;

namespace {

#define BCHK(...)                                                                      \
    do {                                                                               \
        hipError_t e_ = (__VA_ARGS__);                                                 \
        if (e_ != hipSuccess)                                                          \
            return set_error(NYL_E_HIP, "%s: %s (%s:%d)", #__VA_ARGS__,                \
                             hipGetErrorString(e_), __FILE__, __LINE__);               \
    } while (0)

uint64_t al256(uint64_t v) { return (v + 255) & ~uint64_t(255); }

// one launch of a plan
enum LKind {
    L_STAGE, L_DIV, L_CORR, L_DIAG, L_SWEEP, L_RESTR, L_PROL, L_RED, L_OUTER, L_FINAL,
    L_ACC, L_COARSE
};
enum LType { T_D, T_F, T_FD };   // double, float, float fine over double coarse

struct Launch {
    LKind kind;
    LType type;
    uint64_t off;       // item table in the batch workspace
    int nitems;
    int64_t nblk;
    int arg;            // mode / kind / what / periodicity
    int arg2;           // stage: periodicity class
};

}  // namespace

struct nyl_batch {
    int n = 0;
    std::vector<nyl_ctx *> m;
    std::vector<int> cur0;      // members' state buffer at table build
    int mixed = 0;
    int lmax = 0;
    StageTile tile = {256, -1, 0};   // stage kernel tile (threads = cells per tile)
    int use_graphs = 1;
    int phase = 0;              // member cur = cur0 ^ phase
    char *ws = nullptr;
    uint64_t ws_bytes = 0, tab_bytes = 0;
    hipStream_t stream = nullptr;
    MGTabs tabs;
    // device control arrays (offsets)
    uint64_t o_normb, o_res, o_shift, o_tol, o_nite, o_maxite, o_active, o_any, o_dt, o_acc;
    BMG mg;
    // plans
    std::vector<Launch> pre[2][3], vcyc, post[2];
    std::vector<char> blob;     // host image of the item tables
    hipGraphExec_t g_pre[2][3] = {}, g_vcyc = nullptr, g_post[2] = {};
    int *host_any = nullptr;    // pinned: members still iterating, most V-cycles so far
    int vcap[3] = {0, 0, 0};    // V-cycles launched without a poll, per RK stage
    double *host_dt = nullptr;  // pinned, n
    void *host_stats = nullptr; // pinned: nite_acc (int n), res (double n), normb (double n)
};

namespace {

BHead head(const Geo &g, int member, int64_t ncell_for_blocks, int64_t &start,
           int64_t cap = 0) {
    BHead h;
    h.g = g;
    h.member = member;
    int64_t nb = batch_blocks(ncell_for_blocks);
    if (cap > 0 && nb > cap) nb = cap;
    h.nblk = (int)nb;
    h.start = start;
    start += nb;
    return h;
}

int64_t gcells(const Geo &g) { return (int64_t)g.nx * g.ny * g.nz; }

template <typename I>
Launch add_table(nyl_batch *b, LKind kind, LType type, const std::vector<I> &items, int64_t nblk,
                 int arg = 0, int arg2 = 0) {
    Launch L;
    L.kind = kind;
    L.type = type;
    L.off = al256(b->blob.size());
    L.nitems = (int)items.size();
    L.nblk = nblk;
    L.arg = arg;
    L.arg2 = arg2;
    b->blob.resize(L.off + sizeof(I) * items.size());
    memcpy(b->blob.data() + L.off, items.data(), sizeof(I) * items.size());
    return L;
}

// stage-kernel items of tile shape `tile` for members whose periodicity
// class is `perio`, at phase p and RK stage st (1..3)
template <typename T>
std::vector<StageItem<T>> stage_items(nyl_batch *b, int p, int st, int perio,
                                      const StageTile &tile, int64_t &nblk) {
    std::vector<StageItem<T>> v;
    nblk = 0;
    for (int q = 0; q < b->n; q++) {
        nyl_ctx *c = b->m[q];
        if (c->g.px + 2 * c->g.py + 4 * c->g.pz != perio) continue;
        const int s = b->cur0[q] ^ p, t = 1 - s;
        StageItem<T> it;
        memset(&it, 0, sizeof(it));
        it.h.g = c->g;
        it.h.member = q;
        it.h.nblk = (int)stage_tiles(c->g, tile);
        it.h.start = nblk;
        nblk += it.h.nblk;
        StageArgs<T> &a = it.a;
        a.b = F<T>(c, c->state[s][0]); a.u = F<T>(c, c->state[s][1]);
        a.v = F<T>(c, c->state[s][2]); a.w = F<T>(c, c->state[s][3]);
        a.bn = F<T>(c, c->state[t][0]); a.un = F<T>(c, c->state[t][1]);
        a.vn = F<T>(c, c->state[t][2]); a.wn = F<T>(c, c->state[t][3]);
        a.ox = F<T>(c, c->vor[0]); a.oy = F<T>(c, c->vor[1]); a.oz = F<T>(c, c->vor[2]);
        a.ke = F<T>(c, c->ke);
        a.d0b = F<T>(c, c->ds[0][0]); a.d0u = F<T>(c, c->ds[0][1]);
        a.d0v = F<T>(c, c->ds[0][2]); a.d0w = F<T>(c, c->ds[0][3]);
        a.d1b = F<T>(c, c->ds[1][0]); a.d1u = F<T>(c, c->ds[1][1]);
        a.d1v = F<T>(c, c->ds[1][2]); a.d1w = F<T>(c, c->ds[1][3]);
        if (st == 1 || st == 2) {
            a.tb = F<T>(c, c->ds[st - 1][0]); a.tu = F<T>(c, c->ds[st - 1][1]);
            a.tv = F<T>(c, c->ds[st - 1][2]); a.tw = F<T>(c, c->ds[st - 1][3]);
        }
        a.idx2 = c->d.idx2; a.idy2 = c->d.idy2; a.idz2 = c->d.idz2;
        a.dz = c->d.dz;
        if (st == 3 && c->d.viscosity != 0.0) {   // viscosity at the last RK stage
            a.visc = 1;
            a.vcx = c->d.viscosity * c->d.idx2;
            a.vcy = c->d.viscosity * c->d.idy2;
            a.vcz = c->d.viscosity * c->d.idz2;
        }
        v.push_back(it);
    }
    return v;
}

template <typename T>
void build_stage(nyl_batch *b, int p, int st, std::vector<Launch> &out) {
    for (int perio = 0; perio < 8; perio++) {
        int64_t nblk;
        auto v = stage_items<T>(b, p, st, perio, b->tile, nblk);
        if (!v.empty()) out.push_back(add_table(b, L_STAGE, T_D, v, nblk, st, perio));
    }
}

// FinalItem table over all members' level-1 partials (nparts from the items
// of the reduction that fills them)
Launch final_table(nyl_batch *b, const std::vector<int> &nparts, int kind) {
    std::vector<FinalItem> v(b->n);
    for (int q = 0; q < b->n; q++) {
        nyl_ctx *c = b->m[q];
        v[q].partials = F<double>(c, c->partials);
        v[q].nparts = nparts[q];
        v[q].member = q;
        v[q].ncell = (double)gcells(c->lev[0].g);
    }
    return add_table(b, L_FINAL, T_D, v, b->n, kind);
}

// partials per member of the batched reductions: at most half the member's
// partials buffer (kind 2 stores two sets)
constexpr int64_t kRedCap = NYL_RED_BLOCKS / 2;

template <typename T>
void build_pre(nyl_batch *b, int p, int st, std::vector<Launch> &out) {
    if (st == 1) out.push_back(Launch{L_ACC, T_D, 0, 0, 0, 0, 0});   // zero the V-cycle counts
    build_stage<T>(b, p, st, out);
    // divergence of the new state into the FP64 right-hand side
    {
        std::vector<DivItem<T>> v;
        int64_t nb = 0;
        for (int q = 0; q < b->n; q++) {
            nyl_ctx *c = b->m[q];
            const int s = 1 - (b->cur0[q] ^ p);   // the stage wrote the other buffer
            DivItem<T> it;
            it.h = head(c->g, q, gcells(c->g), nb);
            it.a.u = F<T>(c, c->state[s][1]); it.a.v = F<T>(c, c->state[s][2]);
            it.a.w = F<T>(c, c->state[s][3]);
            it.a.idx2 = c->d.idx2; it.a.idy2 = c->d.idy2; it.a.idz2 = c->d.idz2;
            it.a.out = F<double>(c, c->mgb);
            v.push_back(it);
        }
        out.push_back(add_table(b, L_DIV, b->mixed ? T_F : T_D, v, nb));
    }
    std::vector<int> np(b->n);
    if (!b->mixed) {
        // normb = |b|^2, then the residual of the warm-started x
        std::vector<RedItem<double>> v0, v1;
        int64_t nb0 = 0, nb1 = 0;
        for (int q = 0; q < b->n; q++) {
            nyl_ctx *c = b->m[q];
            RedItem<double> it;
            it.h = head(c->lev[0].g, q, gcells(c->lev[0].g), nb0, kRedCap);
            it.v = F<double>(c, c->lev[0].b);
            it.b = nullptr;
            it.partials = F<double>(c, c->partials);
            np[q] = it.h.nblk;
            v0.push_back(it);
            it.h = head(c->lev[0].g, q, gcells(c->lev[0].g), nb1, kRedCap);
            it.v = F<double>(c, c->lev[0].x);
            it.b = F<double>(c, c->lev[0].b);
            v1.push_back(it);
        }
        out.push_back(add_table(b, L_RED, T_D, v0, nb0, 0));
        out.push_back(final_table(b, np, NYL_FIN_NORMB));
        out.push_back(add_table(b, L_RED, T_D, v1, nb1, 1));
        out.push_back(final_table(b, np, NYL_FIN_INIT));
    } else {
        // sum and sum of squares of b (mean removal, normb), then the FP64
        // outer residual of x
        std::vector<RedItem<double>> v0;
        std::vector<OuterItem> v1;
        int64_t nb0 = 0, nb1 = 0;
        std::vector<int> np1(b->n);
        for (int q = 0; q < b->n; q++) {
            nyl_ctx *c = b->m[q];
            RedItem<double> it;
            it.h = head(c->lev[0].g, q, gcells(c->lev[0].g), nb0, kRedCap);
            it.v = F<double>(c, c->mgb);
            it.b = nullptr;
            it.partials = F<double>(c, c->partials);
            np[q] = it.h.nblk;
            v0.push_back(it);
            OuterItem o;
            o.h = head(c->lev[0].g, q, gcells(c->lev[0].g), nb1, kRedCap);
            o.x = F<double>(c, c->mgx); o.b = F<double>(c, c->mgb);
            o.r = F<float>(c, c->lev[0].b); o.y = F<float>(c, c->lev[0].y);
            o.omega = c->d.mg_omega;
            o.partials = F<double>(c, c->partials);
            np1[q] = o.h.nblk;
            v1.push_back(o);
        }
        out.push_back(add_table(b, L_RED, T_D, v0, nb0, 2));
        out.push_back(final_table(b, np, NYL_FIN_SHIFT));
        out.push_back(add_table(b, L_OUTER, T_D, v1, nb1));
        out.push_back(final_table(b, np1, NYL_FIN_INIT));
    }
}

// Level element types of member q: FP64 batch: all double; mixed: FP32
// except the coarsest level.
bool b_lev64(const nyl_batch *b, int q, int l) { return lev64(b->m[q], l); }

// the sweeps x -> y (first) and/or y -> x (second) of level l for members
// whose level l has the type T and that have more than `below` levels below l
template <typename T>
void sweep_launch(nyl_batch *b, int l, bool first, bool second, bool t64,
                  std::vector<Launch> &out, bool skip_first_l0, int below = 0,
                  bool addx_l0 = false) {
    for (int pass = 0; pass < 2; pass++) {
        if ((pass == 0 && !first) || (pass == 1 && !second)) continue;
        std::vector<SweepItem<T>> v;
        int64_t nb = 0;
        for (int q = 0; q < b->n; q++) {
            nyl_ctx *c = b->m[q];
            if (l + below >= c->nlev || b_lev64(b, q, l) != t64) continue;
            if (pass == 0 && l == 0 && skip_first_l0) continue;
            if (c->exact && l == c->nlev - 1) continue;   // solved directly
            const Level &f = c->lev[l];
            SweepItem<T> it;
            it.h = head(f.g, q, gcells(f.g), nb);
            it.src = F<T>(c, pass == 0 ? f.x : f.y);
            it.dst = F<T>(c, pass == 0 ? f.y : f.x);
            it.b = F<T>(c, f.b);
            it.omega = c->d.mg_omega;
            // the mixed solve's last sweep of a V-cycle adds its result to x
            it.xo = (addx_l0 && pass == 1 && l == 0) ? F<double>(c, c->mgx) : nullptr;
            v.push_back(it);
        }
        if (!v.empty()) out.push_back(add_table(b, L_SWEEP, t64 ? T_D : T_F, v, nb));
    }
}

void smooth_launches(nyl_batch *b, int l, std::vector<Launch> &out, bool skip_first_l0) {
    sweep_launch<double>(b, l, true, false, true, out, skip_first_l0);
    sweep_launch<float>(b, l, true, false, false, out, skip_first_l0);
    sweep_launch<double>(b, l, false, true, true, out, skip_first_l0);
    sweep_launch<float>(b, l, false, true, false, out, skip_first_l0);
}

// the direct solve of the members whose coarsest level is l and has an odd
// size (three passes: planes, rows, planes)
void coarse_launches(nyl_batch *b, int l, std::vector<Launch> &out) {
    for (int pass = 0; pass < 3; pass++) {
        std::vector<CoarseItem> v;
        int64_t nb = 0;
        for (int q = 0; q < b->n; q++) {
            nyl_ctx *c = b->m[q];
            if (!c->exact || l != c->nlev - 1) continue;
            const Level &f = c->lev[l];
            CoarseItem it;
            it.h.g = f.g;
            it.h.member = q;
            it.h.nblk = pass == 1 ? f.g.ny : f.g.nz;
            it.h.start = nb;
            nb += it.h.nblk;
            it.a.g = f.g;
            it.a.b = F<double>(c, f.b);
            it.a.s = F<double>(c, f.y);
            it.a.x = F<double>(c, f.x);
            it.a.tab = F<double>(c, c->cs_tab);
            v.push_back(it);
        }
        if (!v.empty()) out.push_back(add_table(b, L_COARSE, T_D, v, nb, pass));
    }
}

template <typename Tf, typename Tc>
void restrict_launch(nyl_batch *b, int l, LType type, std::vector<Launch> &out) {
    std::vector<RestrItem<Tf, Tc>> v;
    int64_t nb = 0;
    const bool f64 = std::is_same<Tf, double>::value, c64 = std::is_same<Tc, double>::value;
    for (int q = 0; q < b->n; q++) {
        nyl_ctx *c = b->m[q];
        if (l + 1 >= c->nlev || b_lev64(b, q, l) != f64 || b_lev64(b, q, l + 1) != c64) continue;
        const Level &f = c->lev[l], &k = c->lev[l + 1];
        RestrItem<Tf, Tc> it;
        it.h = head(f.g, q, gcells(k.g), nb);
        it.gc = k.g;
        it.x = F<Tf>(c, f.x); it.b = F<Tf>(c, f.b);
        it.bc = F<Tc>(c, k.b); it.xc = F<Tc>(c, k.x);
        v.push_back(it);
    }
    if (!v.empty()) out.push_back(add_table(b, L_RESTR, type, v, nb));
}

template <typename Tf, typename Tc>
void prolong_launch(nyl_batch *b, int l, LType type, std::vector<Launch> &out) {
    std::vector<ProlItem<Tf, Tc>> v;
    int64_t nb = 0;
    const bool f64 = std::is_same<Tf, double>::value, c64 = std::is_same<Tc, double>::value;
    for (int q = 0; q < b->n; q++) {
        nyl_ctx *c = b->m[q];
        if (l + 1 >= c->nlev || b_lev64(b, q, l) != f64 || b_lev64(b, q, l + 1) != c64) continue;
        const Level &f = c->lev[l], &k = c->lev[l + 1];
        ProlItem<Tf, Tc> it;
        it.h = head(f.g, q, gcells(f.g), nb);
        it.gc = k.g;
        it.xf = F<Tf>(c, f.x); it.xc = F<Tc>(c, k.x);
        v.push_back(it);
    }
    if (!v.empty()) out.push_back(add_table(b, L_PROL, type, v, nb));
}

// one V-cycle of every active member (core/mgfor/solvers.f90:35-55) and the
// stop rule after it
void build_vcycle(nyl_batch *b, std::vector<Launch> &out) {
    const bool skip = b->mixed != 0;   // mixed: the first level-1 sweep comes from the outer residual
    for (int l = 0; l < b->lmax; l++) {
        smooth_launches(b, l, out, skip);
        coarse_launches(b, l, out);
        restrict_launch<double, double>(b, l, T_D, out);
        restrict_launch<float, float>(b, l, T_F, out);
        restrict_launch<float, double>(b, l, T_FD, out);
    }
    for (int l = b->lmax - 2; l >= 0; l--) {
        prolong_launch<double, double>(b, l, T_D, out);
        prolong_launch<float, float>(b, l, T_F, out);
        prolong_launch<float, double>(b, l, T_FD, out);
        // post-smoothing of the members that have a level below l
        sweep_launch<double>(b, l, true, false, true, out, false, 1);
        sweep_launch<float>(b, l, true, false, false, out, false, 1);
        sweep_launch<double>(b, l, false, true, true, out, false, 1);
        sweep_launch<float>(b, l, false, true, false, out, false, 1, b->mixed != 0);
    }
    std::vector<int> np(b->n);
    if (!b->mixed) {
        std::vector<RedItem<double>> v;
        int64_t nb = 0;
        for (int q = 0; q < b->n; q++) {
            nyl_ctx *c = b->m[q];
            RedItem<double> it;
            it.h = head(c->lev[0].g, q, gcells(c->lev[0].g), nb, kRedCap);
            it.v = F<double>(c, c->lev[0].x);
            it.b = F<double>(c, c->lev[0].b);
            it.partials = F<double>(c, c->partials);
            np[q] = it.h.nblk;
            v.push_back(it);
        }
        out.push_back(add_table(b, L_RED, T_D, v, nb, 1, 1));   // gated by active
    } else {
        // x += e was done by the last level-1 sweep (its result added to x)
        std::vector<OuterItem> vo;
        int64_t nbo = 0;
        for (int q = 0; q < b->n; q++) {
            nyl_ctx *c = b->m[q];
            OuterItem o;
            o.h = head(c->lev[0].g, q, gcells(c->lev[0].g), nbo, kRedCap);
            o.x = F<double>(c, c->mgx); o.b = F<double>(c, c->mgb);
            o.r = F<float>(c, c->lev[0].b); o.y = F<float>(c, c->lev[0].y);
            o.omega = c->d.mg_omega;
            o.partials = F<double>(c, c->partials);
            np[q] = o.h.nblk;
            vo.push_back(o);
        }
        out.push_back(add_table(b, L_OUTER, T_D, vo, nbo, 0, 1));   // gated by active
    }
    out.push_back(final_table(b, np, NYL_FIN_STEP));
}

template <typename T>
void build_post(nyl_batch *b, int p, std::vector<Launch> &out) {
    std::vector<CorrItem<T>> vc;
    std::vector<DiagItem<T>> vd;
    int64_t nbc = 0, nbd = 0;
    for (int q = 0; q < b->n; q++) {
        nyl_ctx *c = b->m[q];
        const int s = 1 - (b->cur0[q] ^ p);
        CorrItem<T> it;
        it.h = head(c->g, q, gcells(c->g), nbc);
        it.a.u = F<T>(c, c->state[s][1]); it.a.v = F<T>(c, c->state[s][2]);
        it.a.w = F<T>(c, c->state[s][3]);
        it.a.x = F<double>(c, c->mgx);
        it.a.dx2 = c->d.dx2;
        vc.push_back(it);
        DiagItem<T> d;
        d.h = head(c->g, q, gcells(c->g), nbd);
        d.a.u = F<T>(c, c->state[s][1]); d.a.v = F<T>(c, c->state[s][2]);
        d.a.w = F<T>(c, c->state[s][3]);
        d.a.Ux = F<T>(c, c->uc[0]); d.a.Uy = F<T>(c, c->uc[1]); d.a.Uz = F<T>(c, c->uc[2]);
        d.a.ox = F<T>(c, c->vor[0]); d.a.oy = F<T>(c, c->vor[1]); d.a.oz = F<T>(c, c->vor[2]);
        d.a.ke = F<T>(c, c->ke);
        d.a.idx2 = c->d.idx2; d.a.idy2 = c->d.idy2; d.a.idz2 = c->d.idz2;
        d.a.fpar = c->d.fparameter > 0.0 ? c->d.fparameter : 0.0;
        vd.push_back(d);
    }
    const LType t = b->mixed ? T_F : T_D;
    out.push_back(add_table(b, L_CORR, t, vc, nbc));
    out.push_back(add_table(b, L_DIAG, t, vd, nbd, 6));
    out.push_back(Launch{L_ACC, T_D, 0, 0, 0, 1, 0});   // add this stage's V-cycles
}

void destroy_graphs(nyl_batch *b) {
    for (int p = 0; p < 2; p++) {
        for (int s = 0; s < 3; s++)
            if (b->g_pre[p][s]) { (void)hipGraphExecDestroy(b->g_pre[p][s]); b->g_pre[p][s] = nullptr; }
        if (b->g_post[p]) { (void)hipGraphExecDestroy(b->g_post[p]); b->g_post[p] = nullptr; }
    }
    if (b->g_vcyc) { (void)hipGraphExecDestroy(b->g_vcyc); b->g_vcyc = nullptr; }
}

// (re)build every plan's item tables for the members' current state buffers
template <typename T>
void build_plans(nyl_batch *b) {
    b->blob.clear();
    for (int q = 0; q < b->n; q++) b->cur0[q] = b->m[q]->cur;
    b->phase = 0;
    for (int p = 0; p < 2; p++) {
        for (int s = 0; s < 3; s++) { b->pre[p][s].clear(); build_pre<T>(b, p, s + 1, b->pre[p][s]); }
        b->post[p].clear();
        build_post<T>(b, p, b->post[p]);
    }
    b->vcyc.clear();
    build_vcycle(b, b->vcyc);
}

}  // namespace

// offset of the item tables inside the batch workspace
static uint64_t tables_off(const nyl_batch *b) { return al256(b->o_acc + sizeof(int) * b->n); }

namespace {

int run_launch(nyl_batch *b, const Launch &L) {
    char *tab = b->ws + tables_off(b) + L.off;
    hipStream_t s = b->stream;
    const int *act = b->mg.active;
    switch (L.kind) {
    case L_STAGE:
        if (b->mixed) BCHK(launch_b_stage<float>((const StageItem<float> *)tab, L.nitems, L.nblk, L.arg, L.arg2, b->tile, b->mg.dt, s));
        else BCHK(launch_b_stage<double>((const StageItem<double> *)tab, L.nitems, L.nblk, L.arg, L.arg2, b->tile, b->mg.dt, s));
        break;
    case L_DIV:
        if (L.type == T_F) BCHK(launch_b_div<float>((const DivItem<float> *)tab, L.nitems, L.nblk, s));
        else BCHK(launch_b_div<double>((const DivItem<double> *)tab, L.nitems, L.nblk, s));
        break;
    case L_CORR:
        if (L.type == T_F) BCHK(launch_b_correct<float>((const CorrItem<float> *)tab, L.nitems, L.nblk, s));
        else BCHK(launch_b_correct<double>((const CorrItem<double> *)tab, L.nitems, L.nblk, s));
        break;
    case L_DIAG:
        if (L.type == T_F) BCHK(launch_b_diag<float>((const DiagItem<float> *)tab, L.nitems, L.nblk, L.arg, s));
        else BCHK(launch_b_diag<double>((const DiagItem<double> *)tab, L.nitems, L.nblk, L.arg, s));
        break;
    case L_SWEEP:
        if (L.type == T_F) BCHK(launch_b_sweep<float>((const SweepItem<float> *)tab, L.nitems, L.nblk, b->tabs, act, s));
        else BCHK(launch_b_sweep<double>((const SweepItem<double> *)tab, L.nitems, L.nblk, b->tabs, act, s));
        break;
    case L_RESTR:
        if (L.type == T_F) BCHK(launch_b_restrict<float, float>((const RestrItem<float, float> *)tab, L.nitems, L.nblk, b->tabs, act, s));
        else if (L.type == T_FD) BCHK(launch_b_restrict<float, double>((const RestrItem<float, double> *)tab, L.nitems, L.nblk, b->tabs, act, s));
        else BCHK(launch_b_restrict<double, double>((const RestrItem<double, double> *)tab, L.nitems, L.nblk, b->tabs, act, s));
        break;
    case L_PROL:
        if (L.type == T_F) BCHK(launch_b_prolong<float, float>((const ProlItem<float, float> *)tab, L.nitems, L.nblk, b->tabs, act, s));
        else if (L.type == T_FD) BCHK(launch_b_prolong<float, double>((const ProlItem<float, double> *)tab, L.nitems, L.nblk, b->tabs, act, s));
        else BCHK(launch_b_prolong<double, double>((const ProlItem<double, double> *)tab, L.nitems, L.nblk, b->tabs, act, s));
        break;
    case L_RED:
        // the initial reductions run for every member, the V-cycle's (arg2) for active ones
        BCHK(launch_b_reduce<double>((const RedItem<double> *)tab, L.nitems, L.nblk, L.arg,
                                     b->tabs, L.arg2 ? act : nullptr, s));
        break;
    case L_OUTER:
        // the initial residual runs for every member, the V-cycle's (arg2) for active ones
        BCHK(launch_b_outer_resid((const OuterItem *)tab, L.nitems, L.nblk, b->mg, b->tabs,
                                  L.arg2 ? act : nullptr, s));
        break;
    case L_FINAL:
        BCHK(launch_b_final((const FinalItem *)tab, L.nitems, L.arg, b->mg, s));
        break;
    case L_ACC:
        BCHK(launch_b_acc(b->n, b->mg, L.arg, s));
        break;
    case L_COARSE:
        BCHK(launch_b_coarse((const CoarseItem *)tab, L.nitems, L.nblk, L.arg, act, s));
        break;
    }
    return NYL_OK;
}

int run_plan(nyl_batch *b, std::vector<Launch> &plan, hipGraphExec_t *g) {
    if (b->use_graphs) {
        if (!*g) {
            hipGraph_t graph;
            BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            int rc = NYL_OK;
            for (auto &L : plan)
                if ((rc = run_launch(b, L))) break;
            hipError_t e = hipStreamEndCapture(b->stream, &graph);
            if (rc) return rc;
            BCHK(e);
            BCHK(hipGraphInstantiate(g, graph, nullptr, nullptr, 0));
            BCHK(hipGraphDestroy(graph));
        }
        BCHK(hipGraphLaunch(*g, b->stream));
        return NYL_OK;
    }
    for (auto &L : plan) {
        int rc = run_launch(b, L);
        if (rc) return rc;
    }
    return NYL_OK;
}

int read_any(nyl_batch *b, int *any) {
    BCHK(hipMemcpyAsync(b->host_any, b->mg.any, 2 * sizeof(int), hipMemcpyDeviceToHost,
                        b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    *any = b->host_any[0];
    return NYL_OK;
}

}  // namespace

namespace {

// item-table room for the autotuner: one stage plan at the smallest tile
uint64_t tune_bytes(const nyl_batch *b) {
    return al256((uint64_t)b->n * std::max(sizeof(StageItem<double>), sizeof(StageItem<float>))) * 8;
}

int upload_tables(nyl_batch *b) {
    BCHK(hipMemcpyAsync(b->ws + tables_off(b), b->blob.data(), b->blob.size(),
                        hipMemcpyHostToDevice, b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    return NYL_OK;
}

int rebuild(nyl_batch *b) {
    destroy_graphs(b);
    if (b->mixed) build_plans<float>(b);
    else build_plans<double>(b);
    return upload_tables(b);
}

// Time RK stage 1 of the batch (it writes only the next state buffer and
// tendency slot 0, which stage 1 of the next step overwrites) for each tile
// shape of kStageTiles and keep the fastest.
template <typename T>
int autotune_t(nyl_batch *b, double *ms_out) {
    const uint64_t tune_off = tables_off(b) + al256(b->blob.size());
    double best = 1e300;
    StageTile best_tile = b->tile;
    hipEvent_t e0, e1;
    BCHK(hipEventCreate(&e0));
    BCHK(hipEventCreate(&e1));
    std::vector<double> zeros(b->n, 0.0);
    for (int q = 0; q < b->n; q++) b->host_dt[q] = 0.0;
    BCHK(hipMemcpyAsync(b->ws + b->o_dt, b->host_dt, sizeof(double) * b->n,
                        hipMemcpyHostToDevice, b->stream));
    for (int ci = 0; ci < NYL_N_TILES; ci++) {
        const StageTile &tile = kStageTiles[ci];
        std::vector<std::vector<char>> tabs;
        std::vector<std::pair<int64_t, int>> meta;   // nblk, perio
        std::vector<char> img;
        std::vector<uint64_t> offs;
        for (int perio = 0; perio < 8; perio++) {
            int64_t nblk;
            auto v = stage_items<T>(b, b->phase, 1, perio, tile, nblk);
            if (v.empty()) continue;
            const uint64_t o = al256(img.size());
            img.resize(o + sizeof(StageItem<T>) * v.size());
            memcpy(img.data() + o, v.data(), sizeof(StageItem<T>) * v.size());
            offs.push_back(o);
            meta.push_back({nblk, perio});
            tabs.push_back({});
        }
        BCHK(hipMemcpyAsync(b->ws + tune_off, img.data(), img.size(), hipMemcpyHostToDevice,
                            b->stream));
        double ms_sum = 0.0;
        for (int rep = 0; rep < 4; rep++) {
            BCHK(hipEventRecord(e0, b->stream));
            for (size_t g = 0; g < offs.size(); g++) {
                int nitems = 0;
                for (int q = 0; q < b->n; q++) {
                    nyl_ctx *c = b->m[q];
                    if (c->g.px + 2 * c->g.py + 4 * c->g.pz == meta[g].second) nitems++;
                }
                BCHK(launch_b_stage<T>((const StageItem<T> *)(b->ws + tune_off + offs[g]), nitems,
                                       meta[g].first, 1, meta[g].second, tile, b->mg.dt,
                                       b->stream));
            }
            BCHK(hipEventRecord(e1, b->stream));
            BCHK(hipEventSynchronize(e1));
            float ms = 0.f;
            BCHK(hipEventElapsedTime(&ms, e0, e1));
            if (rep > 0) ms_sum += ms;   // first launch warms up
        }
        const double ms = ms_sum / 3.0;
        if (ms_out) ms_out[ci] = ms;
        if (ms < best) { best = ms; best_tile = tile; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->tile = best_tile;
    // tables for the chosen tile, for the members' current buffers
    return rebuild(b);
}

}  // namespace

extern "C" {

int nyl_batch_create(nyl_ctx *const *members, int32_t n, nyl_batch **out) {
    if (!members || !out || n <= 0) return set_error(NYL_E_ARG, "null argument or empty batch");
    *out = nullptr;
    for (int q = 0; q < n; q++) {
        if (!members[q]) return set_error(NYL_E_ARG, "member %d is null", q);
        if (!members[q]->ws) return set_error(NYL_E_STATE, "member %d has no workspace bound", q);
        if (members[q]->mixed != members[0]->mixed)
            return set_error(NYL_E_ARG, "members must share the precision mode");
        if (members[q]->ntr > 0)
            return set_error(NYL_E_ARG, "member %d has passive tracers: step it on its own", q);
        if (members[q]->d.model != NYL_MODEL_LES || members[q]->d.upwind_a ||
            members[q]->d.upwind_vf)
            return set_error(NYL_E_ARG, "member %d is not a WENO LES model: step it on its own",
                             q);
        for (int r = 0; r < q; r++)
            if (members[r] == members[q]) return set_error(NYL_E_ARG, "member %d appears twice", q);
    }
    nyl_batch *b = new nyl_batch();
    b->n = n;
    b->m.assign(members, members + n);
    b->cur0.assign(n, 0);
    b->mixed = members[0]->mixed;
    b->tabs = members[0]->tabs;
    for (int q = 0; q < n; q++) b->lmax = std::max(b->lmax, members[q]->nlev);
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) { uint64_t o = off; off += al256(bytes); return o; };
    b->o_normb = take(8 * n); b->o_res = take(8 * n); b->o_shift = take(8 * n);
    b->o_tol = take(8 * n); b->o_dt = take(8 * n);
    b->o_nite = take(4 * n); b->o_maxite = take(4 * n); b->o_active = take(4 * n);
    b->o_any = take(8); b->o_acc = take(4 * n);
    if (b->mixed) build_plans<float>(b);
    else build_plans<double>(b);
    b->ws_bytes = tables_off(b) + al256(b->blob.size()) + tune_bytes(b);
    *out = b;
    return NYL_OK;
}

int nyl_batch_destroy(nyl_batch *b) {
    if (!b) return NYL_OK;
    destroy_graphs(b);
    if (b->host_any) (void)hipHostFree(b->host_any);
    if (b->host_dt) (void)hipHostFree(b->host_dt);
    if (b->host_stats) (void)hipHostFree(b->host_stats);
    delete b;
    return NYL_OK;
}

int nyl_batch_workspace_bytes(const nyl_batch *b, uint64_t *bytes) {
    if (!b || !bytes) return set_error(NYL_E_ARG, "null argument");
    *bytes = b->ws_bytes;
    return NYL_OK;
}

int nyl_batch_set_stream(nyl_batch *b, void *stream) {
    if (!b) return set_error(NYL_E_ARG, "null batch");
    b->stream = (hipStream_t)stream;
    return NYL_OK;
}

int nyl_batch_bind_workspace(nyl_batch *b, void *dptr, uint64_t bytes) {
    if (!b || !dptr) return set_error(NYL_E_ARG, "null argument");
    if (bytes < b->ws_bytes)
        return set_error(NYL_E_ARG, "batch workspace of %llu bytes is smaller than the %llu needed",
                         (unsigned long long)bytes, (unsigned long long)b->ws_bytes);
    if (((uintptr_t)dptr) & 255) return set_error(NYL_E_ARG, "workspace must be 256-byte aligned");
    const int n = b->n;
    if (!b->host_any) {
        BCHK(hipHostMalloc((void **)&b->host_any, 2 * sizeof(int), 0));
        BCHK(hipHostMalloc((void **)&b->host_dt, sizeof(double) * n, 0));
        BCHK(hipHostMalloc(&b->host_stats, (sizeof(int) + 2 * sizeof(double)) * n, 0));
    }
    destroy_graphs(b);
    b->ws = (char *)dptr;
    BMG &mg = b->mg;
    mg.normb = (double *)(b->ws + b->o_normb); mg.res = (double *)(b->ws + b->o_res);
    mg.shift = (double *)(b->ws + b->o_shift); mg.tol = (double *)(b->ws + b->o_tol);
    mg.dt = (const double *)(b->ws + b->o_dt);
    mg.nite = (int *)(b->ws + b->o_nite); mg.maxite = (int *)(b->ws + b->o_maxite);
    mg.active = (int *)(b->ws + b->o_active); mg.any = (int *)(b->ws + b->o_any);
    mg.acc = (int *)(b->ws + b->o_acc);
    std::vector<char> ctl(tables_off(b), 0);
    for (int q = 0; q < n; q++) {
        ((double *)(ctl.data() + b->o_tol))[q] = b->m[q]->d.mg_tol;
        ((int *)(ctl.data() + b->o_maxite))[q] = b->m[q]->d.mg_maxite;
    }
    BCHK(hipMemcpyAsync(b->ws, ctl.data(), ctl.size(), hipMemcpyHostToDevice, b->stream));
    BCHK(hipMemcpyAsync(b->ws + tables_off(b), b->blob.data(), b->blob.size(),
                        hipMemcpyHostToDevice, b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    return NYL_OK;
}

int nyl_batch_set_graphs(nyl_batch *b, int on) {
    if (!b) return set_error(NYL_E_ARG, "null batch");
    b->use_graphs = on ? 1 : 0;
    return NYL_OK;
}

int nyl_batch_autotune(nyl_batch *b, int32_t *tile_out, double *ms_out) {
    if (!b || !b->ws) return set_error(NYL_E_STATE, "batch workspace not bound");
    int rc = b->mixed ? autotune_t<float>(b, ms_out) : autotune_t<double>(b, ms_out);
    if (rc) return rc;
    if (tile_out) {
        const StageTile &t = b->tile;
        tile_out[0] = t.tpb;
        tile_out[1] = t.lx < 0 ? 0 : 1 << t.lx;
        tile_out[2] = t.lx < 0 ? 0 : 1 << t.ly;
    }
    return NYL_OK;
}

int nyl_batch_tile_candidate(int32_t index, int32_t *tile_out) {
    if (index < 0 || index >= NYL_N_TILES || !tile_out)
        return set_error(NYL_E_ARG, "tile candidate index out of range");
    const StageTile &t = kStageTiles[index];
    tile_out[0] = t.tpb;
    tile_out[1] = t.lx < 0 ? 0 : 1 << t.lx;
    tile_out[2] = t.lx < 0 ? 0 : 1 << t.ly;
    return NYL_OK;
}

int nyl_batch_set_tile(nyl_batch *b, int32_t cells, int32_t tx, int32_t ty) {
    if (!b || !b->ws) return set_error(NYL_E_STATE, "batch workspace not bound");
    int found = -1;
    for (int q = 0; q < NYL_N_TILES; q++) {
        const StageTile &t = kStageTiles[q];
        const int x = t.lx < 0 ? 0 : 1 << t.lx, y = t.lx < 0 ? 0 : 1 << t.ly;
        if (t.tpb == cells && x == tx && y == ty) found = q;
    }
    if (found < 0)
        return set_error(NYL_E_ARG, "stage tile %d cells %dx%d is not one of the candidates",
                         cells, tx, ty);
    const StageTile &t = kStageTiles[found];
    if (t.tpb == b->tile.tpb && t.lx == b->tile.lx && t.ly == b->tile.ly) return NYL_OK;
    b->tile = t;
    return rebuild(b);
}

int nyl_batch_step_rk3(nyl_batch *b, const double *dt, nyl_mg_stats *stats) {
    if (!b || !b->ws) return set_error(NYL_E_STATE, "batch workspace not bound");
    if (!dt) return set_error(NYL_E_ARG, "null dt");
    const int n = b->n;
    bool stale = false;
    for (int q = 0; q < n; q++) stale = stale || b->m[q]->cur != (b->cur0[q] ^ b->phase);
    if (stale) {   // a member was stepped on its own since the tables were built
        int rc = rebuild(b);
        if (rc) return rc;
    }
    for (int q = 0; q < n; q++) b->host_dt[q] = dt[q];
    BCHK(hipMemcpyAsync(b->ws + b->o_dt, b->host_dt, sizeof(double) * n, hipMemcpyHostToDevice,
                        b->stream));
    int maxite = 0;
    for (int q = 0; q < n; q++) maxite = std::max(maxite, b->m[q]->d.mg_maxite);
    for (int st = 1; st <= 3; st++) {
        const int p = b->phase;
        int rc = run_plan(b, b->pre[p][st - 1], &b->g_pre[p][st - 1]);
        if (rc) return rc;
        // V-cycles: as many as this stage needed last step are launched
        // back to back with no host read in between (a V-cycle plan is a
        // no-op for members the device stop rule has retired), then one
        // read of the active count; further V-cycles, if any member still
        // iterates, are launched one at a time with a read after each
        int any = 0, it = 0;
        for (; it < std::min(b->vcap[st - 1], maxite + 1); it++)
            if ((rc = run_plan(b, b->vcyc, &b->g_vcyc))) return rc;
        if ((rc = read_any(b, &any))) return rc;
        for (; any > 0 && it <= maxite; it++) {
            if ((rc = run_plan(b, b->vcyc, &b->g_vcyc))) return rc;
            if ((rc = read_any(b, &any))) return rc;
        }
        b->vcap[st - 1] = b->host_any[1];
        if ((rc = run_plan(b, b->post[p], &b->g_post[p]))) return rc;
        b->phase ^= 1;
    }
    for (int q = 0; q < n; q++) b->m[q]->cur = b->cur0[q] ^ b->phase;
    int *acc = (int *)b->host_stats;
    double *res = (double *)((char *)b->host_stats + sizeof(int) * n);
    double *nb = res + n;
    BCHK(hipMemcpyAsync(acc, b->mg.acc, sizeof(int) * n, hipMemcpyDeviceToHost, b->stream));
    BCHK(hipMemcpyAsync(res, b->mg.res, sizeof(double) * n, hipMemcpyDeviceToHost, b->stream));
    BCHK(hipMemcpyAsync(nb, b->mg.normb, sizeof(double) * n, hipMemcpyDeviceToHost, b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    for (int q = 0; q < n; q++) {
        b->m[q]->total_cycles += acc[q];
        if (stats) {
            stats[q].nite = acc[q];
            stats[q].res = res[q];
            stats[q].normb = nb[q];
            stats[q].total_cycles = b->m[q]->total_cycles;
        }
    }
    return NYL_OK;
}

}  // extern "C"
