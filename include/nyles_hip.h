/*
 * nyles_hip.h -- C ABI of libnyles_hip.so, the MI355X (gfx950) time step of
 * the Nyles incompressible Boussinesq LES.
 *
 * The library replaces the native layer the reference's Python host code
 * calls into:
 *   - the f2py stencil modules fortran_upwind, fortran_vortex_force,
 *     fortran_vorticity, fortran_bernoulli, fortran_kinenergy
 *     (core/Makefile:1-12; signatures core/fortran_upwind.f90:3,
 *     core/fortran_vortex_force.f90:10,87, core/fortran_vorticity.f90:2,
 *     core/fortran_bernoulli.f90:2,29,61, core/fortran_kinenergy.f90:3);
 *   - the ctypes ABI of the mgfor multigrid library
 *     (core/mgfordriver.py:5-105 over core/build.py:262-314:
 *     get_ptrmg, solve, set_pyarray, get_pyarray, get_pyshape);
 *   - the one-rank MPI halo exchange (core/mpi/halo.py:123-188).
 *
 * Conventions
 *   - Every entry point returns 0 on success and a negative NYL_E_* code on
 *     failure; nyl_last_error() then describes the failure (thread-local).
 *     Nothing in the library calls exit() or abort().
 *   - The library allocates no device memory.  The caller (the Python host
 *     layer, through PyTorch's allocator) provides one workspace buffer of
 *     nyl_workspace_bytes() bytes, which must outlive the context.
 *   - All work is enqueued on the stream given by nyl_set_stream() (default:
 *     the null stream).  Entry points that return host values
 *     (nyl_reduce, nyl_mg_solve, nyl_diagnose, nyl_step_rk3) synchronise
 *     that stream.  A context must not be used from two host threads at once.
 *   - Fields are stored interior-only, C order [z][y][x] (x contiguous).
 *     The halo of the reference layout is not stored: periodic neighbours
 *     are read at their wrapped index and wall closures are applied in the
 *     kernels, which is value-identical to the reference's filled halos.
 */
#ifndef NYLES_HIP_H
#define NYLES_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NYL_ABI_VERSION 3

/* status codes */
#define NYL_OK 0
#define NYL_E_ARG (-1)      /* invalid argument */
#define NYL_E_STATE (-2)    /* call out of order (e.g. workspace not bound) */
#define NYL_E_HIP (-3)      /* HIP runtime error */
#define NYL_E_SIZE (-4)     /* unsupported grid size */

/* precision modes */
#define NYL_PREC_FP64 0     /* loop-faithful FP64, bit-compatible with the oracle */
/* NYL_PREC_MIXED: model fields (state, tendencies, vor, ke) and their
 * stencils in FP32; the projection is iterative refinement: FP32 V-cycles
 * (FP64 coarsest level) on the correction inside an FP64 outer residual of
 * the right-hand side with its mean removed (summed in FP64),
 * with mgfor's stop rule and V-cycle count (a tolerance tighter than one
 * FP32 inner solve resolves, below 1e-6 of the current squared residual, is
 * met through more outer steps of at most 8 V-cycles); NYL_RED_* reductions
 * accumulate in FP64.  A field must stay below 4 GiB in either mode.  No
 * counterpart in the reference (which is FP64 throughout). */
#define NYL_PREC_MIXED 1

typedef struct nyl_ctx nyl_ctx;

/* Geometry and numerics of one model instance.
 * Mirrors the parameters core/model_les.py:33-95 and core/grid.py:15-53
 * read, plus the mgfor MG_Param defaults (core/mgfor/mg_types.f90:15-26). */
typedef struct nyl_desc {
    int32_t nx, ny, nz;          /* interior cells (global_n* on one rank) */
    int32_t periodic_x, periodic_y, periodic_z;   /* geometry perio_* flags */
    double dx, dy, dz;           /* core/grid.py:30-32 */
    double idx2, idy2, idz2;     /* 1/dx**2 ... as computed by core/grid.py:42-44 */
    double dx2;                  /* dx**2 as computed by core/mgfordriver.py:78 */
    double fparameter;           /* coriolis*dx*dy, 0 if not rotating (core/model_les.py:61-67) */
    int32_t precision;           /* NYL_PREC_* */
    int32_t mg_maxite;           /* 20 (core/mgfor/mg_types.f90:20) */
    double mg_tol;               /* 1e-6 on the squared ratio (core/mgfor/mg_types.f90:21) */
    double mg_omega;             /* 0.9 (core/mgfor/mg_types.f90:22) */
    double viscosity;            /* diff_coef['u'] (core/model_les.py:54-57), 0: none; added
                                    at the last stage (core/model_les.py:140-141) */
    int32_t n_tracers;           /* passive tracers t0.. advected like b
                                    (core/model_les.py:36-43), 0..NYL_MAX_TRACERS */
    int32_t model;               /* NYL_MODEL_*: param.model["modelname"] (core/nyles.py:92-100) */
    int32_t upwind_a;            /* 0: WENO5 advection of b and the tracers (the reference's
                                    build); 1, 3, 5: the linear upwind branch of that order
                                    (core/fortran_upwind.f90:34-67, linear=.false. there) */
    int32_t upwind_vf;           /* the same for the vortex force
                                    (core/fortran_vortex_force.f90:43-65,121-144) */
} nyl_desc;

#define NYL_MAX_TRACERS 8

/* model variants (core/nyles.py:92-100); the variants run each stage as
 * tendencies + stage combination (nyl_rhs, nyl_stage_update) internally */
#define NYL_MODEL_LES 0      /* core/model_les.py */
#define NYL_MODEL_LINEAR 1   /* LES(linear=True): no vortex force, vorticity and ke
                                not diagnosed (core/model_les.py:33-34,122-127,135-136) */
#define NYL_MODEL_EULER 2    /* core/model_les_euler.py: b and the tracers are not
                                advected (:131), Bernoulli without buoyancy
                                (core/bernoulli.py:26-32) */

/* Per-solve statistics (core/mgfor/mg_types.f90:46-49 MG_Stats, plus normb). */
typedef struct nyl_mg_stats {
    int32_t nite;                /* V-cycles of the last solve */
    int32_t total_cycles;        /* V-cycles since context creation */
    double res;                  /* last squared relative residual */
    double normb;                /* squared norm of the right-hand side */
} nyl_mg_stats;

/* field identifiers for nyl_field_offset */
enum {
    NYL_F_B = 0, NYL_F_U, NYL_F_V, NYL_F_W,          /* prognostic state b, u_i, u_j, u_k */
    NYL_F_KE, NYL_F_VORX, NYL_F_VORY, NYL_F_VORZ,    /* diagnostics ke, vor */
    NYL_F_UCX, NYL_F_UCY, NYL_F_UCZ,                 /* contravariant U (core/cov_to_contra.py),
                                                        filled by nyl_apply(NYL_OP_U_FROM_U) */
    NYL_F_P, NYL_F_DIV,                              /* pressure, divergence */
    NYL_F_DS0_B, NYL_F_DS0_U, NYL_F_DS0_V, NYL_F_DS0_W,   /* RK tendencies, slot 0 */
    NYL_F_DS1_B, NYL_F_DS1_U, NYL_F_DS1_V, NYL_F_DS1_W,   /* slot 1 */
    NYL_F_DS2_B, NYL_F_DS2_U, NYL_F_DS2_V, NYL_F_DS2_W,   /* slot 2 */
    NYL_F_MGX,                                       /* MG level-1 solution x (warm start, FP64) */
    NYL_F_MGB,                                       /* MG level-1 right-hand side b (FP64) */
    NYL_F_MGY,                                       /* MG level-1 smoother scratch y */
    NYL_NFIELDS
};
/* passive tracer q (0 <= q < n_tracers): its state, and its tendency in
 * RK slot 0..2 */
#define NYL_F_TRACER(q) (64 + (q))
#define NYL_F_DS_TRACER(slot, q) (80 + NYL_MAX_TRACERS * (slot) + (q))

/* dtype codes returned by nyl_field_offset: model fields are F32 in mixed
 * mode; NYL_F_P, NYL_F_DIV, NYL_F_MGX and NYL_F_MGB are always F64 */
#define NYL_DT_F64 0
#define NYL_DT_F32 1

/* what nyl_reduce computes */
#define NYL_RED_MAXU2 0   /* max over cells of U^2+V^2+W^2 (core/nyles.py:249) */
#define NYL_RED_KE 1      /* sum over the interior of ke evaluated in FP64 from the current
                             u, v, w (core/fortran_kinenergy.f90:47-55) */
#define NYL_RED_BSUM 2    /* sum of b over the interior */
#define NYL_RED_NONFINITE 3  /* number of non-finite values in b,u,v,w */

int nyl_abi_version(void);
const char *nyl_last_error(void);

/* core/mgfordriver.py:6-32 MG.__init__ + core/model_les.py:33-95 LES.__init__ */
int nyl_ctx_create(const nyl_desc *desc, nyl_ctx **out);
int nyl_ctx_destroy(nyl_ctx *ctx);

/* Workspace sizing and binding.  nyl_bind_workspace zero-fills the buffer
 * (the reference starts every array at zero, core/variables.py:146-150). */
int nyl_workspace_bytes(const nyl_ctx *ctx, uint64_t *bytes);
int nyl_bind_workspace(nyl_ctx *ctx, void *dptr, uint64_t bytes);
int nyl_set_stream(nyl_ctx *ctx, void *hip_stream);
/* on (default) = 1: on a non-null stream, the V-cycle below the finest level
 * is captured once into a hipGraph and replayed; 0: direct launches.  No
 * counterpart in the reference. */
int nyl_set_graphs(nyl_ctx *ctx, int on);

/* Byte offset, element count and dtype of a field in the workspace.  The
 * state fields are double-buffered: query again after every step. */
int nyl_field_offset(const nyl_ctx *ctx, int field, uint64_t *offset_bytes,
                     int64_t *nelem, int32_t *dtype);

/* MG hierarchy: number of levels and interior size of a level
 * (core/mgfor/mg_setup.f90:225-307, get_pyshape core/mgfor/pytools.f90:8-15). */
int nyl_mg_levels(const nyl_ctx *ctx, int32_t *nlevels);
int nyl_mg_level_shape(const nyl_ctx *ctx, int level, int32_t *nx, int32_t *ny, int32_t *nz);

/* LES.rhs (core/model_les.py:129-144): tracer advection (b and the passive
 * tracers), vortex force and Bernoulli tendencies of the current state, and
 * with last != 0 the viscosity (core/model_les.py:140-141), into tendency
 * slot 0..2.  The
 * contravariant U is evaluated as u*idx2 ... (the product U_from_u stores,
 * core/cov_to_contra.py:15-17); vor and ke are those of the last diagnose. */
int nyl_rhs(nyl_ctx *ctx, int slot, int last);

/* One RK3_SSP stage (core/timescheme.py:180-221) without the projection:
 * rhs into the stage's slot and the stage update of b,u,v,w.
 * stage 1: s+=c*ds0; 2: s+=c*(ds1-3*ds0); 3: s+=c*(8*ds2-ds0-ds1), with
 * c = dt, dt/4., dt/12. computed by the caller. */
int nyl_rk3_stage(nyl_ctx *ctx, int stage, double c);

/* The stage update of a time scheme from the stored tendencies of every
 * prognostic field (b, u, v, w and the tracers), without recomputing the
 * right-hand side: the update loops of core/timescheme.py:113-221 after
 * LES.rhs and the user forcing (core/model_les.py:143-144) have written the
 * tendencies (nyl_rhs, then the caller's changes).  mode 1..3: RK3_SSP
 * stage k with c = dt, dt/4., dt/12. (tendency slots 0..k-1), mode 1 is
 * also Euler forward; LFAM3: 6 first step (Euler, also saves the state), 4
 * predictor (c2 = 2.*dt), 5 corrector (c = dt): slot 0.  The updated state
 * becomes current. */
int nyl_stage_update(nyl_ctx *ctx, int mode, double c, double c2);

/* LES.diagnose_var (core/model_les.py:97-127): div, MG projection
 * (core/projection.py:38-87, core/mgfordriver.py:66-78), u -= grad p,
 * then vorticity and ke.  U is not stored on the step path (the stage
 * kernel recomputes it from u); nyl_apply(NYL_OP_U_FROM_U) fills NYL_F_UC*. */
int nyl_diagnose(nyl_ctx *ctx, nyl_mg_stats *stats);

/* Timescheme.RK3_SSP forward (core/timescheme.py:180-221) with three
 * diagnose_var calls; stats accumulates the three solves (nite = sum). */
int nyl_step_rk3(nyl_ctx *ctx, double dt, nyl_mg_stats *stats);

/* Timescheme.EulerForward (core/timescheme.py:113-122): rhs, s += dt*ds,
 * diagnose_var. */
int nyl_step_ef(nyl_ctx *ctx, double dt, nyl_mg_stats *stats);

/* Timescheme.LFAM3 (core/timescheme.py:125-177): the first call is an Euler
 * step that also stores the 'before' state; later calls are the
 * leap-frog / Adams-Moulton predictor, diagnose_var, the corrector at n+1/2
 * and diagnose_var (stats.nite sums both solves).  nyl_lfam3_reset makes the
 * next call a first step again. */
int nyl_step_lfam3(nyl_ctx *ctx, double dt, nyl_mg_stats *stats);
int nyl_lfam3_reset(nyl_ctx *ctx);

/* mgfor solve (core/mgfor/solvers.f90:8-33) on the current MG right-hand
 * side NYL_F_MGB, warm-started from NYL_F_MGX. */
int nyl_mg_solve(nyl_ctx *ctx, nyl_mg_stats *stats);

/* Diagnostic pieces, each one reference call:
 *   NYL_OP_U_FROM_U   core/cov_to_contra.py:4-20
 *   NYL_OP_DIV        core/projection.py:15-35 (into NYL_F_DIV and NYL_F_MGB)
 *   NYL_OP_CORRECT    core/projection.py:82-87 with p = x*dx**2
 *   NYL_OP_VORTICITY  core/vorticity.py:7-34
 *   NYL_OP_KINENERGY  core/kinenergy.py:7-24
 *   NYL_OP_PRESSURE   p = x*dx**2 into NYL_F_P (core/mgfordriver.py:78) */
#define NYL_OP_U_FROM_U 0
#define NYL_OP_DIV 1
#define NYL_OP_CORRECT 2
#define NYL_OP_VORTICITY 3
#define NYL_OP_KINENERGY 4
#define NYL_OP_PRESSURE 5
int nyl_apply(nyl_ctx *ctx, int op);

/* Reductions (synchronous). */
int nyl_reduce(nyl_ctx *ctx, int what, double *out);

/* nyl_reduce on caller-owned fields, with no context, no workspace and no
 * device allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under hipGraph capture, tight
 * serving loops).  desc gives the grid (nx, ny, nz) and idx2, idy2, idz2;
 * b, u, v, w are device fields of nx*ny*nz elements of dtype NYL_DT_F64 or
 * NYL_DT_F32 (b only for NYL_RED_BSUM, u, v, w only for NYL_RED_MAXU2 and
 * NYL_RED_KE); the result is written to dev_out[0] (device memory) on
 * `stream`, asynchronously.  The partial sums live in a static device array
 * of 16384 doubles (128 KiB per GPU, held until the process ends),
 * so calls that overlap on one GPU race.  A field must stay below 4 GiB
 * (larger grids are refused with NYL_E_SIZE). */
int nyl_reduce_noalloc(const nyl_desc *desc, int32_t dtype, const void *b, const void *u,
                       const void *v, const void *w, int32_t what, double *dev_out,
                       void *stream);

/* Kernel-class timing with HIP events on the context stream: every launch
 * of the class between start and stop is bracketed by two events; stop
 * synchronises and returns the summed kernel time, the launch count and the
 * summed algorithmic HBM bytes of those launches (each input field read
 * once, each output written once).  No counterpart in the reference; the
 * benchmark uses it to price the dominant kernel against the HBM roofline. */
#define NYL_K_STAGE 1        /* fused rhs + RK update */
#define NYL_K_MG_SWEEP 2     /* one damped-Jacobi sweep */
#define NYL_K_MG_RESTRICT 3  /* residual + restriction */
#define NYL_K_MG_PROLONG 4   /* prolongation */
#define NYL_K_DIAG 5         /* U, vorticity, ke */
#define NYL_K_DIV 6          /* divergence */
#define NYL_K_CORRECT 7      /* u -= grad p */
#define NYL_K_RESNORM 8      /* residual norm (two launches); mixed mode: outer FP64 residual */
#define NYL_K_MG_UPDATE 9    /* mixed mode: x += e of the defect correction */
#define NYL_K_MG_SWEEP_FINE 10  /* the NYL_K_MG_SWEEP launches on the finest level only */
#define NYL_K_MG_SWEEP2 11   /* two sweeps fused in one pass (pre- or post-smoothing) */
#define NYL_K_MG_COARSE 12   /* direct solve of a coarsest level with an odd size */
#define NYL_K_MG_SWEEP2_FINE 13  /* the NYL_K_MG_SWEEP2 launches on the finest level only */
int nyl_prof_start(nyl_ctx *ctx, int kernel_class);
int nyl_prof_stop(nyl_ctx *ctx, double *total_ms, int64_t *launches, double *alg_bytes);

/* Kernel-level checks of the level-1 multigrid kernels (synchronous):
 * 0: one fsmoother3d sweep NYL_F_MGX -> NYL_F_MGY with the per-cell kernel;
 * 1: the same sweep with the tiled kernel; 2: tiled sweep that also reduces
 * the squared residual of NYL_F_MGX (read it with nyl_reduce_last); 3: that
 * residual norm alone; 4: two sweeps NYL_F_MGX -> NYL_F_MGY in one tiled
 * pass; 5: the same with the residual norm of NYL_F_MGX; 6: two sweeps from
 * zero into NYL_F_MGY (NYL_F_MGX not read).  Used by the parity tests; no
 * reference counterpart beyond the sweep itself
 * (core/mgfor/basicoperators.f90:363-400). */
int nyl_debug_mg(nyl_ctx *ctx, int what);
int nyl_reduce_last(nyl_ctx *ctx, double *out);

/* Wait for all work of the context's stream. */
int nyl_sync(nyl_ctx *ctx);

/* ---------------------------------------------------------------------------
 * Ensembles (no counterpart in the reference, which runs one model per
 * process).  A batch advances existing contexts (the members: bound, any
 * grid size and geometry, one precision mode) together by RK3_SSP steps:
 * each RK stage is one persistent stage-kernel launch per periodicity class
 * plus one launch per kernel class for the projection of every member, the
 * multigrid stop rule of each member (core/mgfor/solvers.f90:17-30) runs on
 * the device and every plan is replayed from a hipGraph.  A member's result
 * does not depend on the other members.  The batch workspace (size from
 * nyl_batch_workspace_bytes, 256-byte aligned) holds the item tables; the
 * members keep their own workspaces.  Members may still be used on their
 * own between batch steps.  Mixed-precision members run one FP32 V-cycle per
 * FP64 outer residual (defect correction) in a batch. */
typedef struct nyl_batch nyl_batch;
int nyl_batch_create(nyl_ctx *const *members, int32_t n, nyl_batch **out);
int nyl_batch_destroy(nyl_batch *batch);
int nyl_batch_workspace_bytes(const nyl_batch *batch, uint64_t *bytes);
int nyl_batch_bind_workspace(nyl_batch *batch, void *dptr, uint64_t bytes);
int nyl_batch_set_stream(nyl_batch *batch, void *hip_stream);
/* on = 0: launch the plans directly instead of from hipGraphs */
int nyl_batch_set_graphs(nyl_batch *batch, int on);
/* Stage-kernel tiles: {cells, tx, ty}; tx = ty = 0 is a row tile of
 * `cells` consecutive cells of a member (x fastest), otherwise a brick of
 * tx x ty x (cells / (tx ty)) cells.  NYL_TILE_CANDIDATES shapes of 128,
 * 256 and 512 cells, listed by nyl_batch_tile_candidate(index, tile[3]). */
#define NYL_TILE_CANDIDATES 15
int nyl_batch_tile_candidate(int32_t index, int32_t *tile_out);
/* Times RK stage 1 of the batch for every candidate tile (ms_out in
 * candidate order, NYL_TILE_CANDIDATES entries, may be NULL) and keeps the
 * fastest (tile_out[3], may be NULL).  It overwrites the members' RK3
 * scratch (next state buffer, tendency slot 0) only. */
int nyl_batch_autotune(nyl_batch *batch, int32_t *tile_out, double *ms_out);
/* Sets the stage tile (one of the candidates) without timing. */
int nyl_batch_set_tile(nyl_batch *batch, int32_t cells, int32_t tx, int32_t ty);
/* One RK3_SSP step of every member, member q with time step dt[q];
 * stats[q] (may be NULL): V-cycles of the step, last residual and normb.
 * Work queued on the members' own streams must be complete before the call;
 * the call returns when the step is complete. */
int nyl_batch_step_rk3(nyl_batch *batch, const double *dt, nyl_mg_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* NYLES_HIP_H */
