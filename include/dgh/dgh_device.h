/*
 * dgh_device.h -- C ABI of libdgh_device.so, the MI355X (gfx950) implementation
 * of the implicit 2D tensor-product DG / statically condensed HDG Newton-Krylov
 * hot path of YueWu2002/DG_HDG_Euler_Cpp.
 *
 * The ABI replaces the reference's in-process operator seam (all paths relative
 * to the reference repository):
 *   - Problem residual methods       dg_hdg_2d/include/euler/discretization.hpp:7-45
 *   - DGOperator refresh/apply/block dg_hdg_2d/include/euler/dg_operator.hpp:5-17
 *   - BlockJacobi refresh/apply      common/include/euler/linalg/block_jacobi.hpp:12-39
 *   - GMRES::solve (Action callbacks) common/include/euler/linalg/gmres.hpp:5-35
 *   - HybridOperator                 dg_hdg_2d/include/euler/hybrid.hpp:6-32
 *   - TimeIntegrator::advance        dg_hdg_2d/include/euler/time_integrator.hpp:21-40
 *   - Config parse/validate          dg_hdg_2d/src/config.cpp:6-122
 *
 * Conventions
 *   - Every call returns DGH_OK (0) or DGH_ERROR (1); dgh_last_error() returns the
 *     message, using the reference's exception text where the reference throws
 *     (types.hpp:18-21 `require`). A host wrapper rethrows it as std::runtime_error.
 *   - Vectors are DEVICE pointers (hipMalloc'd or any device allocation) in the
 *     reference layout: volume index cell*4n^2 + var*n^2 + (i + n*j), i the x node
 *     (types.hpp:22-31); trace index face*4n + var*n + q (discretization.cpp:148).
 *   - A context owns every device buffer it needs (caches, factors, Krylov basis)
 *     and one HIP stream; calls are synchronous with respect to the host (each
 *     call waits for its own work before returning). A context is not
 *     thread-safe (the reference operators are not reentrant, DESIGN.md:43 of
 *     the reference).
 *   - Linearisation snapshots follow the reference lifecycle: refresh after the
 *     state, step size or frozen speeds change; apply/block reuse the snapshot.
 */
#ifndef DGH_DEVICE_H
#define DGH_DEVICE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DGH_OK 0
#define DGH_ERROR 1

typedef struct dgh_ctx dgh_ctx;

enum dgh_precision {
    DGH_PRECISION_FP64 = 0,       /* reference arithmetic */
    DGH_PRECISION_MIXED = 1,      /* FP32 block LU + FP32 inner GMRES in FP64 FGMRES */
    DGH_PRECISION_MIXED_BF16 = 2, /* as MIXED with BF16-stored LU factors */
    DGH_PRECISION_MIXED_FP16 = 3  /* as MIXED with FP16-stored LU factors (U scaled by 2^-e per block) */
};
enum dgh_testcase { DGH_DIAGONAL = 0, DGH_WAVE = 1, DGH_AXIS = 2, DGH_SEEDED = 3 };
enum dgh_method { DGH_METHOD_DG = 0, DGH_METHOD_HDG = 1, DGH_METHOD_BOTH = 2 };

/* Mirrors euler::Config (config.hpp:7-18) plus the device-path additions. */
typedef struct dgh_config {
    int nx, ny, degree, quadrature;
    int restart, max_linear, max_newton, steps;
    int newton_steps, warmup_updates;
    double cfl, final_time, safety;
    double nonlinear_tolerance, linear_tolerance;
    int direct, colamd, timing_only, check_update, save_solution;
    int integrator; /* 0 be, 1 ssp-dirk3 */
    int method;     /* enum dgh_method */
    int testcase;   /* enum dgh_testcase */
    int precision;  /* enum dgh_precision */
    int inner_restart;       /* mixed: restart of the FP32 inner GMRES */
    int inner_max;           /* mixed: iteration cap of one inner solve */
    double inner_tolerance;  /* mixed: relative tolerance of one inner solve */
    unsigned long long seed;
    int graph;               /* 1 (default, --graph): each GMRES Arnoldi step, FP64 or the inner FP32
                                solve, replays a captured hipGraph; 0 (--no-graph): host-driven loop */
    int precond;             /* 0 block-jacobi, 1 block-ilu0 (block_ilu.cpp; DG, FP64, red-black order) */
    int ordering;            /* 0 natural, 1 colamd, 2 red-black (block-ilu0 on the device) */
    char prefix[512];
} dgh_config;

/* Mirrors euler::UpdateStatistics (time_integrator.hpp:9-13) plus the
 * mixed-precision error report against the FP64 update. */
typedef struct dgh_update_stats {
    int step, stage, iteration, krylov_iterations;
    double time, stage_time, dt, cfl, residual_physical_rms, damping;
    double linear_relative_residual, dg_linear_relative_residual, iteration_s;
    int inner_iterations, outer_iterations;
    /* mixed precision with check_update: ||du - du_fp64||_M / ||du_fp64||_M of
     * the same Newton system solved by the FP64 path; -1 when not measured */
    double precision_error;
} dgh_update_stats;

/* Mirrors euler::RunStatistics (time_integrator.hpp:14-20) without the vector. */
typedef struct dgh_run_stats {
    double final_time, conservation_drift, seconds, max_linear_residual;
    double max_dg_linear_residual, minimum_step, maximum_step;
    double measured_newton_seconds;
    int steps, newton_iterations, krylov_iterations, measured_updates;
} dgh_run_stats;

/* Reverse sweep of the discrete adjoint: transposed linear solves, one per
 * stage, their GMRES iterations and the largest final relative residual. */
typedef struct dgh_adjoint_stats {
    int krylov_iterations, linear_solves;
    double max_linear_residual;
} dgh_adjoint_stats;

/* ---- configuration (config.cpp:6-122) ---------------------------------- */
size_t dgh_config_size(void);
int dgh_config_default(dgh_config *out);
/* Parses reference CLI flags (argv[0] is skipped) plus --device gpu,
 * --precision fp64|mixed|mixed-bf16|mixed-fp16, --inner-restart N, --inner-max N,
 * --inner-tol R and --seed S. On error
 * writes the reference message into err. */
int dgh_config_parse(int argc, const char *const *argv, dgh_config *out, char *err, size_t errlen);

/* ---- lifetime --------------------------------------------------------- */
/* Problem::Problem (discretization.cpp:4-25) + operator construction. */
int dgh_create(const dgh_config *cfg, dgh_ctx **out);
void dgh_destroy(dgh_ctx *ctx);
const char *dgh_last_error(const dgh_ctx *ctx); /* ctx may be NULL (create errors) */
/* out[0..9]: volume size, trace size, cells, nodes, faces, block width,
 * face width, nx, ny, device bytes currently allocated by the context */
int dgh_dims(const dgh_ctx *ctx, int64_t out[10]);
/* out[0..4]: hx, hy, lx, ly, step_size */
int dgh_geometry(const dgh_ctx *ctx, double out[5]);
/* Gauss nodes x[n] and weights w[n] of the reference element (host buffers;
 * nodal_basis.hpp:13-25). */
int dgh_basis(const dgh_ctx *ctx, double *x, double *w);

/* ---- state set-up ----------------------------------------------------- */
/* Problem::initial (initial_and_error.cpp:4-54), computed on the device. */
int dgh_initial(dgh_ctx *ctx, double *u);
/* Seeded admissible state (no reference counterpart): rho~U[.5,1.5],
 * u,v~U[-.5,.5], p~U[.5,1.5] per node (splitmix64 counters), then the
 * reference's positivity scaling (initial_and_error.cpp:19-51). */
int dgh_seeded(dgh_ctx *ctx, uint64_t seed, double *u);

/* ---- residual side (Problem) ------------------------------------------ */
/* Problem::freeze (discretization.cpp:163-195): HLL speeds + CFL rate; old = u. */
int dgh_freeze(dgh_ctx *ctx, const double *u, double *rate);
/* Problem::step_size and Problem::old (time_integrator.cpp:86-87); old may be NULL. */
int dgh_set_step(dgh_ctx *ctx, double tau, const double *old);
/* Problem::dg_residual (discretization.cpp:123-140). */
int dgh_dg_residual(dgh_ctx *ctx, const double *u, double *r);
/* Problem::admissible (discretization.cpp:211-233). */
int dgh_admissible(dgh_ctx *ctx, const double *u, int *ok);
/* Problem::residual_norm (discretization.cpp:234-236): RMS = ||r||_{M^-1}/sqrt(lx ly). */
int dgh_residual_norm(dgh_ctx *ctx, const double *r, double *rms);
/* TimeIntegrator::dual_norm / state_norm (time_integrator.cpp:42-47). */
int dgh_dual_norm(dgh_ctx *ctx, const double *x, double *out);
int dgh_state_norm(dgh_ctx *ctx, const double *x, double *out);
/* Problem::integral (discretization.cpp:237-242) and Problem::error (initial_and_error.cpp:55-73). */
int dgh_integral(dgh_ctx *ctx, const double *u, double out[4]);
int dgh_error_l2(dgh_ctx *ctx, const double *u, double time, double out[4]);
/* Frozen speeds as [face*n + q][2] = (a, b) into a device buffer. */
int dgh_speeds(dgh_ctx *ctx, double *out);

/* ---- DG linearisation (DGOperator) ------------------------------------ */
/* DGOperator::refresh (dg_operator.cpp:6-9). */
int dgh_dg_refresh(dgh_ctx *ctx, const double *u);
/* DGOperator::apply (dg_operator.cpp:10-21): y = J x. */
int dgh_dg_apply(dgh_ctx *ctx, const double *x, double *y);
/* J^T w (no reference counterpart; the adjoint of dgh_dg_apply). */
int dgh_dg_apply_transpose(dgh_ctx *ctx, const double *w, double *z);
/* DGOperator::block(row, row) (dg_operator.cpp:22-48), column-major m x m, into
 * a HOST buffer; off-diagonal blocks are not materialised on the device. */
int dgh_dg_block(dgh_ctx *ctx, int row, double *out_host);

/* ---- block Jacobi (BlockJacobi) ---------------------------------------- */
/* BlockJacobi::refresh with the DG diagonal blocks (block_jacobi.hpp:12-23),
 * assembled and LU-factorised on the device in the context's precision. */
int dgh_bj_refresh(dgh_ctx *ctx);
/* BlockJacobi::apply (block_jacobi.hpp:24-39): z = P^-1 r. */
int dgh_bj_apply(dgh_ctx *ctx, const double *r, double *z);

/* ---- block ILU(0) (BlockPreconditioner with Preconditioner::BlockILU0) ----- */
/* BlockPreconditioner::refresh + factorize (block_ilu.hpp:14-19,
 * block_ilu.cpp:48-61) on the DG element graph (BlockGraph::cartesian,
 * block_graph.cpp:20-29) in red-black order: the cells with x + y even first.
 * The reference's 2D CLI accepts only block-jacobi (config.cpp:100-104); its
 * 1D solver runs this preconditioner (time_integrator.cpp:29-35 of dg_hdg_1d).
 * Needs even nx, ny >= 4. Shares the factor storage with dgh_bj_refresh. */
int dgh_ilu_refresh(dgh_ctx *ctx);
/* BlockPreconditioner::apply (block_ilu.cpp:62-84): z = (L U)^-1 r; r != z. */
int dgh_ilu_apply(dgh_ctx *ctx, const double *r, double *z);

/* ---- mixed precision (no reference counterpart) ---------------------------- */
/* Block Jacobi factors in FP32 (and 2-byte BF16 / FP16 storage for
 * DGH_PRECISION_MIXED_BF16 / _FP16) from the same element blocks as dgh_bj_refresh. */
int dgh_bj_refresh_mixed(dgh_ctx *ctx);
/* FP32 vectors: z = P32^-1 r and y = J x (Jacobian snapshot of dgh_dg_refresh). */
int dgh_bj_apply_f32(dgh_ctx *ctx, const float *r, float *z);
int dgh_dg_apply_f32(dgh_ctx *ctx, const float *x, float *y);
/* FP64 flexible GMRES(restart) whose preconditioner is an FP32 GMRES
 * (inner_restart, inner_tolerance, inner_max) with the FP32 block Jacobi.
 * iterations = inner iterations, outer = FGMRES iterations. */
int dgh_dg_fgmres(dgh_ctx *ctx, const double *rhs, double *x, int max_iterations, double tolerance,
                  int *iterations, int *outer, double *relative_residual);

/* ---- Krylov ------------------------------------------------------------- */
/* GMRES::solve (gmres.cpp:17-86) with A = DG Jv and M = block Jacobi, metric
 * M^-1, restart from the config. x may alias nothing. If residual != NULL the
 * final true residual b - A x is written there. */
int dgh_dg_gmres(dgh_ctx *ctx, const double *rhs, double *x, int max_iterations, double tolerance,
                 int *iterations, double *relative_residual, double *residual);

/* ---- condensed HDG (HybridOperator) ------------------------------------- */
/* HybridOperator::refresh_projected (hybrid.cpp:28-30, 56-72): linearise at u,
 * condense every element with the DG residual r as volume residual and a
 * zero skeleton residual (element blocks factorised in batches). */
int dgh_hdg_refresh_projected(dgh_ctx *ctx, const double *u, const double *r);
/* HybridOperator::rhs (hybrid.hpp:27): -g + sum_K C_K y_K, trace vector. */
int dgh_hdg_rhs(dgh_ctx *ctx, double *out);
/* HybridOperator::apply (hybrid.cpp:73-82): y = D x + sum_K scatter(S_K gather(x)). */
int dgh_hdg_apply(dgh_ctx *ctx, const double *x, double *y);
/* dgh_hdg_apply without any device allocation (no reference counterpart; for
 * callers that cannot allocate, e.g. inside a serving loop or a region that
 * forbids hipMalloc). Its two intermediate vectors live in a fixed 64 MiB
 * array per GPU, so the mesh must satisfy 2 cells 4n (rounded up to 32) +
 * cells 16n <= 8388608 doubles (e.g. 128 x 128 at k = 8, 256 x 256 at k = 4);
 * larger meshes get DGH_ERROR with the sizes in dgh_last_error. Results equal
 * dgh_hdg_apply bit for bit. Calls on one GPU share that array: calls that
 * overlap in time on one GPU (several host threads, several contexts) race. */
int dgh_hdg_apply_noalloc(dgh_ctx *ctx, const double *x, double *y);
/* HybridOperator::block(face, face) (hybrid.cpp:83-96), column-major 4n x 4n, HOST buffer. */
int dgh_hdg_block(dgh_ctx *ctx, int face, double *out_host);
/* BlockJacobi over the face blocks: refresh (LU of every face block) and apply. */
int dgh_hdg_bj_refresh(dgh_ctx *ctx);
int dgh_hdg_bj_apply(dgh_ctx *ctx, const double *r, double *z);
/* HybridOperator::recover (hybrid.cpp:97-105): du = -X dlambda - y. */
int dgh_hdg_recover(dgh_ctx *ctx, const double *dlambda, double *du);
/* HybridOperator::recovered_residual_from_schur (hybrid.cpp:106-121). */
int dgh_hdg_recovered_residual(dgh_ctx *ctx, const double *schur_error, double *out);
/* GMRES on the Schur system with the face block Jacobi and the inverse face
 * mass metric (time_integrator.cpp:27, 175-176); residual may be NULL. */
int dgh_hdg_gmres(dgh_ctx *ctx, const double *rhs, double *x, int max_iterations, double tolerance,
                  int *iterations, double *relative_residual, double *residual);

/* ---- whole Newton / time loop (TimeIntegrator) -------------------------- */
/* TimeIntegrator::advance (time_integrator.cpp:48-213) on the device for the
 * configured method. u is advanced in place. updates may be NULL. */
int dgh_advance(dgh_ctx *ctx, double *u, dgh_run_stats *run, dgh_update_stats *updates, int capacity,
                int *count);

/* One damped Newton update of the current stage, the body of the Newton loop
 * of TimeIntegrator::advance (time_integrator.cpp:92-179): with the speeds
 * frozen (dgh_freeze) and the stage's step size and base state set
 * (dgh_set_step), evaluates r(u), refreshes the Jacobian and preconditioner,
 * solves for the direction to the reference's linear target (DG or condensed
 * HDG, in the configured precision), line-searches with the admissibility
 * check and overwrites u with the accepted iterate. out (may be NULL) gets
 * the RMS of r(u) before the update, the Krylov count, damping, linear
 * residuals and the update's wall time; step/stage/time fields are 0. */
int dgh_newton_update(dgh_ctx *ctx, double *u, dgh_update_stats *out);

/* ---- discrete adjoint (no reference counterpart) -------------------------- */
/* The reference only linearises forward (Jv); these entry points are the
 * reverse of the same operators, built on dgh_dg_apply_transpose.
 * Transposed block Jacobi: LU of every transposed diagonal block B_i^T, so the
 * apply returns B_i^-T r (shares the factor storage with dgh_bj_refresh). */
int dgh_bj_refresh_transposed(dgh_ctx *ctx);
int dgh_bj_apply_transposed(dgh_ctx *ctx, const double *r, double *z);
/* GMRES on J^T x = rhs with the transposed block Jacobi, metric M^-1. */
int dgh_dg_gmres_transposed(dgh_ctx *ctx, const double *rhs, double *x, int max_iterations, double tolerance,
                            int *iterations, double *relative_residual, double *residual);
/* dgh_advance (BE or DIRK3, --steps N) that also checkpoints every converged
 * stage state and the wave speeds that stage froze, with each step size. */
int dgh_trajectory_record(dgh_ctx *ctx, double *u, dgh_run_stats *run, dgh_update_stats *updates, int capacity,
                          int *count);
/* dgh_advance from a new initial state on the recorded schedule: the recorded
 * step sizes and frozen speeds are reused instead of freezing again. This is
 * the map whose gradient dgh_adjoint_backward returns. */
int dgh_trajectory_replay(dgh_ctx *ctx, double *u, dgh_run_stats *run, dgh_update_stats *updates, int capacity,
                          int *count);
/* Reverse sweep over the recorded trajectory: given final_gradient = dJ/du_N,
 * writes initial_gradient = dJ/du_0; every stage's Newton operator is rebuilt
 * from its checkpoint and solved transposed. */
int dgh_adjoint_backward(dgh_ctx *ctx, const double *final_gradient, double *initial_gradient,
                         dgh_adjoint_stats *stats);
/* Output functional: integral of density over the cells whose centre lies in
 * x > 0, y > 0 (the exact quadrant for even nx, ny), and its gradient (may be
 * NULL). */
int dgh_quadrant_density(dgh_ctx *ctx, const double *u, double *value, double *gradient);

/* ---- timing hooks --------------------------------------------------------- */
/* Accumulated device time (HIP events on the context stream) and launch count
 * of a named kernel group since the last reset: "bj_apply", "jv", "residual",
 * "lu", "assemble", "gmres_ortho". */
int dgh_kernel_time(dgh_ctx *ctx, const char *name, double *total_ms, long long *launches);
int dgh_reset_kernel_times(dgh_ctx *ctx);
int dgh_set_timing(dgh_ctx *ctx, int enabled);

#ifdef __cplusplus
}
#endif
#endif /* DGH_DEVICE_H */
