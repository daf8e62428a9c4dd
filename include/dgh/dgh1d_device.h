/*
 * dgh1d_device.h -- C ABI of the 1D periodic Euler DG path of libdgh_device.so
 * (MI355X, gfx950): the reference's dg_hdg_1d solver on the device.
 *
 * The ABI replaces the reference's 1D in-process operator seam (paths relative
 * to the reference repository):
 *   - Problem initial/freeze/residual/admissible/integral/error
 *                                     dg_hdg_1d/include/euler/discretization.hpp:7-33
 *   - DGOperator refresh/apply/block dg_hdg_1d/include/euler/dg_operator.hpp:5-20
 *   - BlockJacobi refresh/apply      common/include/euler/linalg/block_jacobi.hpp:12-39
 *   - GMRES::solve                   common/include/euler/linalg/gmres.hpp:5-35
 *   - TimeIntegrator::advance        dg_hdg_1d/include/euler/time_integrator.hpp:21-41
 *   - Config parse/validate          dg_hdg_1d/src/config.cpp:6-116
 *
 * Conventions as in dgh_device.h: DGH_OK / DGH_ERROR returns with the
 * reference's message text from dgh1_last_error(); vectors are DEVICE pointers
 * in the reference layout cell*3n + var*n + i (types.hpp:24-41 of dg_hdg_1d);
 * one context owns its buffers and one HIP stream, is synchronous and not
 * thread-safe. The device path runs method dg or hdg with GMRES(restart) and
 * block Jacobi (--linear gmres --precond block-jacobi), the reference's 1D
 * protocol.
 */
#ifndef DGH1D_DEVICE_H
#define DGH1D_DEVICE_H

#include "dgh_device.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dgh1_ctx dgh1_ctx;

/* Mirrors the 1D euler::Config (dg_hdg_1d/include/euler/config.hpp:7-18). */
typedef struct dgh1_config {
    int cells, degree, quadrature;
    int restart, max_linear, max_newton, steps;
    int newton_steps, warmup_updates;
    double cfl, final_time, safety;
    double nonlinear_tolerance, linear_tolerance;
    int integrator; /* 0 be, 1 ssp-dirk3 */
    int method;     /* enum dgh_method; a context runs dg or hdg */
    int timing_only;
    char prefix[512];
    /* initial state (no reference counterpart beside the default): 0 the
     * reference's simple wave (simple_wave.cpp:3-25); 1 `--testcase density`,
     * rho = 1 + 0.2 sin(pi x), u = 2 - rho, p = rho^3 / 3 at t = 0, the
     * simple wave with Riemann invariant u + c = 2 */
    int testcase;
} dgh1_config;

/* parse() (dg_hdg_1d/src/config.cpp:24-116): flags to a validated config.
 * argv excludes the program name. On error returns DGH_ERROR and
 * dgh1_create_error() holds the reference's message. */
int dgh1_parse_config(int argc, const char *const *argv, dgh1_config *out);
/* Problem + DG operator + block Jacobi + GMRES workspaces for one config
 * (discretization.cpp:4-14, time_integrator.cpp:20-37). NULL on error. */
dgh1_ctx *dgh1_create(const dgh1_config *config);
const char *dgh1_create_error(void);
void dgh1_destroy(dgh1_ctx *ctx);
const char *dgh1_last_error(const dgh1_ctx *ctx);
/* number of unknowns cells*3*(k+1) */
long long dgh1_size(const dgh1_ctx *ctx);

/* reference-element Gauss nodes (nodal_basis.hpp:34-45), n = k+1 values into
 * HOST memory, for the node coordinates of the solution output (main.cpp:91-94) */
int dgh1_nodes(const dgh1_ctx *ctx, double *x);
/* the matching Gauss weights (nodal_basis.hpp:34-45), n values into HOST
 * memory, for the mass-weighted DG/HDG difference (main.cpp:86-94) */
int dgh1_weights(const dgh1_ctx *ctx, double *w);
/* Problem::initial (discretization.cpp:31-73): projected simple wave with
 * the cell-mean scaling. */
int dgh1_initial(dgh1_ctx *ctx, double *u);
/* Problem::freeze (discretization.cpp:75-92): old = u, Einfeldt speeds per
 * face, returns the CFL rate. speeds (host, 2*cells, may be NULL) gets a, b. */
int dgh1_freeze(dgh1_ctx *ctx, const double *u, double *rate, double *speeds);
/* step size and stage base of the next residual / linearisation */
int dgh1_set_step(dgh1_ctx *ctx, double tau, const double *old);
/* Problem::dg_residual (discretization.cpp:102-127) */
int dgh1_residual(dgh1_ctx *ctx, const double *u, double *r);
/* Problem::admissible (discretization.cpp:128-144) */
int dgh1_admissible(dgh1_ctx *ctx, const double *u, int *ok);
/* DGOperator::refresh (dg_operator.cpp:7-23, volume_jacobian.cpp:17-37) */
int dgh1_refresh(dgh1_ctx *ctx, const double *u);
/* DGOperator::apply (dg_operator.cpp:24-49): y = J x */
int dgh1_apply(dgh1_ctx *ctx, const double *x, double *y);
/* DGOperator::block(row, row) (dg_operator.cpp:50-79), column-major
 * 3n x 3n into HOST memory */
int dgh1_block(dgh1_ctx *ctx, int row, double *out);
/* BlockJacobi::refresh / apply with the diagonal blocks (block_jacobi.hpp:12-39) */
int dgh1_bj_refresh(dgh1_ctx *ctx);
int dgh1_bj_apply(dgh1_ctx *ctx, const double *r, double *z);
/* GMRES::solve (gmres.cpp:17-86) with the block Jacobi, metric M^-1 */
int dgh1_gmres(dgh1_ctx *ctx, const double *rhs, double *x, int max_iterations, double tolerance,
               int *iterations, double *relative_residual);
/* Problem::error (discretization.cpp:156-168) at time t, three L2 norms */
int dgh1_error(dgh1_ctx *ctx, const double *u, double t, double *error3);
/* Problem::integral (discretization.cpp:148-155) */
int dgh1_integral(dgh1_ctx *ctx, const double *u, double *integral3);
/* ---- condensed HDG (dg_hdg_1d/src/hybrid.cpp); trace vectors are face*3 + var,
 * face f between cells f and f+1 ---------------------------------------- */
/* HybridOperator::refresh_projected (hybrid.cpp:37-81) from the state and its
 * DG residual, plus the face block-Jacobi LUs */
int dgh1_hdg_refresh(dgh1_ctx *ctx, const double *u, const double *r);
/* HybridOperator::rhs(), 3*cells values */
int dgh1_hdg_rhs(dgh1_ctx *ctx, double *rhs);
/* HybridOperator::apply (hybrid.cpp:82-101): the Schur operator */
int dgh1_hdg_apply(dgh1_ctx *ctx, const double *x, double *y);
/* HybridOperator::block(face, face) (hybrid.cpp:102-115), 3x3 column-major HOST */
int dgh1_hdg_block(dgh1_ctx *ctx, int face, double *out);
/* BlockJacobi over the face blocks (block_jacobi.hpp:24-39) */
int dgh1_hdg_bj_apply(dgh1_ctx *ctx, const double *r, double *z);
/* HybridOperator::recover (hybrid.cpp:128-137): du = -y - X dl */
int dgh1_hdg_recover(dgh1_ctx *ctx, const double *dl, double *du);
/* HybridOperator::recovered_residual (hybrid.cpp:116-127) */
int dgh1_hdg_recovered_residual(dgh1_ctx *ctx, const double *dl, double *out);
/* GMRES::solve on the Schur system with the face block Jacobi, unweighted
 * (time_integrator.cpp:36) */
int dgh1_hdg_gmres(dgh1_ctx *ctx, const double *rhs, double *x, int max_iterations, double tolerance,
                   int *iterations, double *relative_residual);
/* TimeIntegrator::advance (time_integrator.cpp:45-207) for the configured
 * method (hdg: condensed HDG with the Schur refinement of :125-153; dg and
 * both: DG): u is advanced in place; updates (host) may be NULL. The CLI runs
 * --method both as one context per method. */
int dgh1_advance(dgh1_ctx *ctx, double *u, dgh_run_stats *run, dgh_update_stats *updates, int capacity,
                 int *count);

#ifdef __cplusplus
}
#endif
#endif
