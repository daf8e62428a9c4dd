// Internal declarations shared by the HIP kernel translation units and the
// host driver of libdgh_device.so. Not part of the public C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace dgh {

constexpr int NVAR = 4;
constexpr int MAXN = 9;    // degree <= 8 (reference config.cpp:7)
constexpr int MAXEQ = 21;  // max(20, 2n+3) error/projection points
constexpr int MAXQ = 16;   // over-integration: volume quadrature points per direction

// Reference-element tables of one context (nodal_basis.hpp:34-65), uploaded
// once; kernels read them through uniform (scalar-cache) loads.
struct Basis {
    double wd[MAXN * MAXN];      // weighted derivative d(q, i) at [q + n*i]
    double w[MAXN], x[MAXN], left[MAXN], right[MAXN];
    double mass[MAXN * MAXN];    // .25*hx*hy*w_i*w_j at node i + n*j
    double inv_mass[MAXN * MAXN];
    double proj[MAXN * MAXEQ];   // L2 projection n x EQ, column-major [i + n*q]
    double ev[MAXEQ * MAXN];     // basis at error points EQ x n, [q + EQ*i]
    double ex[MAXEQ], ew[MAXEQ];
    // over-integration (quad Q > n, nodal_basis.hpp:34-65): Gauss weights of
    // the Q-point rule, basis values vq(q, k) and weighted derivatives
    // dq(q, i) = sum_k qw_q vq(q, k) D(k, i), both at [q + Q*k]
    double qw[MAXQ], vq[MAXQ * MAXN], dq[MAXQ * MAXN];
};

// Mesh / physics scalars passed by value to every kernel.
struct Geo {
    int nx, ny, n, cells, eq;
    int nq;        // volume quadrature points per direction (n when collocated)
    int testcase;  // 0 diagonal, 1 wave, 2 axis, 3 seeded
    double hx, hy, lx, ly, gamma, safety, cosine, sine;
};

// Face-node linearisation cache (face_jacobian.cpp:16-37), structure of
// arrays over face nodes f*n + q.
enum FaceField { F_UL, F_VL, F_HL, F_UR, F_VR, F_HR, F_BL, F_BR, F_PEN, F_TAUL, F_TAUR, F_NFIELDS };

// Red-black colouring of an even nx x ny periodic grid (the block ILU(0) row
// order): colour c holds the cells with (x + y) % 2 == c, numbered in natural
// order; its k-th cell sits in row y = k / (nx/2) at x = 2 (k % (nx/2)) + ((y + c) & 1).
__host__ __device__ inline int colour_cell(int k, int colour, int nx) {
    const int half = nx / 2, y = k / half;
    return 2 * (k % half) + ((y + colour) & 1) + nx * y;
}
__host__ __device__ inline int colour_index(int cell, int nx) { return (cell / nx) * (nx / 2) + (cell % nx) / 2; }

// The wavefront's index in its workgroup, marked uniform so that what is
// derived from it (columns, cells, their addresses, loop bounds) stays scalar.
__device__ __forceinline__ int wave_index() {
    return __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / 64));
}

// ------------------------------------------------------------ launchers
// All launchers enqueue on `stream` and return hipError_t of the launch.
// Vectors use the reference layout cell*4n^2 + v*n^2 + (i + n*j).
hipError_t launch_restrict(int n, const double *x, double *traces, const Basis *b, const Geo &g, hipStream_t s);
hipError_t launch_restrict(int n, const float *x, double *traces, const Basis *b, const Geo &g, hipStream_t s);
hipError_t launch_freeze_speeds(int n, const double *traces, double *speeds, int *flag, const Geo &g,
                                hipStream_t s);
hipError_t launch_freeze_rate(int n, const double *u, const double *speeds, unsigned long long *rate_bits,
                              int *flag, const Geo &g, hipStream_t s);
hipError_t launch_residual_flux(int n, const double *traces, const double *speeds, double *flux,
                                const Basis *b, const Geo &g, hipStream_t s);
hipError_t launch_element_residual(int n, const double *u, const double *old, const double *flux, double *r,
                                   double tau, const Basis *b, const Geo &g, hipStream_t s);
hipError_t launch_admissible(int n, const double *u, int *flag, const Basis *b, const Geo &g, hipStream_t s);
hipError_t launch_volume_cache(int n, const double *u, double *vcache, const Geo &g, hipStream_t s);
hipError_t launch_face_cache(int n, const double *traces, const double *speeds, double *fcache,
                             const Basis *b, const Geo &g, hipStream_t s);
hipError_t launch_jv_flux(int n, const double *traces, const double *fcache, double *flux, const Geo &g,
                          hipStream_t s);
hipError_t launch_element_jv(int n, const double *x, const double *vcache, const double *flux, double *y,
                             double inv_step, const Basis *b, const Geo &g, hipStream_t s);
hipError_t launch_element_jv(int n, const float *x, const double *vcache, const double *flux, float *y,
                             double inv_step, const Basis *b, const Geo &g, hipStream_t s);
hipError_t launch_element_jv(int n, const float *x, const double *vcache, const double *flux, double *y,
                             double inv_step, const Basis *b, const Geo &g, hipStream_t s);
// Adjoint: z = J^T w, volume part and the face-flux adjoint.
// Over-integrated volume terms (quad Q = g.nq > n): residual (volume + mass +
// lifts), (u, v, H) cache at the Q^2 points ([cell][3][Q^2]), Jv / J^T w
// element parts, the CFL rate and the point admissibility check.
hipError_t launch_oi_element_residual(int n, const double *u, const double *old, const double *fq, double *r,
                                      double tau, const Basis *b, const Geo &g, hipStream_t s);
hipError_t launch_oi_volume_cache(int n, const double *u, double *vcache, const Basis *b, const Geo &g, hipStream_t s);
hipError_t launch_oi_element_jv(int n, const double *x, const double *vcache, const double *fq, double *y,
                                double inv_step, const Basis *b, const Geo &g, hipStream_t s);
hipError_t launch_oi_element_jv(int n, const float *x, const double *vcache, const double *fq, float *y,
                                double inv_step, const Basis *b, const Geo &g, hipStream_t s);
hipError_t launch_oi_element_jv(int n, const float *x, const double *vcache, const double *fq, double *y,
                                double inv_step, const Basis *b, const Geo &g, hipStream_t s);
hipError_t launch_oi_element_jtv(int n, const double *w, const double *vcache, const double *gtraces, double *z,
                                 double inv_step, const Basis *b, const Geo &g, hipStream_t s);
hipError_t launch_oi_freeze_rate(int n, const double *u, const double *speeds, unsigned long long *rate_bits,
                                 int *flag, const Basis *b, const Geo &g, hipStream_t s);
hipError_t launch_oi_admissible(int n, const double *u, int *flag, const Basis *b, const Geo &g, hipStream_t s);
hipError_t launch_adjoint_face(int n, const double *wtraces, const double *fcache, double *gtraces,
                               const Geo &g, hipStream_t s);
hipError_t launch_element_jtv(int n, const double *w, const double *vcache, const double *gtraces, double *z,
                              double inv_step, const Basis *b, const Geo &g, hipStream_t s);
hipError_t launch_initial(int n, double *u, int *flag, const Basis *b, const Geo &g, hipStream_t s);
hipError_t launch_seeded(int n, double *u, unsigned long long seed, int *flag, const Basis *b, const Geo &g,
                         hipStream_t s);
hipError_t launch_error_partials(int n, const double *u, double time, double *partials, int nblocks,
                                 const Basis *b, const Geo &g, hipStream_t s);

// Reductions (two-pass, fixed order, deterministic).
constexpr int REDUCE_BLOCKS = 1024;
// mode 0: sum x^2 * inv_mass(node); 1: sum x^2 * mass(node); 2: plain sum x^2;
// 3: sum x^2 * weights[i]
hipError_t launch_weighted_sq(int n, const double *x, long long count, int mode, const double *weights,
                              double *partials, const Basis *b, hipStream_t s);
hipError_t launch_integral(int n, const double *u, long long cells, double *partials, const Basis *b,
                           hipStream_t s);
hipError_t launch_finish_sum(const double *partials, int nparts, int width, double *out, hipStream_t s);

// Block Jacobi (DG element blocks m = 4n^2, HDG face blocks 4n).
// Blocks of cells [first, first+count) into blocks[(cell-first) * m*m].
hipError_t launch_assemble_blocks(int n, const double *vcache, const double *fcache, double *blocks,
                                  double inv_step, const Basis *b, const Geo &g, int first, int count,
                                  hipStream_t s);
hipError_t launch_assemble_blocks(int n, const double *vcache, const double *fcache, float *blocks,
                                  double inv_step, const Basis *b, const Geo &g, int first, int count,
                                  hipStream_t s);
// Element blocks of the cells of one colour, compact: blocks[k * m*m] for the
// colour's k-th cell.
hipError_t launch_assemble_colour(int n, const double *vcache, const double *fcache, double *blocks,
                                  double inv_step, const Basis *b, const Geo &g, int colour, hipStream_t s);
hipError_t launch_assemble_colour(int n, const double *vcache, const double *fcache, float *blocks,
                                  double inv_step, const Basis *b, const Geo &g, int colour, hipStream_t s);
hipError_t launch_lu(int m, double *blocks, int *rowsrc, int nblocks, int *flag, hipStream_t s);
// Square LU blocks into the packed layout that k_bj_apply reads (every block
// solved by launch_bj_apply* must be packed), in place, `chunk` blocks at a
// time through `scratch` (chunk * m*m entries).
hipError_t launch_pack(int m, double *blocks, int nblocks, double *scratch, int chunk, hipStream_t s);
hipError_t launch_pack(int m, float *blocks, int nblocks, float *scratch, int chunk, hipStream_t s);
// umax (optional): max |U| of every block, for the FP16 factor scale
hipError_t launch_lu(int m, float *blocks, int *rowsrc, int nblocks, int *flag, hipStream_t s,
                     float *umax = nullptr);
hipError_t launch_bj_apply(int m, const double *factors, const int *rowsrc, const double *r, double *z,
                           int nblocks, hipStream_t s);
hipError_t launch_bj_apply(int m, const float *factors, const int *rowsrc, const float *r, float *z, int nblocks,
                           hipStream_t s);
// Compact factors of one colour's cells; r and z in the full vector layout,
// only that colour's cells read and written.
hipError_t launch_bj_apply_colour(int m, const double *factors, const int *rowsrc, const double *r, double *z,
                                  int nblocks, int colour, int nx, hipStream_t s);
hipError_t launch_bj_apply_colour(int m, const float *factors, const int *rowsrc, const float *r, float *z,
                                  int nblocks, int colour, int nx, hipStream_t s);
// FP64 <-> FP32 vector conversion
hipError_t launch_convert(const double *src, float *dst, long long count, hipStream_t s);
hipError_t launch_convert(const float *src, double *dst, long long count, hipStream_t s);
// 2-byte factors (BF16, or FP16 with a per-block scale of U), converted from
// the square column-major FP32 LU factors into the column-pair layout, or
// into the row-pair layout when rowpairs is set (a comparison path, chosen
// with DGH_BJ_ROWPAIRS=1); conversion, storage and apply must use one layout
bool two_byte_rowpairs_requested();
long long pair_factor_stride(int m, bool rowpairs);
// (FP16 takes max |U| per block from umax, as launch_lu records it)
hipError_t launch_to_pairs(int m, const float *square, void *out, float *sigma, float *rdiag, const float *umax,
                           int nblocks, int fp16, bool rowpairs, int *flag, hipStream_t s);
hipError_t launch_bj_apply_pairs(int m, const void *factors, const float *sigma, const float *rdiag,
                                 const int *rowsrc, const float *r, float *z, int nblocks, int fp16, bool rowpairs,
                                 hipStream_t s);

// GMRES building blocks over vectors of length `count` with the DG metric
// inv_mass(node) (weights == nullptr) or a per-entry weight vector; dot
// products accumulate in FP64 for both vector types.
hipError_t launch_multi_dot(int n, const double *const *basis, int nvec, const double *w, const double *weights,
                            long long count, double *partials, const Basis *b, hipStream_t s);
hipError_t launch_multi_dot(int n, const float *const *basis, int nvec, const float *w, const double *weights,
                            long long count, double *partials, const Basis *b, hipStream_t s);
hipError_t launch_update_dot(int n, const double *const *basis, int nvec, const double *h, double *w,
                             const double *weights, long long count, double *partials, int mode,
                             const Basis *b, hipStream_t s);
hipError_t launch_update_dot(int n, const float *const *basis, int nvec, const double *h, float *w,
                             const double *weights, long long count, double *partials, int mode,
                             const Basis *b, hipStream_t s);
hipError_t launch_combine(const double *const *basis, int nvec, const double *y, double *out, long long count,
                          hipStream_t s);
hipError_t launch_combine(const float *const *basis, int nvec, const double *y, float *out, long long count,
                          hipStream_t s);
// out += sum_k V_k y_k with FP32 basis and FP64 accumulation target
hipError_t launch_combine_add(const float *const *basis, int nvec, const double *y, double *out, long long count,
                              hipStream_t s);
hipError_t launch_axpby(double a, const double *x, double b, const double *y, double *out, long long count,
                        hipStream_t s);
hipError_t launch_axpby(double a, const float *x, double b, const float *y, float *out, long long count,
                        hipStream_t s);
// out = x / d
hipError_t launch_scale_div(const double *x, double d, double *out, long long count, hipStream_t s);
hipError_t launch_scale_div(const float *x, double d, float *out, long long count, hipStream_t s);
hipError_t launch_scale_div(const double *x, double d, float *out, long long count, hipStream_t s);
// out = (x - y) / d
hipError_t launch_sub_div(const double *x, const double *y, double d, double *out, long long count, hipStream_t s);

// Transposed element blocks B^T (adjoint block Jacobi).
hipError_t launch_assemble_blocks_transposed(int n, const double *vcache, const double *fcache, double *blocks,
                                             double inv_step, const Basis *b, const Geo &g, hipStream_t s);
// Graph-captured GMRES: one Arnoldi step's Hessenberg / Givens update on the
// device (H(k, j) from scalars[k] (+ scalars[32 + k] for two Gram-Schmidt
// passes), the squared norm of the next vector in scalars[64]), and
// out = x / *beta with beta in device memory.
hipError_t launch_arnoldi_givens(int j, int restart, const double *scalars, double *H, double *cs, double *sn,
                                 double *g, double *next, double *out, int passes, hipStream_t s);
hipError_t launch_scale_div_dev(const double *x, const double *beta, double *out, long long count, hipStream_t s);
hipError_t launch_scale_div_dev(const float *x, const double *beta, float *out, long long count, hipStream_t s);
// gradient of the density integral over the quadrant x > 0, y > 0
hipError_t launch_quadrant_weights(int n, double *g, const Basis *b, const Geo &geo, hipStream_t s);
// partials[REDUCE_BLOCKS] of sum a_i b_i
hipError_t launch_dot(const double *a, const double *b, long long count, double *partials, hipStream_t s);
// out = (M lambda - k) / tau
hipError_t launch_mass_combine(int n, const double *lambda, const double *k, double tau, double *out, long long count,
                               const Basis *b, hipStream_t s);

// Block ILU(0) of the DG element graph in red-black order (block_ilu.cpp:48-84
// with the symmetric permutation red cells first). Y holds, per red cell, the
// block solve A^-1 E against its 16n lift columns (m x 16n) and S, per red
// cell and side s, the 4n x 4n coupling R_s A^-1 E_s.
hipError_t launch_ilu_coupling(int n, const double *factors_red, const int *rowsrc_red, double *Y, double *S,
                               const Basis *b, const Geo &g, hipStream_t s);
hipError_t launch_ilu_coupling(int n, const float *factors_red, const int *rowsrc_red, float *Y, double *S,
                               const Basis *b, const Geo &g, hipStream_t s);
// z_r = z0_r - A_rr^-1 sum_b A_rb x_b for the red cells through Y = A_rr^-1 E
// (x given by its side traces); black entries of z untouched
hipError_t launch_ilu_red_back(int n, const double *Y, const double *traces, const double *fcache, const double *z0,
                               double *z, const Geo &g, hipStream_t s);
hipError_t launch_ilu_red_back(int n, const float *Y, const double *traces, const double *fcache, const float *z0,
                               float *z, const Geo &g, hipStream_t s);
// A_bb -= sum over the four red neighbours of L_bj U_jb (rank-4n per side)
hipError_t launch_ilu_update(int n, double *blocks_black, const double *S, const double *fcache, const Basis *b,
                             const Geo &g, hipStream_t s);
hipError_t launch_ilu_update(int n, float *blocks_black, const double *S, const double *fcache, const Basis *b,
                             const Geo &g, hipStream_t s);
// out_c = rhs_c - sum over the four neighbours of A_{c,nb} x_nb for the cells of
// one colour, with x given by its side traces
hipError_t launch_ilu_offdiag(int n, int colour, const double *traces, const double *fcache, const double *rhs,
                              double *out, const Basis *b, const Geo &g, hipStream_t s);
hipError_t launch_ilu_offdiag(int n, int colour, const double *traces, const double *fcache, const float *rhs,
                              float *out, const Basis *b, const Geo &g, hipStream_t s);

// Statically condensed HDG (hybrid.cpp:31-121).
hipError_t launch_assemble_hdg(int n, const double *vcache, const double *fcache, double *blocks, double inv_step,
                               const Basis *b, const Geo &g, int first, int count, hipStream_t s);
// With a workspace of hdg_solve_workspace(n, count) doubles the solve runs as
// MFMA tile products on precomputed diagonal-block inverses; with nullptr,
// as the column-streamed substitution.
size_t hdg_solve_workspace(int n, int count);
hipError_t launch_hdg_solve(int n, const double *factors, const int *rowsrc, const double *fcache,
                            const double *resid, double *X, double *Y, int first, int count, const Basis *b,
                            const Geo &g, hipStream_t s, double *workspace);
hipError_t launch_hdg_cx(int n, const double *X, const double *Y, const double *fcache, double *S, double *fres,
                         const Basis *b, const Geo &g, hipStream_t s);
hipError_t launch_hdg_face_sum(int n, const double *base, double base_scale, const double *diag_x, const double *loc,
                               double *out, const Geo &g, hipStream_t s);
hipError_t launch_hdg_diag_x(int n, const double *fcache, const double *x, double *out, const Basis *b, const Geo &g,
                             hipStream_t s);
hipError_t launch_hdg_schur(int n, const double *S, const double *x, double *loc, const Geo &g, hipStream_t s);
hipError_t launch_hdg_face_blocks(int n, const double *S, const double *fcache, double *blocks, const Basis *b,
                                  const Geo &g, hipStream_t s);
hipError_t launch_hdg_recover(int n, const double *X, const double *Y, const double *dl, double *du, const Geo &g,
                              hipStream_t s);
hipError_t launch_hdg_recovered_residual(int n, const double *e, const double *fcache, double *out, const Basis *b,
                                         const Geo &g, hipStream_t s);
hipError_t launch_inverse_trace_mass(int n, double *out, const Basis *b, const Geo &g, hipStream_t s);

}  // namespace dgh
