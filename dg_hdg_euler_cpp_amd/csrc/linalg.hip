// Block-Jacobi assembly / batched partial-pivot LU / triangular apply, and the
// reductions and vector kernels of GMRES.
#include "dgh_internal.hpp"

#include <hip/hip_bf16.h>

#include <cfloat>

namespace dgh {
namespace {

constexpr int BLOCK = 256;

unsigned grid_for(long long threads, int block) { return static_cast<unsigned>((threads + block - 1) / block); }

__device__ __forceinline__ int neighbor(int cell, int side, int nx, int ny) {
    int x = cell % nx, y = cell / nx;
    if (side < 2) x = (x + (side == 0 ? nx - 1 : 1)) % nx;
    else y = (y + (side == 2 ? ny - 1 : 1)) % ny;
    return x + nx * y;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-wide sum of `v` (all 256 threads), result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double *smem) {
    v = wave_sum(v);
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    __syncthreads();
    if (lane == 0) smem[wave] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < BLOCK / 64; ++w) t += smem[w];
    return t;
}

// ------------------------------------------------------------- reductions
// Weighted sums of squares over the DG metric: mode 0 x^2/M (dual norm,
// time_integrator.cpp:42-44; discretization.cpp:234-236), mode 1 x^2 M (state
// norm, :45-47), mode 2 plain x^2.
template <int N>
__global__ __launch_bounds__(BLOCK) void k_weighted_sq(const double *__restrict__ x, long long count, int mode,
                                                       const double *__restrict__ weights, double *partials,
                                                       const Basis *__restrict__ B) {
    constexpr int P = N * N;
    __shared__ double smem[BLOCK / 64];
    double acc = 0;
    for (long long i = static_cast<long long>(blockIdx.x) * BLOCK + threadIdx.x; i < count;
         i += static_cast<long long>(gridDim.x) * BLOCK) {
        const double v = x[i];
        const int node = static_cast<int>(i % P);
        acc += mode == 0 ? v * v * B->inv_mass[node]
             : mode == 1 ? v * v * B->mass[node]
             : mode == 3 ? v * v * weights[i]
                         : v * v;
    }
    const double t = block_sum(acc, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// discretization.cpp:237-242: integral of each conserved variable.
template <int N>
__global__ __launch_bounds__(BLOCK) void k_integral(const double *__restrict__ u, long long cells, double *partials,
                                                    const Basis *__restrict__ B) {
    constexpr int P = N * N, M = 4 * P;
    __shared__ double smem[BLOCK / 64];
    double acc[4] = {0, 0, 0, 0};
    for (long long i = static_cast<long long>(blockIdx.x) * BLOCK + threadIdx.x; i < cells * P;
         i += static_cast<long long>(gridDim.x) * BLOCK) {
        const long long cell = i / P;
        const int node = static_cast<int>(i % P);
        for (int v = 0; v < 4; ++v) acc[v] += u[cell * M + v * P + node] * B->mass[node];
    }
    for (int v = 0; v < 4; ++v) {
        const double t = block_sum(acc[v], smem);
        if (threadIdx.x == 0) partials[blockIdx.x * 4 + v] = t;
    }
}

// Second pass: out[k] = sum_b partials[b*width + k], fixed order.
__global__ __launch_bounds__(BLOCK) void k_finish_sum(const double *__restrict__ partials, int nparts, int width,
                                                      double *out) {
    __shared__ double smem[BLOCK / 64];
    for (int k = 0; k < width; ++k) {
        double acc = 0;
        for (int b = threadIdx.x; b < nparts; b += BLOCK) acc += partials[static_cast<long long>(b) * width + k];
        const double t = block_sum(acc, smem);
        if (threadIdx.x == 0) out[k] = t;
    }
}

// ------------------------------------------------------------ block assembly
// dg_operator.cpp:22-48 with volume_jacobian.cpp:72-89, 117-126 for row == col:
// entry (row (v,i,j), col (w,jx,jy)) =
//   [j==jy] -(.5 hy w_jy Ax[v][w](jx,jy)) d(jx,i)
// + [i==jx] (-.5 hx w_jx Ay[v][w](jx,jy)) d(jy,j)
// + [row==col] M/tau
// + [j==jy] (-l_jx H0[v][w](q=j)) l_i + (r_jx H1[v][w]) r_i
// + [i==jx] (-l_jy H2[v][w](q=i)) l_j + (r_jy H3[v][w]) r_j
// with H_s the HLL derivative of the face on side s (face_jacobian.cpp:69-73).
// With HDG = true the face terms are those of HybridOperator::assemble_cell
// (hybrid.cpp:31-55): H_s = w_q ((+-1) A_phys + tau I) on every side with the
// same shapes, added with sign +1 (face_jacobian.cpp:77-82).
template <int N, typename T, bool HDG>
__global__ __launch_bounds__(BLOCK) void k_assemble(const double *__restrict__ vc, const double *__restrict__ fc,
                                                    T *__restrict__ blocks, double inv_step,
                                                    const Basis *__restrict__ B, Geo g, int first) {
    constexpr int P = N * N, M = 4 * P;
    __shared__ double ax[P * 16], ay[P * 16];  // [node][v*4+w]
    __shared__ double hs[4 * N * 16];          // [side][q][v*4+w]
    const int cell = first + blockIdx.x;
    const double gamma = g.gamma, gm = gamma - 1.;
    for (int node = threadIdx.x; node < P; node += BLOCK) {
        const double u = vc[static_cast<long long>(cell) * 3 * P + node];
        const double v = vc[static_cast<long long>(cell) * 3 * P + P + node];
        const double H = vc[static_cast<long long>(cell) * 3 * P + 2 * P + node];
        const double s2 = u * u + v * v;
        double *a = ax + node * 16, *b = ay + node * 16;
        // physics.cpp:42-57, axis 0 then axis 1, row-major [v][w]
        a[0] = 0; a[1] = 1.; a[2] = 0; a[3] = 0;
        a[4] = .5 * (gamma - 3) * u * u + .5 * gm * v * v; a[5] = (3 - gamma) * u; a[6] = -gm * v; a[7] = gm;
        a[8] = -u * v; a[9] = v; a[10] = u; a[11] = 0.;
        a[12] = u * (.5 * gm * s2 - H); a[13] = H - gm * u * u; a[14] = -gm * u * v; a[15] = gamma * u;
        b[0] = 0; b[1] = 0; b[2] = 1.; b[3] = 0;
        b[4] = -u * v; b[5] = v; b[6] = u; b[7] = 0.;
        b[8] = .5 * gm * u * u + .5 * (gamma - 3) * v * v; b[9] = -gm * u; b[10] = (3 - gamma) * v; b[11] = gm;
        b[12] = v * (.5 * gm * s2 - H); b[13] = -gm * u * v; b[14] = H - gm * v * v; b[15] = gamma * v;
    }
    const long long nodes = 2LL * g.cells * N;
    for (int t = threadIdx.x; t < 4 * N; t += BLOCK) {
        const int side = t / N, q = t % N;
        const int face = side == 0 ? neighbor(cell, 0, g.nx, g.ny)
                       : side == 1 ? cell
                       : side == 2 ? g.cells + neighbor(cell, 2, g.nx, g.ny)
                                   : g.cells + cell;
        const long long fn = static_cast<long long>(face) * N + q;
        const bool left = side % 2 != 0;
        const double u = fc[(left ? F_UL : F_UR) * nodes + fn], v = fc[(left ? F_VL : F_VR) * nodes + fn];
        const double H = fc[(left ? F_HL : F_HR) * nodes + fn];
        const double blend = fc[(left ? F_BL : F_BR) * nodes + fn], pen = fc[F_PEN * nodes + fn];
        const double s2 = u * u + v * v;
        double a[16];
        if (side < 2) {
            a[0] = 0; a[1] = 1.; a[2] = 0; a[3] = 0;
            a[4] = .5 * (gamma - 3) * u * u + .5 * gm * v * v; a[5] = (3 - gamma) * u; a[6] = -gm * v; a[7] = gm;
            a[8] = -u * v; a[9] = v; a[10] = u; a[11] = 0.;
            a[12] = u * (.5 * gm * s2 - H); a[13] = H - gm * u * u; a[14] = -gm * u * v; a[15] = gamma * u;
        } else {
            a[0] = 0; a[1] = 0; a[2] = 1.; a[3] = 0;
            a[4] = -u * v; a[5] = v; a[6] = u; a[7] = 0.;
            a[8] = .5 * gm * u * u + .5 * (gamma - 3) * v * v; a[9] = -gm * u; a[10] = (3 - gamma) * v; a[11] = gm;
            a[12] = v * (.5 * gm * s2 - H); a[13] = -gm * u * v; a[14] = H - gm * v * v; a[15] = gamma * v;
        }
        if (HDG) {
            const double tau = fc[(left ? F_TAUL : F_TAUR) * nodes + fn];
            const double weight = .5 * (side < 2 ? g.hy : g.hx) * B->w[q];
            for (int k = 0; k < 16; ++k) a[k] *= left ? 1. : -1.;
            for (int k = 0; k < 4; ++k) a[k * 5] += tau;
            for (int k = 0; k < 16; ++k) a[k] *= weight;
        } else {
            for (int k = 0; k < 16; ++k) a[k] *= blend;
            for (int k = 0; k < 4; ++k) a[k * 5] += left ? pen : -pen;
            // dg_operator.cpp:31: sign -1 on the -x/-y sides (exact, folded into H)
            if (!left)
                for (int k = 0; k < 16; ++k) a[k] = -a[k];
        }
        for (int k = 0; k < 16; ++k) hs[(side * N + q) * 16 + k] = a[k];
    }
    __syncthreads();
    T *out = blocks + static_cast<long long>(blockIdx.x) * M * M;
    const double sx = .5 * g.hy, sy = -.5 * g.hx;
    for (int col = 0; col < M; ++col) {
        const int w = col / P, cnode = col % P, jx = cnode % N, jy = cnode / N;
        for (int row = threadIdx.x; row < M; row += BLOCK) {
            const int v = row / P, rnode = row % P, i = rnode % N, j = rnode / N;
            const int vw = v * 4 + w;
            double val = 0;
            if (N > 1) {
                if (j == jy) val -= (sx * B->w[jy] * ax[cnode * 16 + vw]) * B->wd[jx + N * i];
                if (i == jx) val += (sy * B->w[jx] * ay[cnode * 16 + vw]) * B->wd[jy + N * j];
            }
            if (row == col) val += B->mass[rnode] * inv_step;
            if (j == jy) {
                val += (B->left[jx] * hs[(0 * N + j) * 16 + vw]) * B->left[i];
                val += (B->right[jx] * hs[(1 * N + j) * 16 + vw]) * B->right[i];
            }
            if (i == jx) {
                val += (B->left[jy] * hs[(2 * N + i) * 16 + vw]) * B->left[j];
                val += (B->right[jy] * hs[(3 * N + i) * 16 + vw]) * B->right[j];
            }
            out[static_cast<long long>(col) * M + row] = static_cast<T>(val);
        }
    }
}

// -------------------------------------------------------------- batched LU
__device__ __forceinline__ double readlane_d(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float readlane_d(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Partial-pivot LU of one column-major M x M block per workgroup with Eigen
// PartialPivLU / LAPACK semantics (block_jacobi.hpp:12-23): pivot = max |a| in
// the current column (first row in the current order on ties), rows
// interchanged across the whole matrix, unit lower L. rowsrc[i] = original row
// that ends at row i, so (P b)[i] = b[rowsrc[i]].
//
// Right-looking with NB-column panels. Almost every row of these blocks is
// pivoted, so row interchanges are never applied as strided row swaps:
//  * the panel (rows k0..M-1) is factorised in LDS;
//  * the panel's interchanges, the U12 triangular solve and the trailing update
//    are one pass over each trailing column: a wavefront gathers the column's
//    rows through the panel's row permutation, solves the NB-row U12 segment
//    in registers (readlane broadcasts) and writes the updated column back;
//  * interchanges of later panels are applied to the finished L columns in one
//    final gather pass.
template <int M, int NB, typename T>
__global__ __launch_bounds__(BLOCK) void k_lu(T *__restrict__ blocks, int *__restrict__ rowsrc, int nblocks,
                                              int *flag) {
    constexpr int NA = (M + 63) / 64;  // row chunks per lane
    __shared__ T panel[NB * M];         // [c][r], rows k0..M-1 of the current panel
    __shared__ int perm[M];             // original row held by each logical row
    __shared__ int piv[M];              // row interchanged with row k at step k
    __shared__ int sigma[M];            // gather map of the current panel's interchanges
    __shared__ int tmp[M];
    __shared__ double red_val[BLOCK / 64];
    __shared__ int red_idx[BLOCK / 64];
    __shared__ int bad;
    const int blk = blockIdx.x;
    T *A = blocks + static_cast<long long>(blk) * M * M;
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const double tiny = sizeof(T) == 8 ? DBL_MIN : FLT_MIN;
    for (int r = threadIdx.x; r < M; r += BLOCK) perm[r] = r;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int k0 = 0; k0 < M; k0 += NB) {
        const int nb = min(NB, M - k0), rows = M - k0;
        // (a) panel columns k0..k0+nb-1, rows k0..M-1 (already in the current row order)
        for (int idx = threadIdx.x; idx < nb * rows; idx += BLOCK) {
            const int c = idx / rows, r = k0 + idx % rows;
            panel[c * M + r] = A[static_cast<long long>(k0 + c) * M + r];
        }
        __syncthreads();
        // (b) unblocked factorisation of the panel
        for (int c = 0; c < nb; ++c) {
            const int col = k0 + c;
            double best = -1.0;
            int bidx = M;
            for (int r = col + threadIdx.x; r < M; r += BLOCK) {
                const double a = fabs(static_cast<double>(panel[c * M + r]));
                if (a > best || (a == best && r < bidx)) { best = a; bidx = r; }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double ob = __shfl_xor(best, off, 64);
                const int oi = __shfl_xor(bidx, off, 64);
                if (ob > best || (ob == best && oi < bidx)) { best = ob; bidx = oi; }
            }
            if (lane == 0) { red_val[wave] = best; red_idx[wave] = bidx; }
            __syncthreads();
            if (threadIdx.x == 0) {
                double b = red_val[0];
                int bi = red_idx[0];
                for (int w = 1; w < BLOCK / 64; ++w)
                    if (red_val[w] > b || (red_val[w] == b && red_idx[w] < bi)) { b = red_val[w]; bi = red_idx[w]; }
                if (bi >= M) bi = col;  // all-NaN column: keep the diagonal, flagged below
                piv[col] = bi;
                if (bi != col) { const int t = perm[col]; perm[col] = perm[bi]; perm[bi] = t; }
            }
            __syncthreads();
            const int p = piv[col];
            if (p != col)
                for (int cc = threadIdx.x; cc < nb; cc += BLOCK) {
                    const T t = panel[cc * M + col];
                    panel[cc * M + col] = panel[cc * M + p];
                    panel[cc * M + p] = t;
                }
            __syncthreads();
            const T pivot = panel[c * M + col];
            if (pivot != T(0))
                for (int r = col + 1 + threadIdx.x; r < M; r += BLOCK) panel[c * M + r] /= pivot;
            __syncthreads();
            for (int idx = threadIdx.x; idx < (nb - c - 1) * (M - col - 1); idx += BLOCK) {
                const int cc = c + 1 + idx / (M - col - 1), r = col + 1 + idx % (M - col - 1);
                panel[cc * M + r] -= panel[c * M + r] * panel[cc * M + col];
            }
            __syncthreads();
        }
        // (c) panel write-back; (d) gather map of the panel's interchanges
        for (int idx = threadIdx.x; idx < nb * rows; idx += BLOCK) {
            const int c = idx / rows, r = k0 + idx % rows;
            const T val = panel[c * M + r];
            if (!isfinite(static_cast<double>(val)) || (r == k0 + c && !(fabs(static_cast<double>(val)) > tiny)))
                bad = 1;
            A[static_cast<long long>(k0 + c) * M + r] = val;
        }
        if (threadIdx.x == 0) {
            for (int r = k0; r < M; ++r) sigma[r] = r;
            for (int c = 0; c < nb; ++c) {
                const int a = k0 + c, b = piv[a];
                const int t = sigma[a]; sigma[a] = sigma[b]; sigma[b] = t;
            }
        }
        __syncthreads();
        // (e) per trailing column j: gather rows through sigma, U12 segment by
        //     forward substitution with the unit L11, A22 -= L21 U12, write back.
        for (int j = k0 + nb + wave; j < M; j += BLOCK / 64) {
            T *colp = A + static_cast<long long>(j) * M;
            T x[NA];
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                const int row = k0 + a * 64 + lane;
                x[a] = row < M ? colp[sigma[row]] : T(0);
            }
            T u[NB];
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                u[c] = T(0);
                if (c < nb) {
                    u[c] = readlane_d(x[0], c);
                    if (lane > c && lane < nb) x[0] -= panel[c * M + k0 + lane] * u[c];
                }
            }
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                const int row = k0 + a * 64 + lane;
                if (row >= k0 + nb && row < M) {
                    T acc = x[a];
#pragma unroll
                    for (int c = 0; c < NB; ++c)
                        if (c < nb) acc -= panel[c * M + row] * u[c];
                    x[a] = acc;
                }
            }
            if (lane < nb && !isfinite(static_cast<double>(x[0]))) bad = 1;
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                const int row = k0 + a * 64 + lane;
                if (row < M) colp[row] = x[a];
            }
        }
        __syncthreads();
    }
    // (f) apply the interchanges of later panels to the finished L columns:
    //     tau = gather map of all interchanges after the panel, built backwards.
    for (int r = threadIdx.x; r < M; r += BLOCK) sigma[r] = r;
    __syncthreads();
    const int npanels = (M + NB - 1) / NB;
    for (int q = npanels - 1; q >= 0; --q) {
        const int k0 = q * NB, nb = min(NB, M - k0), lo = k0 + nb;
        if (q < npanels - 1)
            for (int c = k0 + wave; c < k0 + nb; c += BLOCK / 64) {
                T *colp = A + static_cast<long long>(c) * M;
                T x[NA];
#pragma unroll
                for (int a = 0; a < NA; ++a) {
                    const int row = lo + a * 64 + lane;
                    x[a] = row < M ? colp[sigma[row]] : T(0);
                }
#pragma unroll
                for (int a = 0; a < NA; ++a) {
                    const int row = lo + a * 64 + lane;
                    if (row < M) colp[row] = x[a];
                }
            }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int r = k0; r < M; ++r) tmp[r] = r;
            for (int c = 0; c < nb; ++c) {
                const int a = k0 + c, b = piv[a];
                const int t = tmp[a]; tmp[a] = tmp[b]; tmp[b] = t;
            }
        }
        __syncthreads();
        for (int r = k0 + threadIdx.x; r < M; r += BLOCK) {
            const int src = sigma[r];
            sigma[r] = tmp[src];
        }
        __syncthreads();
    }
    for (int r = threadIdx.x; r < M; r += BLOCK) rowsrc[static_cast<long long>(blk) * M + r] = perm[r];
    if (threadIdx.x == 0 && bad) atomicOr(flag, 8);
}


// block_jacobi.hpp:24-39: z = U^-1 L^-1 P r per block; one wavefront per block,
// lane l owns rows l, l+64, ...; the value of each pivot row is broadcast with
// readlane and the factor columns stream once from HBM (column-major: one
// coalesced load per column and row chunk). The loads of column j+1 are issued
// before the FMAs of column j so every wave keeps two columns in flight.
__device__ __forceinline__ float to_compute(__hip_bfloat16 v) { return __bfloat162float(v); }
__device__ __forceinline__ float to_compute(float v) { return v; }
__device__ __forceinline__ double to_compute(double v) { return v; }

// TF: stored factor type (double, float or bfloat16); TV: vector and
// arithmetic type (double for the FP64 path, float for the inner FP32 solve).
template <int M, typename TF, typename TV>
__global__ __launch_bounds__(BLOCK) void k_bj_apply(const TF *__restrict__ factors, const int *__restrict__ rowsrc,
                                                    const TV *__restrict__ r, TV *__restrict__ z, int nblocks) {
    using T = TF;
    constexpr int NA = (M + 63) / 64;
    const int lane = threadIdx.x % 64;
    const long long blk = static_cast<long long>(blockIdx.x) * (BLOCK / 64) + threadIdx.x / 64;
    if (blk >= nblocks) return;
    const T *F = factors + blk * M * M;
    TV x[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const int row = a * 64 + lane;
        x[a] = row < M ? r[blk * M + rowsrc[blk * M + row]] : TV(0);
    }
    // forward substitution with the unit lower L: x_i -= L_ij x_j for i > j
    T cur[NA], nxt[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const int row = a * 64 + lane;
        nxt[a] = (row > 0 && row < M) ? F[row] : T(0);
    }
#pragma unroll
    for (int ja = 0; ja < NA; ++ja) {
        const int jend = min(64, M - ja * 64);
        for (int jl = 0; jl < jend; ++jl) {
            const int j = ja * 64 + jl;
#pragma unroll
            for (int a = ja; a < NA; ++a) cur[a] = nxt[a];
            if (j + 1 < M) {
                const T *colp = F + static_cast<long long>(j + 1) * M;
#pragma unroll
                for (int a = ja; a < NA; ++a) {
                    const int row = a * 64 + lane;
                    nxt[a] = (row > j + 1 && row < M) ? colp[row] : T(0);
                }
            }
            const TV xj = readlane_d(x[ja], jl);
#pragma unroll
            for (int a = ja; a < NA; ++a) x[a] -= static_cast<TV>(to_compute(cur[a])) * xj;
        }
    }
    // backward substitution with U: x_j /= U_jj, x_i -= U_ij x_j for i < j
    {
        const T *colp = F + static_cast<long long>(M - 1) * M;
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            const int row = a * 64 + lane;
            nxt[a] = row < M ? colp[row] : T(0);
        }
    }
#pragma unroll
    for (int ja = NA - 1; ja >= 0; --ja) {
        const int jend = min(64, M - ja * 64);
        for (int jl = jend - 1; jl >= 0; --jl) {
            const int j = ja * 64 + jl;
#pragma unroll
            for (int a = 0; a <= ja; ++a) cur[a] = nxt[a];
            if (j > 0) {
                const T *colp = F + static_cast<long long>(j - 1) * M;
#pragma unroll
                for (int a = 0; a <= ja; ++a) {
                    const int row = a * 64 + lane;
                    nxt[a] = row <= j - 1 ? colp[row] : T(0);
                }
            }
            const TV diag = readlane_d(static_cast<TV>(to_compute(cur[ja])), jl);
            const TV xj = readlane_d(x[ja], jl) / diag;
            if (lane == jl) x[ja] = xj;
#pragma unroll
            for (int a = 0; a <= ja; ++a) {
                const int row = a * 64 + lane;
                if (row < j) x[a] -= static_cast<TV>(to_compute(cur[a])) * xj;
            }
        }
    }
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const int row = a * 64 + lane;
        if (row < M) z[blk * M + row] = x[a];
    }
}

template <typename TS, typename TD>
__global__ void k_convert(const TS *__restrict__ src, TD *__restrict__ dst, long long count) {
    for (long long i = static_cast<long long>(blockIdx.x) * BLOCK + threadIdx.x; i < count;
         i += static_cast<long long>(gridDim.x) * BLOCK)
        dst[i] = static_cast<TD>(src[i]);
}
__global__ void k_float_to_bf16(const float *__restrict__ src, __hip_bfloat16 *__restrict__ dst, long long count) {
    for (long long i = static_cast<long long>(blockIdx.x) * BLOCK + threadIdx.x; i < count;
         i += static_cast<long long>(gridDim.x) * BLOCK)
        dst[i] = __float2bfloat16(src[i]);
}

// ----------------------------------------------------------------- GMRES
constexpr int MAXV = 32;  // restart + 1 <= 32 basis vectors per launch

template <typename T> struct PtrsT { const T *p[MAXV]; };
using Ptrs = PtrsT<double>;

// partials[b][k] = sum_i V_k[i] * (W_i * w_i); W = inv_mass(node) (metric
// weights == nullptr) or weights[i] (gmres.cpp:39-43 projection).
template <int N, typename TV>
__global__ __launch_bounds__(BLOCK) void k_multi_dot(PtrsT<TV> V, int nvec, const TV *__restrict__ w,
                                                     const double *__restrict__ weights, long long count,
                                                     double *partials, const Basis *__restrict__ B) {
    constexpr int P = N * N;
    __shared__ double smem[BLOCK / 64][MAXV];
    double acc[MAXV];
#pragma unroll
    for (int k = 0; k < MAXV; ++k) acc[k] = 0;
    for (long long i = static_cast<long long>(blockIdx.x) * BLOCK + threadIdx.x; i < count;
         i += static_cast<long long>(gridDim.x) * BLOCK) {
        const double ww = (weights ? weights[i] : B->inv_mass[i % P]) * static_cast<double>(w[i]);
#pragma unroll
        for (int k = 0; k < MAXV; ++k)
            if (k < nvec) acc[k] += static_cast<double>(V.p[k][i]) * ww;
    }
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
#pragma unroll
    for (int k = 0; k < MAXV; ++k) {
        if (k < nvec) {
            const double t = wave_sum(acc[k]);
            if (lane == 0) smem[wave][k] = t;
        }
    }
    __syncthreads();
    if (threadIdx.x < nvec) {
        double t = 0;
        for (int wv = 0; wv < BLOCK / 64; ++wv) t += smem[wv][threadIdx.x];
        partials[static_cast<long long>(blockIdx.x) * nvec + threadIdx.x] = t;
    }
}

// w -= sum_k V_k h_k (gmres.cpp:44-45), then mode 0: partial projections of
// the updated w (next CGS pass), mode 1: partial weighted sum of w^2, mode 2: none.
template <int N, typename TV>
__global__ __launch_bounds__(BLOCK) void k_update_dot(PtrsT<TV> V, int nvec, const double *__restrict__ h,
                                                      TV *__restrict__ w, const double *__restrict__ weights,
                                                      long long count, double *partials, int mode,
                                                      const Basis *__restrict__ B) {
    constexpr int P = N * N;
    __shared__ double smem[BLOCK / 64][MAXV];
    double hk[MAXV], acc[MAXV];
#pragma unroll
    for (int k = 0; k < MAXV; ++k) {
        hk[k] = k < nvec ? h[k] : 0.0;
        acc[k] = 0;
    }
    for (long long i = static_cast<long long>(blockIdx.x) * BLOCK + threadIdx.x; i < count;
         i += static_cast<long long>(gridDim.x) * BLOCK) {
        double s = 0;
#pragma unroll
        for (int k = 0; k < MAXV; ++k)
            if (k < nvec) s += static_cast<double>(V.p[k][i]) * hk[k];
        const TV wt = static_cast<TV>(static_cast<double>(w[i]) - s);
        w[i] = wt;
        const double wi = static_cast<double>(wt);
        if (mode != 2) {
            const double ww = (weights ? weights[i] : B->inv_mass[i % P]) * wi;
            if (mode == 0) {
#pragma unroll
                for (int k = 0; k < MAXV; ++k)
                    if (k < nvec) acc[k] += static_cast<double>(V.p[k][i]) * ww;
            } else {
                acc[0] += wi * ww;
            }
        }
    }
    if (mode == 2) return;
    const int width = mode == 0 ? nvec : 1;
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
#pragma unroll
    for (int k = 0; k < MAXV; ++k) {
        if (k < width) {
            const double t = wave_sum(acc[k]);
            if (lane == 0) smem[wave][k] = t;
        }
    }
    __syncthreads();
    if (threadIdx.x < width) {
        double t = 0;
        for (int wv = 0; wv < BLOCK / 64; ++wv) t += smem[wv][threadIdx.x];
        partials[static_cast<long long>(blockIdx.x) * width + threadIdx.x] = t;
    }
}

// out = sum_k V_k y_k (gmres.cpp:75)
template <typename TV, typename TO>
__global__ __launch_bounds__(BLOCK) void k_combine(PtrsT<TV> V, int nvec, const double *__restrict__ y,
                                                   TO *__restrict__ out, long long count, int accumulate) {
    double yk[MAXV];
#pragma unroll
    for (int k = 0; k < MAXV; ++k) yk[k] = k < nvec ? y[k] : 0.0;
    for (long long i = static_cast<long long>(blockIdx.x) * BLOCK + threadIdx.x; i < count;
         i += static_cast<long long>(gridDim.x) * BLOCK) {
        double s = 0;
#pragma unroll
        for (int k = 0; k < MAXV; ++k)
            if (k < nvec) s += static_cast<double>(V.p[k][i]) * yk[k];
        out[i] = accumulate ? static_cast<TO>(static_cast<double>(out[i]) + s) : static_cast<TO>(s);
    }
}

template <typename TX, typename TO>
__global__ __launch_bounds__(BLOCK) void k_scale_div(const TX *__restrict__ x, double d,
                                                     TO *__restrict__ out, long long count) {
    for (long long i = static_cast<long long>(blockIdx.x) * BLOCK + threadIdx.x; i < count;
         i += static_cast<long long>(gridDim.x) * BLOCK)
        out[i] = static_cast<TO>(static_cast<double>(x[i]) / d);
}
__global__ __launch_bounds__(BLOCK) void k_sub_div(const double *__restrict__ x, const double *__restrict__ y,
                                                   double d, double *__restrict__ out, long long count) {
    for (long long i = static_cast<long long>(blockIdx.x) * BLOCK + threadIdx.x; i < count;
         i += static_cast<long long>(gridDim.x) * BLOCK)
        out[i] = (x[i] - y[i]) / d;
}

template <typename T>
__global__ __launch_bounds__(BLOCK) void k_axpby(double a, const T *__restrict__ x, double b,
                                                 const T *__restrict__ y, T *__restrict__ out,
                                                 long long count) {
    for (long long i = static_cast<long long>(blockIdx.x) * BLOCK + threadIdx.x; i < count;
         i += static_cast<long long>(gridDim.x) * BLOCK)
        out[i] = (b == 0.0 || y == nullptr) ? static_cast<T>(a * x[i]) : static_cast<T>(a * x[i] + b * y[i]);
}

}  // namespace

#define DGH_DISPATCH_N(n, ...)                          \
    switch (n) {                                        \
    case 1: { constexpr int N = 1; __VA_ARGS__; } break; \
    case 2: { constexpr int N = 2; __VA_ARGS__; } break; \
    case 3: { constexpr int N = 3; __VA_ARGS__; } break; \
    case 4: { constexpr int N = 4; __VA_ARGS__; } break; \
    case 5: { constexpr int N = 5; __VA_ARGS__; } break; \
    case 6: { constexpr int N = 6; __VA_ARGS__; } break; \
    case 7: { constexpr int N = 7; __VA_ARGS__; } break; \
    case 8: { constexpr int N = 8; __VA_ARGS__; } break; \
    case 9: { constexpr int N = 9; __VA_ARGS__; } break; \
    default: return hipErrorInvalidValue;               \
    }                                                   \
    return hipGetLastError();

// Block widths: DG element blocks 4n^2 and HDG face blocks 4n.
#define DGH_DISPATCH_M(m, ...)                              \
    switch (m) {                                            \
    case 4: { constexpr int M = 4; __VA_ARGS__; } break;    \
    case 8: { constexpr int M = 8; __VA_ARGS__; } break;    \
    case 12: { constexpr int M = 12; __VA_ARGS__; } break;  \
    case 16: { constexpr int M = 16; __VA_ARGS__; } break;  \
    case 20: { constexpr int M = 20; __VA_ARGS__; } break;  \
    case 24: { constexpr int M = 24; __VA_ARGS__; } break;  \
    case 28: { constexpr int M = 28; __VA_ARGS__; } break;  \
    case 32: { constexpr int M = 32; __VA_ARGS__; } break;  \
    case 36: { constexpr int M = 36; __VA_ARGS__; } break;  \
    case 64: { constexpr int M = 64; __VA_ARGS__; } break;  \
    case 100: { constexpr int M = 100; __VA_ARGS__; } break; \
    case 144: { constexpr int M = 144; __VA_ARGS__; } break; \
    case 196: { constexpr int M = 196; __VA_ARGS__; } break; \
    case 256: { constexpr int M = 256; __VA_ARGS__; } break; \
    case 324: { constexpr int M = 324; __VA_ARGS__; } break; \
    default: return hipErrorInvalidValue;                   \
    }                                                       \
    return hipGetLastError();

hipError_t launch_weighted_sq(int n, const double *x, long long count, int mode, const double *weights,
                              double *partials, const Basis *b, hipStream_t s) {
    DGH_DISPATCH_N(n, { k_weighted_sq<N><<<REDUCE_BLOCKS, BLOCK, 0, s>>>(x, count, mode, weights, partials, b); })
}
hipError_t launch_integral(int n, const double *u, long long cells, double *partials, const Basis *b,
                           hipStream_t s) {
    DGH_DISPATCH_N(n, { k_integral<N><<<REDUCE_BLOCKS, BLOCK, 0, s>>>(u, cells, partials, b); })
}
hipError_t launch_finish_sum(const double *partials, int nparts, int width, double *out, hipStream_t s) {
    k_finish_sum<<<1, BLOCK, 0, s>>>(partials, nparts, width, out);
    return hipGetLastError();
}
hipError_t launch_assemble_blocks(int n, const double *vcache, const double *fcache, double *blocks,
                                  double inv_step, const Basis *b, const Geo &g, int first, int count,
                                  hipStream_t s) {
    DGH_DISPATCH_N(n, {
        k_assemble<N, double, false><<<count, BLOCK, 0, s>>>(vcache, fcache, blocks, inv_step, b, g, first);
    })
}
hipError_t launch_assemble_blocks(int n, const double *vcache, const double *fcache, float *blocks,
                                  double inv_step, const Basis *b, const Geo &g, int first, int count,
                                  hipStream_t s) {
    DGH_DISPATCH_N(n, {
        k_assemble<N, float, false><<<count, BLOCK, 0, s>>>(vcache, fcache, blocks, inv_step, b, g, first);
    })
}
hipError_t launch_assemble_hdg(int n, const double *vcache, const double *fcache, double *blocks, double inv_step,
                               const Basis *b, const Geo &g, int first, int count, hipStream_t s) {
    DGH_DISPATCH_N(n, {
        k_assemble<N, double, true><<<count, BLOCK, 0, s>>>(vcache, fcache, blocks, inv_step, b, g, first);
    })
}
hipError_t launch_lu(int m, double *blocks, int *rowsrc, int nblocks, int *flag, hipStream_t s) {
    DGH_DISPATCH_M(m, { k_lu<M, 16, double><<<nblocks, BLOCK, 0, s>>>(blocks, rowsrc, nblocks, flag); })
}
hipError_t launch_lu(int m, float *blocks, int *rowsrc, int nblocks, int *flag, hipStream_t s) {
    DGH_DISPATCH_M(m, { k_lu<M, 16, float><<<nblocks, BLOCK, 0, s>>>(blocks, rowsrc, nblocks, flag); })
}
hipError_t launch_bj_apply(int m, const double *factors, const int *rowsrc, const double *r, double *z,
                           int nblocks, hipStream_t s) {
    DGH_DISPATCH_M(m, {
        k_bj_apply<M, double, double><<<grid_for(nblocks, BLOCK / 64), BLOCK, 0, s>>>(factors, rowsrc, r, z, nblocks);
    })
}
hipError_t launch_bj_apply(int m, const float *factors, const int *rowsrc, const float *r, float *z, int nblocks,
                           hipStream_t s) {
    DGH_DISPATCH_M(m, {
        k_bj_apply<M, float, float><<<grid_for(nblocks, BLOCK / 64), BLOCK, 0, s>>>(factors, rowsrc, r, z, nblocks);
    })
}
hipError_t launch_bj_apply_bf16(int m, const void *factors, const int *rowsrc, const float *r, float *z,
                                int nblocks, hipStream_t s) {
    const auto *f = static_cast<const __hip_bfloat16 *>(factors);
    DGH_DISPATCH_M(m, {
        k_bj_apply<M, __hip_bfloat16, float><<<grid_for(nblocks, BLOCK / 64), BLOCK, 0, s>>>(f, rowsrc, r, z, nblocks);
    })
}
hipError_t launch_convert(const double *src, float *dst, long long count, hipStream_t s) {
    k_convert<double, float><<<4096, BLOCK, 0, s>>>(src, dst, count);
    return hipGetLastError();
}
hipError_t launch_convert(const float *src, double *dst, long long count, hipStream_t s) {
    k_convert<float, double><<<4096, BLOCK, 0, s>>>(src, dst, count);
    return hipGetLastError();
}
hipError_t launch_to_bf16(const float *src, void *dst, long long count, hipStream_t s) {
    k_float_to_bf16<<<4096, BLOCK, 0, s>>>(src, static_cast<__hip_bfloat16 *>(dst), count);
    return hipGetLastError();
}

template <typename T> static PtrsT<T> make_ptrs(const T *const *basis, int nvec) {
    PtrsT<T> p{};
    for (int k = 0; k < nvec && k < MAXV; ++k) p.p[k] = basis[k];
    return p;
}

template <typename TV>
static hipError_t multi_dot_any(int n, const TV *const *basis, int nvec, const TV *w, const double *weights,
                                long long count, double *partials, const Basis *b, hipStream_t s) {
    if (nvec > MAXV) return hipErrorInvalidValue;
    const auto p = make_ptrs(basis, nvec);
    DGH_DISPATCH_N(n, { k_multi_dot<N, TV><<<REDUCE_BLOCKS, BLOCK, 0, s>>>(p, nvec, w, weights, count, partials, b); })
}
hipError_t launch_multi_dot(int n, const double *const *basis, int nvec, const double *w, const double *weights,
                            long long count, double *partials, const Basis *b, hipStream_t s) {
    return multi_dot_any(n, basis, nvec, w, weights, count, partials, b, s);
}
hipError_t launch_multi_dot(int n, const float *const *basis, int nvec, const float *w, const double *weights,
                            long long count, double *partials, const Basis *b, hipStream_t s) {
    return multi_dot_any(n, basis, nvec, w, weights, count, partials, b, s);
}
template <typename TV>
static hipError_t update_dot_any(int n, const TV *const *basis, int nvec, const double *h, TV *w,
                                 const double *weights, long long count, double *partials, int mode,
                                 const Basis *b, hipStream_t s) {
    if (nvec > MAXV) return hipErrorInvalidValue;
    const auto p = make_ptrs(basis, nvec);
    DGH_DISPATCH_N(n, {
        k_update_dot<N, TV><<<REDUCE_BLOCKS, BLOCK, 0, s>>>(p, nvec, h, w, weights, count, partials, mode, b);
    })
}
hipError_t launch_update_dot(int n, const double *const *basis, int nvec, const double *h, double *w,
                             const double *weights, long long count, double *partials, int mode, const Basis *b,
                             hipStream_t s) {
    return update_dot_any(n, basis, nvec, h, w, weights, count, partials, mode, b, s);
}
hipError_t launch_update_dot(int n, const float *const *basis, int nvec, const double *h, float *w,
                             const double *weights, long long count, double *partials, int mode, const Basis *b,
                             hipStream_t s) {
    return update_dot_any(n, basis, nvec, h, w, weights, count, partials, mode, b, s);
}
hipError_t launch_combine(const double *const *basis, int nvec, const double *y, double *out, long long count,
                          hipStream_t s) {
    if (nvec > MAXV) return hipErrorInvalidValue;
    k_combine<double, double><<<4096, BLOCK, 0, s>>>(make_ptrs(basis, nvec), nvec, y, out, count, 0);
    return hipGetLastError();
}
hipError_t launch_combine(const float *const *basis, int nvec, const double *y, float *out, long long count,
                          hipStream_t s) {
    if (nvec > MAXV) return hipErrorInvalidValue;
    k_combine<float, float><<<4096, BLOCK, 0, s>>>(make_ptrs(basis, nvec), nvec, y, out, count, 0);
    return hipGetLastError();
}
hipError_t launch_combine_add(const float *const *basis, int nvec, const double *y, double *out, long long count,
                              hipStream_t s) {
    if (nvec > MAXV) return hipErrorInvalidValue;
    k_combine<float, double><<<4096, BLOCK, 0, s>>>(make_ptrs(basis, nvec), nvec, y, out, count, 1);
    return hipGetLastError();
}
hipError_t launch_axpby(double a, const double *x, double b, const double *y, double *out, long long count,
                        hipStream_t s) {
    k_axpby<double><<<4096, BLOCK, 0, s>>>(a, x, b, y, out, count);
    return hipGetLastError();
}
hipError_t launch_axpby(double a, const float *x, double b, const float *y, float *out, long long count,
                        hipStream_t s) {
    k_axpby<float><<<4096, BLOCK, 0, s>>>(a, x, b, y, out, count);
    return hipGetLastError();
}
hipError_t launch_scale_div(const double *x, double d, double *out, long long count, hipStream_t s) {
    k_scale_div<double, double><<<4096, BLOCK, 0, s>>>(x, d, out, count);
    return hipGetLastError();
}
hipError_t launch_scale_div(const float *x, double d, float *out, long long count, hipStream_t s) {
    k_scale_div<float, float><<<4096, BLOCK, 0, s>>>(x, d, out, count);
    return hipGetLastError();
}
hipError_t launch_scale_div(const double *x, double d, float *out, long long count, hipStream_t s) {
    k_scale_div<double, float><<<4096, BLOCK, 0, s>>>(x, d, out, count);
    return hipGetLastError();
}
hipError_t launch_sub_div(const double *x, const double *y, double d, double *out, long long count, hipStream_t s) {
    k_sub_div<<<4096, BLOCK, 0, s>>>(x, y, d, out, count);
    return hipGetLastError();
}

}  // namespace dgh
