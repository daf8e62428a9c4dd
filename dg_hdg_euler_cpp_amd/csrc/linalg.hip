// Block-Jacobi assembly / batched partial-pivot LU / triangular apply, and the
// reductions and vector kernels of GMRES.
#include "dgh_internal.hpp"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <cfloat>
#include <climits>
#include <cstdlib>
#include <cstdint>
#include <type_traits>

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
// Face terms of an element block (dg_operator.cpp:22-48 / face_jacobian.cpp
// :34, hybrid.cpp:33-38 for the HDG variant): the 4x4 matrix per side and
// face node, [side][q][v*4+w], computed by the whole block.
template <int N, bool HDG>
__device__ void face_block_terms(int cell, const double *__restrict__ fc, double *hs, const Basis *__restrict__ B,
                                 const Geo &g) {
    const double gamma = g.gamma, gm = gamma - 1.;
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
}

template <int N, typename T, bool HDG, bool TRANSPOSE = false>
__global__ __launch_bounds__(BLOCK) void k_assemble(const double *__restrict__ vc, const double *__restrict__ fc,
                                                    T *__restrict__ blocks, double inv_step,
                                                    const Basis *__restrict__ B, Geo g, int first,
                                                    int colour = -1) {
    constexpr int P = N * N, M = 4 * P;
    __shared__ double ax[P * 16], ay[P * 16];  // [node][v*4+w]
    __shared__ double hs[4 * N * 16];          // [side][q][v*4+w]
    // reference-element tables, read once per entry below
    __shared__ double twd[P], tmass[P], tw[N], tl[N], tr[N];
    for (int k = threadIdx.x; k < P; k += BLOCK) {
        twd[k] = B->wd[k];
        tmass[k] = B->mass[k];
    }
    for (int k = threadIdx.x; k < N; k += BLOCK) {
        tw[k] = B->w[k];
        tl[k] = B->left[k];
        tr[k] = B->right[k];
    }
    const int cell = colour < 0 ? first + static_cast<int>(blockIdx.x) : colour_cell(blockIdx.x, colour, g.nx);
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
    face_block_terms<N, HDG>(cell, fc, hs, B, g);
    __syncthreads();
    T *out = blocks + static_cast<long long>(blockIdx.x) * M * M;
    const double sx = .5 * g.hy, sy = -.5 * g.hx;
    // entry (row, col); a thread keeps its rows' indices across the columns,
    // the column's are uniform
    auto entry = [&](int row, int v, int rnode, int i, int j, int col, int w, int cnode, int jx, int jy) {
        const int vw = v * 4 + w;
        double val = 0;
        if (N > 1) {
            if (j == jy) val -= (sx * tw[jy] * ax[cnode * 16 + vw]) * twd[jx + N * i];
            if (i == jx) val += (sy * tw[jx] * ay[cnode * 16 + vw]) * twd[jy + N * j];
        }
        if (row == col) val += tmass[rnode] * inv_step;
        if (j == jy) {
            val += (tl[jx] * hs[(0 * N + j) * 16 + vw]) * tl[i];
            val += (tr[jx] * hs[(1 * N + j) * 16 + vw]) * tr[i];
        }
        if (i == jx) {
            val += (tl[jy] * hs[(2 * N + i) * 16 + vw]) * tl[j];
            val += (tr[jy] * hs[(3 * N + i) * 16 + vw]) * tr[j];
        }
        return static_cast<T>(val);
    };
    if constexpr (!TRANSPOSE && M % 2 == 0 && sizeof(T) == 8) {
        // FP64 column-major: a thread writes two consecutive rows of a column
        // with one 16-byte store (FP64 blocks: 17.8 -> 15.1 ms at 256^2, k = 8;
        // FP32 blocks are faster with a row per thread, 12.9 against 13.9 ms)
        using T2 = double2;
        for (int pr = threadIdx.x; pr < M / 2; pr += BLOCK) {
            const int r0 = 2 * pr, r1 = r0 + 1;
            const int v0 = r0 / P, n0 = r0 % P, i0 = n0 % N, j0 = n0 / N;
            const int v1 = r1 / P, n1 = r1 % P, i1 = n1 % N, j1 = n1 / N;
            for (int col = 0; col < M; ++col) {
                const int w = col / P, cnode = col % P, jx = cnode % N, jy = cnode / N;
                T2 pair;
                pair.x = entry(r0, v0, n0, i0, j0, col, w, cnode, jx, jy);
                pair.y = entry(r1, v1, n1, i1, j1, col, w, cnode, jx, jy);
                *reinterpret_cast<T2 *>(out + static_cast<long long>(col) * M + r0) = pair;
            }
        }
    } else {
        for (int row = threadIdx.x; row < M; row += BLOCK) {
            const int v = row / P, rnode = row % P, i = rnode % N, j = rnode / N;
            for (int col = 0; col < M; ++col) {
                const int w = col / P, cnode = col % P, jx = cnode % N, jy = cnode / N;
                const T val = entry(row, v, rnode, i, j, col, w, cnode, jx, jy);
                if (TRANSPOSE) out[static_cast<long long>(row) * M + col] = val;
                else out[static_cast<long long>(col) * M + row] = val;
            }
        }
    }
}

// Entry (v, w) of the cached flux Jacobian dF_axis/dU from (u, v, H)
// (physics.cpp:42-57), row-major [v][w].
__device__ __forceinline__ double jacobian_entry(double u, double v, double H, int axis, int r, int c, double gamma) {
    const double gm = gamma - 1., s2 = u * u + v * v;
    if (axis == 0) {
        switch (r * 4 + c) {
        case 1: return 1.;
        case 4: return .5 * (gamma - 3) * u * u + .5 * gm * v * v;
        case 5: return (3 - gamma) * u;
        case 6: return -gm * v;
        case 7: return gm;
        case 8: return -u * v;
        case 9: return v;
        case 10: return u;
        case 12: return u * (.5 * gm * s2 - H);
        case 13: return H - gm * u * u;
        case 14: return -gm * u * v;
        case 15: return gamma * u;
        default: return 0.;
        }
    }
    switch (r * 4 + c) {
    case 2: return 1.;
    case 4: return -u * v;
    case 5: return v;
    case 6: return u;
    case 8: return .5 * gm * u * u + .5 * (gamma - 3) * v * v;
    case 9: return -gm * u;
    case 10: return (3 - gamma) * v;
    case 11: return gm;
    case 12: return v * (.5 * gm * s2 - H);
    case 13: return -gm * u * v;
    case 14: return H - gm * v * v;
    case 15: return gamma * v;
    default: return 0.;
    }
}

// Element block with an over-integrated volume part (volume_jacobian.cpp
// :90-116 + dg_operator.cpp:22-48): per (v, w) pair, the line tables
// lx[qb][i + n k] = -sum_qa d(qa, i) (A_x(qa, qb)[v][w] vq(qa, k)) and
// ly[qb][i + n k] = -sum_qa (vq(qa, i) qw_qa) (A_y(qa, qb)[v][w] vq(qa, k)),
// then entry (v,i,j; w,k,l) = sum_qb cx(qb, j, l) lx + cy(qb, j, l) ly with
// cx = hy/2 qw_qb vq(qb, j) vq(qb, l), cy = hx/2 d(qb, j) vq(qb, l); the mass
// diagonal and the face terms follow as in k_assemble.
template <int N, typename T, bool HDG, bool TRANSPOSE = false>
__global__ __launch_bounds__(BLOCK) void k_assemble_oi(const double *__restrict__ vc, const double *__restrict__ fc,
                                                       T *__restrict__ blocks, double inv_step,
                                                       const Basis *__restrict__ B, Geo g, int first,
                                                       int colour = -1) {
    constexpr int P = N * N, M = 4 * P;
    __shared__ double uvh[3 * MAXQ * MAXQ], hs[4 * N * 16], lx[MAXQ * P], ly[MAXQ * P];
    const int cell = colour < 0 ? first + static_cast<int>(blockIdx.x) : colour_cell(blockIdx.x, colour, g.nx);
    const int Q = g.nq, QQ = Q * Q;
    for (int t = threadIdx.x; t < 3 * QQ; t += BLOCK) uvh[t] = vc[static_cast<long long>(cell) * 3 * QQ + t];
    face_block_terms<N, HDG>(cell, fc, hs, B, g);
    T *out = blocks + static_cast<long long>(blockIdx.x) * M * M;
    for (int vw = 0; vw < 16; ++vw) {
        const int v = vw / 4, w = vw % 4;
        __syncthreads();
        for (int t = threadIdx.x; t < Q * P; t += BLOCK) {
            const int qb = t / P, r = t % P, i = r % N, k = r / N;
            double sx = 0, sy = 0;
            for (int qa = 0; qa < Q; ++qa) {
                const int q = qb * Q + qa;
                const double uu = uvh[q], vv = uvh[QQ + q], H = uvh[2 * QQ + q];
                const double wx = jacobian_entry(uu, vv, H, 0, v, w, g.gamma) * B->vq[qa + Q * k];
                const double wy = jacobian_entry(uu, vv, H, 1, v, w, g.gamma) * B->vq[qa + Q * k];
                sx += B->dq[qa + Q * i] * wx;
                sy += (B->vq[qa + Q * i] * B->qw[qa]) * wy;
            }
            lx[t] = -sx;
            ly[t] = -sy;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < P * P; t += BLOCK) {
            const int rnode = t % P, cnode = t / P;
            const int i = rnode % N, j = rnode / N, kx = cnode % N, ly_ = cnode / N;
            double val = 0;
            for (int qb = 0; qb < Q; ++qb) {
                const double cx = .5 * g.hy * B->qw[qb] * B->vq[qb + Q * j] * B->vq[qb + Q * ly_];
                const double cy = .5 * g.hx * B->dq[qb + Q * j] * B->vq[qb + Q * ly_];
                val += cx * lx[qb * P + i + N * kx] + cy * ly[qb * P + i + N * kx];
            }
            const int row = v * P + rnode, col = w * P + cnode;
            if (row == col) val += B->mass[rnode] * inv_step;
            if (j == ly_) {
                val += (B->left[kx] * hs[(0 * N + j) * 16 + vw]) * B->left[i];
                val += (B->right[kx] * hs[(1 * N + j) * 16 + vw]) * B->right[i];
            }
            if (i == kx) {
                val += (B->left[ly_] * hs[(2 * N + i) * 16 + vw]) * B->left[j];
                val += (B->right[ly_] * hs[(3 * N + i) * 16 + vw]) * B->right[j];
            }
            if (TRANSPOSE) out[static_cast<long long>(row) * M + col] = static_cast<T>(val);
            else out[static_cast<long long>(col) * M + row] = static_cast<T>(val);
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

// Wavefront max / min through DPP row shifts and row broadcasts instead of
// LDS permutes: after row_shr 1, 2, 4, 8 lane 15 of each row holds its row's
// result; row_bcast 15 and 31 carry it into rows 1 and 3, so lane 63 ends with
// the result over all 64 lanes. Lanes without a source keep the identity.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_dmax(double v, double identity) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(identity), __double2loint(v), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(identity), __double2hiint(v), CTRL, ROWS, 0xf, false);
    return fmax(v, __hiloint2double(hi, lo));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_imin(int v) {
    return min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, CTRL, ROWS, 0xf, false));
}
// Pivot rule of one wavefront: the largest `best` (>= -1, never NaN), the
// smallest `idx` among the lanes holding it; the result in every lane.
__device__ __forceinline__ void wave_argmax(double &best, int &idx) {
    double m = best;
    m = dpp_dmax<0x111, 0xf>(m, -1.0);
    m = dpp_dmax<0x112, 0xf>(m, -1.0);
    m = dpp_dmax<0x114, 0xf>(m, -1.0);
    m = dpp_dmax<0x118, 0xf>(m, -1.0);
    m = dpp_dmax<0x142, 0xa>(m, -1.0);
    m = dpp_dmax<0x143, 0xc>(m, -1.0);
    const double mx = readlane_d(m, 63);
    int k = best == mx ? idx : INT_MAX;
    k = dpp_imin<0x111, 0xf>(k);
    k = dpp_imin<0x112, 0xf>(k);
    k = dpp_imin<0x114, 0xf>(k);
    k = dpp_imin<0x118, 0xf>(k);
    k = dpp_imin<0x142, 0xa>(k);
    k = dpp_imin<0x143, 0xc>(k);
    best = mx;
    idx = __builtin_amdgcn_readlane(k, 63);
}

// Partial-pivot LU of one column-major M x M block per workgroup with Eigen
// PartialPivLU / LAPACK semantics (block_jacobi.hpp:12-23): pivot = max |a| in
// the current column (first row in the current order on ties), rows
// interchanged across the whole matrix, unit lower L. rowsrc[i] = original row
// that ends at row i, so (P b)[i] = b[rowsrc[i]].
//
// Right-looking with NB-column panels. Almost every row of these blocks is
// pivoted, so row interchanges are never applied as strided row swaps:
//  * the panel (rows k0..M-1) is factorised in LDS (lu_panel);
//  * the panel's interchanges, the U12 triangular solve and the trailing update
//    are one pass over each trailing column: the column's rows are gathered
//    through the panel's row permutation, the NB-row U12 segment is solved in
//    registers and the updated column is written back. k_lu does this for CB
//    columns per wavefront pass with vector FMAs and L21 read from LDS (each
//    LDS read feeds CB FMAs); k_lu_mfma does it for 16-column tiles with the
//    A22 -= L21 U12 product on MFMA;
//  * interchanges of later panels are applied to the finished L columns in one
//    final gather pass (lu_finish).
// Row indices in LDS are 16-bit: with the panel they keep an FP32 block at
// about 25 KB, six workgroups per CU.
using lu_row = short;
template <int M, int NB, typename T, int LT> struct LuShared {
    static_assert(M <= 32767, "row indices are 16-bit");
    T panel[NB * M];    // [c][r], rows k0..M-1 of the current panel
    lu_row perm[M];     // original row held by each logical row
    lu_row piv[M];      // row interchanged with row k at step k
    lu_row sigma[M];    // gather map of the current panel's interchanges
    double red_val[LT / 64];
    int red_idx[LT / 64];
    int bad;
    unsigned umax;  // bits of max |U| over the finished U entries (non-negative floats order as integers)
};

template <int M, int NB, typename T, int LT> __device__ void lu_begin(LuShared<M, NB, T, LT> &S) {
    for (int r = threadIdx.x; r < M; r += LT) S.perm[r] = r;
    if (threadIdx.x == 0) {
        S.bad = 0;
        S.umax = 0;
    }
    __syncthreads();
}

// Gather map of one panel's interchanges (row k0 + c with row piv[k0 + c],
// c = 0..nb-1 in order): the place whose row ends at place r, found by undoing
// the interchanges from the last one back.
__device__ __forceinline__ int lu_undo_interchanges(const lu_row *piv, int k0, int nb, int r) {
    for (int c = nb - 1; c >= 0; --c) {
        const int a = k0 + c, b = piv[a];
        r = r == a ? b : (r == b ? a : r);
    }
    return r;
}

// Panel columns k0..k0+nb-1 (rows k0..M-1, already in the current row order):
// load into LDS, unblocked factorisation with partial pivoting, write back,
// and the gather map sigma of the panel's interchanges. Ends synchronised.
// Two barriers per column: the pivot search of column c + 1 runs inside the
// scale-and-update pass of column c (each thread owns its rows), and every
// thread combines the per-wave maxima itself.
template <int M, int NB, typename T, int LT>
__device__ void lu_panel(T *__restrict__ A, LuShared<M, NB, T, LT> &S, int k0, int nb) {
    const int lane = threadIdx.x % 64, wave = wave_index();
    const double tiny = sizeof(T) == 8 ? DBL_MIN : FLT_MIN;
    T *panel = S.panel;
    // each thread's rows, all columns' loads in flight together
#pragma unroll
    for (int j = 0; j < (M + LT - 1) / LT; ++j) {
        const int r = k0 + threadIdx.x + j * LT;
        if (r < M) {
            T v[NB];
#pragma unroll
            for (int c = 0; c < NB; ++c) v[c] = c < nb ? A[static_cast<long long>(k0 + c) * M + r] : T(0);
#pragma unroll
            for (int c = 0; c < NB; ++c)
                if (c < nb) panel[c * M + r] = v[c];
        }
    }
    __syncthreads();
    // max |a| of the next column over this thread's rows, first row on ties
    double best = -1.0;
    int bidx = M;
    auto consider = [&](double a, int r) {
        if (a > best || (a == best && r < bidx)) {
            best = a;
            bidx = r;
        }
    };
    for (int r = k0 + threadIdx.x; r < M; r += LT) consider(fabs(static_cast<double>(panel[r])), r);
    for (int c = 0; c < nb; ++c) {
        const int col = k0 + c;
        wave_argmax(best, bidx);
        if (lane == 0) { S.red_val[wave] = best; S.red_idx[wave] = bidx; }
        __syncthreads();
        double b = S.red_val[0];
        int p = S.red_idx[0];
        for (int w = 1; w < LT / 64; ++w)
            if (S.red_val[w] > b || (S.red_val[w] == b && S.red_idx[w] < p)) { b = S.red_val[w]; p = S.red_idx[w]; }
        if (p >= M) p = col;  // all-NaN column: keep the diagonal, flagged below
        if (threadIdx.x == 0) {
            S.piv[col] = p;
            if (p != col) { const lu_row t = S.perm[col]; S.perm[col] = S.perm[p]; S.perm[p] = t; }
        }
        if (p != col)
            for (int cc = threadIdx.x; cc < nb; cc += LT) {
                const T t = panel[cc * M + col];
                panel[cc * M + col] = panel[cc * M + p];
                panel[cc * M + p] = t;
            }
        __syncthreads();
        // scale column c, update columns c+1.., search column c+1 (rows > col)
        const T pivot = panel[c * M + col];
        best = -1.0;
        bidx = M;
        for (int r = col + 1 + threadIdx.x; r < M; r += LT) {
            T l = panel[c * M + r];
            if (pivot != T(0)) {
                l /= pivot;
                panel[c * M + r] = l;
            }
            for (int cc = c + 1; cc < nb; ++cc) panel[cc * M + r] -= l * panel[cc * M + col];
            if (c + 1 < nb) consider(fabs(static_cast<double>(panel[(c + 1) * M + r])), r);
        }
    }
    __syncthreads();
    float um = 0.f;  // max |U11| of this thread's entries
#pragma unroll
    for (int j = 0; j < (M + LT - 1) / LT; ++j) {
        const int r = k0 + threadIdx.x + j * LT;
        if (r >= M) continue;
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            if (c >= nb) continue;
            const T val = panel[c * M + r];
            if (!isfinite(static_cast<double>(val)) || (r == k0 + c && !(fabs(static_cast<double>(val)) > tiny)))
                S.bad = 1;
            if (r <= k0 + c) um = fmaxf(um, static_cast<float>(fabs(static_cast<double>(val))));
            A[static_cast<long long>(k0 + c) * M + r] = val;
        }
        S.sigma[r] = lu_undo_interchanges(S.piv, k0, nb, r);
    }
    if (um > 0.f) atomicMax(&S.umax, __float_as_uint(um));
    __syncthreads();
}

// Interchanges of later panels applied to the finished L columns (sigma here:
// the gather map of all interchanges after the panel, composed from the last
// panel back), then rowsrc and the singularity flag.
// umax (optional): max |U| per block, complete only when every finished U
// entry went through lu_panel (U11) or the caller's trailing pass (U12).
template <int M, int NB, typename T, int LT>
__device__ void lu_finish(T *__restrict__ A, LuShared<M, NB, T, LT> &S, int *__restrict__ rowsrc, int *flag,
                          float *umax = nullptr) {
    constexpr int NA = (M + 63) / 64;
    constexpr int G = (NB + LT / 64 - 1) / (LT / 64);  // panel columns per wavefront
    const int lane = threadIdx.x % 64, wave = wave_index();
    lu_row *sigma = S.sigma;
    for (int r = threadIdx.x; r < M; r += LT) sigma[r] = r;
    __syncthreads();
    const int npanels = (M + NB - 1) / NB;
    for (int q = npanels - 1; q >= 0; --q) {
        const int k0 = q * NB, nb = min(NB, M - k0), lo = k0 + nb;
        if (q < npanels - 1) {
            // all of this wavefront's columns gathered before any is written
            T x[G][NA];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int c = k0 + wave + g * (LT / 64);
                const T *colp = A + static_cast<long long>(c < lo ? c : k0) * M;
#pragma unroll
                for (int a = 0; a < NA; ++a) {
                    const int row = lo + a * 64 + lane;
                    x[g][a] = c < lo && row < M ? colp[sigma[row]] : T(0);
                }
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int c = k0 + wave + g * (LT / 64);
                if (c >= lo) continue;
                T *colp = A + static_cast<long long>(c) * M;
#pragma unroll
                for (int a = 0; a < NA; ++a) {
                    const int row = lo + a * 64 + lane;
                    if (row < M) colp[row] = x[g][a];
                }
            }
        }
        __syncthreads();
        // compose with this panel's interchanges (every sigma[r] lies in k0..M-1)
        for (int r = k0 + threadIdx.x; r < M; r += LT) sigma[r] = lu_undo_interchanges(S.piv, k0, nb, sigma[r]);
        __syncthreads();
    }
    for (int r = threadIdx.x; r < M; r += LT) rowsrc[static_cast<long long>(blockIdx.x) * M + r] = S.perm[r];
    if (threadIdx.x == 0 && S.bad) atomicOr(flag, 8);
    if (threadIdx.x == 0 && umax) umax[blockIdx.x] = __uint_as_float(S.umax);
}

// CB trailing columns per wavefront pass: every LDS read of L21 (and of
// L11 in the U12 solve) serves CB columns' FMAs.
template <int M, int NB, typename T, int LT, int CB = 1>
__global__ __launch_bounds__(LT) void k_lu(T *__restrict__ blocks, int *__restrict__ rowsrc, int nblocks,
                                              int *flag, float *umax) {
    constexpr int NA = (M + 63) / 64;  // row chunks per lane
    __shared__ LuShared<M, NB, T, LT> S;
    T *A = blocks + static_cast<long long>(blockIdx.x) * M * M;
    const int lane = threadIdx.x % 64, wave = wave_index();
    const T *panel = S.panel;
    const lu_row *sigma = S.sigma;
    float um = 0.f;  // max |U12| of this thread's entries
    lu_begin(S);
    for (int k0 = 0; k0 < M; k0 += NB) {
        const int nb = min(NB, M - k0);
        lu_panel(A, S, k0, nb);
        // per trailing column j: gather rows through sigma, U12 segment by
        // forward substitution with the unit L11, A22 -= L21 U12, write back.
        for (int j0 = k0 + nb + CB * wave; j0 < M; j0 += CB * (LT / 64)) {
            T *colp[CB];
            bool live[CB];
            T x[CB][NA];
#pragma unroll
            for (int b = 0; b < CB; ++b) {
                live[b] = j0 + b < M;
                colp[b] = A + static_cast<long long>(live[b] ? j0 + b : j0) * M;
#pragma unroll
                for (int a = 0; a < NA; ++a) {
                    const int row = k0 + a * 64 + lane;
                    x[b][a] = row < M && live[b] ? colp[b][sigma[row]] : T(0);
                }
            }
            T u[CB][NB];
#pragma unroll
            for (int c = 0; c < NB; ++c) {
#pragma unroll
                for (int b = 0; b < CB; ++b) u[b][c] = T(0);
                if (c < nb) {
#pragma unroll
                    for (int b = 0; b < CB; ++b) u[b][c] = readlane_d(x[b][0], c);
                    if (lane > c && lane < nb) {
                        const T l = panel[c * M + k0 + lane];
#pragma unroll
                        for (int b = 0; b < CB; ++b) x[b][0] -= l * u[b][c];
                    }
                }
            }
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                const int row = k0 + a * 64 + lane;
                if (row >= k0 + nb && row < M) {
                    T acc[CB];
#pragma unroll
                    for (int b = 0; b < CB; ++b) acc[b] = x[b][a];
#pragma unroll
                    for (int c = 0; c < NB; ++c)
                        if (c < nb) {
                            const T l = panel[c * M + row];
#pragma unroll
                            for (int b = 0; b < CB; ++b) acc[b] -= l * u[b][c];
                        }
#pragma unroll
                    for (int b = 0; b < CB; ++b) x[b][a] = acc[b];
                }
            }
#pragma unroll
            for (int b = 0; b < CB; ++b) {
                if (!live[b]) continue;
                if (lane < nb && !isfinite(static_cast<double>(x[b][0]))) S.bad = 1;
                if (umax && lane < nb) um = fmaxf(um, static_cast<float>(fabs(static_cast<double>(x[b][0]))));  // U12
#pragma unroll
                for (int a = 0; a < NA; ++a) {
                    const int row = k0 + a * 64 + lane;
                    if (row < M) colp[b][row] = x[b][a];
                }
            }
        }
        __syncthreads();
    }
    if (um > 0.f) atomicMax(&S.umax, __float_as_uint(um));
    __syncthreads();
    lu_finish(A, S, rowsrc, flag, umax);
}

// MFMA 16x16x4 tiles (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32): lane l
// supplies A[l % 16][l / 16] and B[l / 16][l % 16]; the 16 x 16 accumulator
// holds column l % 16 in every lane, rows l / 16 + 4 r (FP64) or 4 (l / 16) + r
// (FP32) in register r.
typedef double mfma_d4 __attribute__((ext_vector_type(4)));
typedef float mfma_f4 __attribute__((ext_vector_type(4)));
template <typename T> struct Mfma16;
template <> struct Mfma16<double> {
    using acc = mfma_d4;
    static __device__ __forceinline__ acc mac(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return lane / 16 + 4 * r; }
};
template <> struct Mfma16<float> {
    using acc = mfma_f4;
    static __device__ __forceinline__ acc mac(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane / 16) + r; }
};

// k_lu with the trailing update on MFMA: each wavefront takes 16-column tiles
// of the trailing matrix; it solves U12 = L11^-1 A12 by forward substitution
// in the MFMA B layout (the value of row k is shuffled within each column's
// four lanes), writes U12, and updates the 16-row tiles of A22 with
// A22 -= L21 U12: NB / 4 MFMAs per tile, L21 read from the LDS panel. No write
// may overtake a read of a row the panel's interchanges moved:
//  * RESIDENT: the wavefront gathers its whole tile (U12 rows as B operands,
//    every A22 tile as an accumulator) through sigma into registers before any
//    write, which keeps (M - NB) / 4 accumulator values per lane;
//  * otherwise the source entries of the at most 2 nb moved rows are first
//    copied to LDS (slot[r], srcrow[q]), every other row is read in place, and
//    the A22 tiles are streamed one after the other (the next tile's loads
//    issued before the current tile's MFMAs).
template <int M, int NB, typename T, int LT, bool RESIDENT>
__global__ __launch_bounds__(LT) void k_lu_mfma(T *__restrict__ blocks, int *__restrict__ rowsrc, int nblocks,
                                                int *flag) {
    static_assert(NB % 4 == 0, "MFMA k-slices of 4 rows");
    using MF = Mfma16<T>;
    using Acc = typename MF::acc;
    constexpr int KS = NB / 4;                   // k-slices of the panel
    constexpr int TT = (M - NB + 15) / 16;       // A22 tiles below a full panel
    constexpr int WAVES = LT / 64;
    constexpr int SLOTS = 2 * NB;  // rows moved by one panel's interchanges
    __shared__ LuShared<M, NB, T, LT> S;
    __shared__ T saved[WAVES][RESIDENT ? 1 : SLOTS * 16];  // [wave][slot][column in tile]
    __shared__ int srcrow[SLOTS];
    __shared__ int slot[M];  // rows >= k0: slot of a moved row, -1 otherwise
    __shared__ int nslots;
    T *A = blocks + static_cast<long long>(blockIdx.x) * M * M;
    const int lane = threadIdx.x % 64, wave = wave_index();
    const int cl = lane % 16, kl = lane / 16;  // column in the tile; k (B) / row quarter (C)
    const T *panel = S.panel;
    const lu_row *sigma = S.sigma;
    lu_begin(S);
    for (int k0 = 0; k0 < M; k0 += NB) {
        const int nb = min(NB, M - k0);
        lu_panel(A, S, k0, nb);
        const int first = k0 + nb, tiles = (M - first + 15) / 16;
        if (tiles == 0) continue;
        if constexpr (!RESIDENT) {
            // slots of the moved rows in row order, by one wavefront with
            // ballots (hipcc 7.2 lowers the serial form of this loop, a uniform
            // select in a loop remainder, to s_cselect on a stale SCC)
            if (wave == 0) {
                int base = 0;
                for (int r0 = k0; r0 < M; r0 += 64) {
                    const int r = r0 + lane;
                    const int src = r < M ? sigma[r] : r;
                    const bool moved = src != r;
                    const unsigned long long mask = __ballot(moved);
                    const int at = base + __popcll(mask & ((1ull << lane) - 1ull));
                    if (r < M) slot[r] = moved ? at : -1;
                    if (moved) srcrow[at] = src;
                    base += __popcll(mask);
                }
                if (lane == 0) nslots = base;
            }
            __syncthreads();
        }
        const int ns = RESIDENT ? 0 : nslots;
        for (int ct = wave; ct < tiles; ct += WAVES) {
            const int col = first + 16 * ct + cl;
            const bool colok = col < M;
            T *colp = A + static_cast<long long>(colok ? col : 0) * M;
            if constexpr (!RESIDENT) {
                for (int q = kl; q < ns; q += 4) saved[wave][q * 16 + cl] = colok ? colp[srcrow[q]] : T(0);
                asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory");
            }
            // entry of logical row r after the panel's interchanges (streamed
            // form: both loads are issued; the moved rows take the saved value)
            auto value = [&](int r) -> T {
                if constexpr (RESIDENT) {
                    return colp[sigma[r]];
                } else {
                    const int q = slot[r];
                    const T kept = saved[wave][(q >= 0 ? q : 0) * 16 + cl];
                    const T here = colp[r];
                    return q >= 0 ? kept : here;
                }
            };
            T u[KS];
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const int k = 4 * s + kl;
                u[s] = (colok && k < nb) ? value(k0 + k) : T(0);
            }
            auto load_tile = [&](int r0, Acc &c) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = r0 + MF::row(lane, r);
                    c[r] = (colok && row < M) ? value(row) : T(0);
                }
            };
            auto update_tile = [&](int r0, Acc &c) {
                const int arow = r0 + cl;
#pragma unroll
                for (int s = 0; s < KS; ++s) {
                    const T a = (arow < M && 4 * s + kl < nb) ? -panel[(4 * s + kl) * M + arow] : T(0);
                    c = MF::mac(a, u[s], c);
                }
            };
            auto store_tile = [&](int r0, const Acc &c) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = r0 + MF::row(lane, r);
                    if (colok && row < M) colp[row] = c[r];
                }
            };
            Acc res[RESIDENT ? TT : 1];
            if constexpr (RESIDENT) {
#pragma unroll
                for (int t = 0; t < TT; ++t) load_tile(first + 16 * t, res[t]);
            }
            // U12: x_i -= L_ik x_k for k < i < nb, row k taken from lane cl + 16 (k % 4)
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                if (k < nb) {
                    const T xk = __shfl(u[k / 4], cl + 16 * (k % 4), 64);
#pragma unroll
                    for (int s = k / 4; s < KS; ++s) {
                        const int i = 4 * s + kl;
                        if (i > k && i < nb) u[s] -= panel[k * M + k0 + i] * xk;
                    }
                }
            }
            bool bad = false;
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const int k = 4 * s + kl;
                if (colok && k < nb) {
                    colp[k0 + k] = u[s];
                    bad = bad || !isfinite(static_cast<double>(u[s]));
                }
            }
            if (bad) S.bad = 1;
            if constexpr (RESIDENT) {
#pragma unroll
                for (int t = 0; t < TT; ++t)
                    if (first + 16 * t < M) {
                        update_tile(first + 16 * t, res[t]);
                        store_tile(first + 16 * t, res[t]);
                    }
            } else {
                Acc cur;
                load_tile(first, cur);
                for (int r0 = first; r0 < M; r0 += 16) {
                    Acc next;
                    if (r0 + 16 < M) load_tile(r0 + 16, next);
                    update_tile(r0, cur);
                    store_tile(r0, cur);
                    cur = next;
                }
                // the next tile of this wave overwrites `saved` only after these reads
                asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            }
        }
        __syncthreads();
    }
    lu_finish(A, S, rowsrc, flag);
}


// block_jacobi.hpp:24-39: z = U^-1 L^-1 P r per block; one wavefront per block,
// lane l owns rows l, l+64, ...; the value of each pivot row is broadcast with
// readlane and the factor columns stream once from HBM (column-major: one
// coalesced load per column and row chunk). A ring of U column buffers keeps
// the loads of the next U - 1 columns in flight while a column is applied;
// the ring slot of a column is fixed at compile time (the chunk loops are
// unrolled), so no registers move.
// columns in flight per wave in k_bj_apply, by stored factor type
#ifndef DGH_BJ_RING64
#define DGH_BJ_RING64 4
#endif
#ifndef DGH_BJ_RING32
#define DGH_BJ_RING32 4
#endif
constexpr int BJ_RING64 = DGH_BJ_RING64, BJ_RING32 = DGH_BJ_RING32;
// Blocks of at most this many 64-row chunks (128 rows in the 2-byte kernel)
// run branch-free: every lane loads every factor entry and the sweeps mask at
// the use instead of predicating the loads, and the ring is refilled with a
// clamped column index instead of a test for the end. Measured at 256^2, k=8:
// the 36-row HDG face blocks gain (condensed HDG update 1.33 -> 1.27 s); the
// 324-row element blocks lose 1-6% (the extra full-width loads and selects),
// so they keep predicated loads and the tested refill.
// Block-Jacobi applies: factor loads as raw buffer loads with scalar column
// offsets, bit-identical to the predicated loads. Measured at 256^2, k=8: the
// FP16 apply 2.68-2.71 -> 2.60 ms, then 2.72 -> 2.48 ms without the
// non-temporal bit (kept); FP64 factors 9.27-9.35 -> 8.94 ms with it (kept for
// the large FP64 blocks); FP32 unchanged either way (DGH_BJ_BUFFER=1 forces
// the buffer loads for every factor type).
#ifndef DGH_BJ_PAIRS_BUFFER
#define DGH_BJ_PAIRS_BUFFER 1
#endif
#ifndef DGH_BJ_BUFFER
#define DGH_BJ_BUFFER 0
#endif
#ifndef DGH_BJ_BRANCH_FREE_CHUNKS
#define DGH_BJ_BRANCH_FREE_CHUNKS 1
#endif
__device__ __forceinline__ float to_compute(float v) { return v; }
__device__ __forceinline__ double to_compute(double v) { return v; }

// Factor loads: each factor is read once per apply (55 GB in FP64 at 256^2,
// k=8) and never reused before it is evicted, so the loads are marked
// non-temporal (measured: FP64 apply 9.98 -> 9.18 ms, FP32 5.33 -> 4.87 ms).
#ifndef DGH_BJ_NONTEMPORAL
#define DGH_BJ_NONTEMPORAL 1
#endif
template <typename T> __device__ __forceinline__ T factor_load(const T *p) {
    if constexpr (!DGH_BJ_NONTEMPORAL) {
        return *p;
    } else {
        return __builtin_nontemporal_load(p);
    }
}

// Offsets of the packed LU layout of one M x M block: column c's strict lower
// part (rows c+1..M-1) starts at packed_lower(M, c); column c's upper part
// (rows 0..c) at packed_upper(M, c), after all lower parts, columns in
// descending order.
__device__ __forceinline__ long long packed_lower(int M, int c) {
    return static_cast<long long>(c) * (M - 1) - static_cast<long long>(c) * (c - 1) / 2;
}
__device__ __forceinline__ long long packed_upper(int M, int c) {
    return static_cast<long long>(M) * (M - 1) / 2 + static_cast<long long>(M) * (M + 1) / 2 -
           static_cast<long long>(c + 1) * (c + 2) / 2;
}
// Square column-major LU blocks (from `square`) into the packed layout (into
// `packed`, same block slots).
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_pack(const T *__restrict__ square, T *__restrict__ packed, int M) {
    const long long mm = static_cast<long long>(M) * M;
    const T *src = square + blockIdx.x * mm;
    T *dst = packed + blockIdx.x * mm;
    for (long long t = threadIdx.x; t < mm; t += BLOCK) {
        const int c = static_cast<int>(t / M), row = static_cast<int>(t % M);
        dst[row > c ? packed_lower(M, c) + (row - c - 1) : packed_upper(M, c) + row] = src[t];
    }
}

// TF: stored factor type (double or float; 2-byte factors use
// k_bj_apply_pairs); TV: vector and
// arithmetic type (double for the FP64 path, float for the inner FP32 solve).
template <int M, typename TF, typename TV, int U>
__global__ __launch_bounds__(BLOCK) void k_bj_apply(const TF *__restrict__ factors, const int *__restrict__ rowsrc,
                                                    const TV *__restrict__ r, TV *__restrict__ z, int nblocks,
                                                    int colour, int nx) {
    static_assert(64 % U == 0, "ring depth must divide the chunk width");
    using T = TF;
    constexpr int NA = (M + 63) / 64;
    const int lane = threadIdx.x % 64;
    // the wave index is uniform across the wave: say so, so that the block's
    // addresses stay in scalar registers
    const long long blk = static_cast<long long>(blockIdx.x) * (BLOCK / 64) +
                          __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / 64));
    if (blk >= nblocks) return;
    const T *F = factors + blk * M * M;
    // vector position of the block: its cell (compact colour storage) or itself
    const long long vb = colour < 0 ? blk : colour_cell(static_cast<int>(blk), colour, nx);
    TV x[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const int row = a * 64 + lane;
        x[a] = row < M ? r[vb * M + rowsrc[blk * M + row]] : TV(0);
    }
    T buf[U][NA];
    // Packed factors (k_pack): the strict lower parts of columns 0..M-2 back
    // to back, then the upper parts (diagonal included) of columns M-1..0, so
    // that both sweeps stream their columns contiguously.
    // Small blocks (BF): every lane loads, lanes outside the column's segment
    // read its first entry (inside the block: for c = M-1 the first upper
    // entry), and the sweeps mask those lanes where they use the value, so
    // that no load sits under a branch. Otherwise the loads are predicated and
    // the lanes outside the segment hold zeros.
    constexpr bool BF = NA <= DGH_BJ_BRANCH_FREE_CHUNKS;
    // Raw buffer loads (as in k_bj_apply_pairs): scalar column offsets, fixed
    // lane offsets, and an offset past the end for lanes outside the column's
    // segment, which load 0. The descriptor starts M entries before the block
    // so that the lower part's column offset stays non-negative. Used for FP64
    // factors of the large blocks (measured at 256^2, k = 8: 8.94 ms per apply
    // against 9.27-9.35 ms with non-temporal global loads, same box); FP32
    // factors ran the same either way and keep the predicated global loads, as
    // do the small blocks. DGH_BJ_BUFFER=1 selects the buffer loads for all.
    constexpr bool BUF = DGH_BJ_BUFFER || (sizeof(T) == 8 && !BF);
    constexpr unsigned OOB = 0x80000000u;
    constexpr int S = sizeof(T);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<T *>(F - M), 0, (M * M + M) * S, 0x00020000);
    unsigned voff[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) voff[a] = 64 * a + lane < M ? static_cast<unsigned>((64 * a + lane) * S) : OOB;
    auto raw_load = [&](unsigned vo, int soff) -> T {
        if constexpr (sizeof(T) == 8) {
            const auto v = __builtin_amdgcn_raw_buffer_load_b64(rs, vo, soff, 2);
            return __builtin_bit_cast(T, v);
        } else {
            return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rs, vo, soff, 2));
        }
    };
    // strictly lower part of column c (rows c+1..M-1), chunks a >= a0; buffer
    // loads: the segment starts in chunk a0 or a0 + 1 (c < 64 (a0 + 1) + U)
    auto load_lower = [&](T(&b)[NA], int c, int a0) {
        if constexpr (BUF) {
            const int soff = static_cast<int>(packed_lower(M, c) - c - 1 + M) * S;
#pragma unroll
            for (int a = 0; a < NA; ++a)
                if (a >= a0) b[a] = raw_load((a <= a0 + 1 && 64 * a + lane <= c) ? OOB : voff[a], soff);
        } else {
            const T *colp = F + packed_lower(M, c);
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                const int row = a * 64 + lane;
                const bool in = row > c && row < M;
                if (a >= a0) {
                    if constexpr (BF) b[a] = factor_load(colp + (in ? row - c - 1 : 0));
                    else b[a] = in ? factor_load(colp + (row - c - 1)) : T(0);
                }
            }
        }
    };
    // upper part (rows 0..c, diagonal included) of column c, chunks a <= a1;
    // buffer loads: c lies in chunk a1 or a1 - 1
    auto load_upper = [&](T(&b)[NA], int c, int a1) {
        if constexpr (BUF) {
            const int soff = static_cast<int>(packed_upper(M, c) + M) * S;
#pragma unroll
            for (int a = 0; a < NA; ++a)
                if (a <= a1) b[a] = raw_load((a >= a1 - 1 && 64 * a + lane > c) ? OOB : voff[a], soff);
        } else {
            const T *colp = F + packed_upper(M, c);
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                const int row = a * 64 + lane;
                if (a <= a1) {
                    if constexpr (BF) b[a] = factor_load(colp + (row <= c ? row : 0));
                    else b[a] = row <= c ? factor_load(colp + row) : T(0);
                }
            }
        }
    };
    // forward substitution with the unit lower L: x_i -= L_ij x_j for i > j;
    // column j sits in ring slot j % U
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (u < M) load_lower(buf[u], u, 0);
#pragma unroll
    for (int ja = 0; ja < NA; ++ja) {
        const int jend = min(64, M - ja * 64);
#pragma clang loop unroll(disable)
        for (int jl = 0; jl < jend; jl += U) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (jend % U == 0 || jl + u < jend) {
                    const int j = ja * 64 + jl + u;
                    const TV xj = readlane_d(x[ja], jl + u);
#pragma unroll
                    for (int a = ja; a < NA; ++a) {
                        const int row = a * 64 + lane;
                        const TV y = x[a] - static_cast<TV>(to_compute(buf[u][a])) * xj;
                        // BF: rows j+1..M-1 only, other lanes hold an unrelated entry
                        if (!BF || ((a > ja || row > j) && (a < NA - 1 || row < M))) x[a] = y;
                    }
                    if (BF) load_lower(buf[u], min(j + U, M - 1), ja);  // spare in-bounds load at the end
                    else if (j + U < M) load_lower(buf[u], j + U, ja);
                }
            }
        }
    }
    // backward substitution with U: x_j /= U_jj, x_i -= U_ij x_j for i < j;
    // column j sits in ring slot (M - 1 - j) % U
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (M - 1 - u >= 0) load_upper(buf[u], M - 1 - u, NA - 1);
#pragma unroll
    for (int ja = NA - 1; ja >= 0; --ja) {
        const int jend = min(64, M - ja * 64);
        const int t0 = M - ja * 64 - jend;  // ring position of the chunk's last column
#pragma clang loop unroll(disable)
        for (int g = 0; g < jend; g += U) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int jl = jend - 1 - g - u;
                if (jend % U == 0 || jl >= 0) {
                    const int j = ja * 64 + jl;
                    T(&cur)[NA] = buf[(t0 + u) % U];
                    const TV diag = readlane_d(static_cast<TV>(to_compute(cur[ja])), jl);
                    const TV xj = readlane_d(x[ja], jl) / diag;
                    if (lane == jl) x[ja] = xj;
#pragma unroll
                    for (int a = 0; a <= ja; ++a) {
                        const int row = a * 64 + lane;
                        if (row < j) x[a] -= static_cast<TV>(to_compute(cur[a])) * xj;
                    }
                    if (BF) load_upper(cur, max(j - U, 0), ja);  // spare in-bounds load at the end
                    else if (j - U >= 0) load_upper(cur, j - U, ja);
                }
            }
        }
    }
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const int row = a * 64 + lane;
        if (row < M) z[vb * M + row] = x[a];
    }
}

template <typename TS, typename TD>
__global__ void k_convert(const TS *__restrict__ src, TD *__restrict__ dst, long long count) {
    for (long long i = static_cast<long long>(blockIdx.x) * BLOCK + threadIdx.x; i < count;
         i += static_cast<long long>(gridDim.x) * BLOCK)
        dst[i] = static_cast<TD>(src[i]);
}

// ------------------------------------------------ two-byte factor layout
// Factors stored in 2 bytes (BF16, or FP16 with a power-of-two scale of U per
// block) are read as aligned 4-byte pairs of consecutive rows, so that a wave
// load moves 256 bytes instead of 128. Each column segment of the packed
// layout starts at an even row and has even length: the lower part of column c
// holds rows s(c) = 2 floor((c+1)/2) .. M-1 (row c itself as a zero when c is
// even), the upper part rows 0 .. c rounded up to even (a zero pad row). M is
// even (4 n^2). Block stride: pair_stride(M) entries.
__host__ __device__ __forceinline__ long long pair_lower(int M, int c) {
    return static_cast<long long>(c) * M - 2LL * ((static_cast<long long>(c) * c) / 4);
}
__host__ __device__ __forceinline__ long long pair_cumulative_upper(long long n) {  // columns 0..n
    return 2 * (n + 1) + 2 * ((n * n) / 4);
}
__host__ __device__ __forceinline__ long long pair_upper(int M, int c) {
    return pair_lower(M, M) + pair_cumulative_upper(M - 1) - pair_cumulative_upper(c);
}
__host__ __device__ __forceinline__ long long pair_stride(int M) {
    return ((pair_lower(M, M) + pair_cumulative_upper(M - 1)) + 63) / 64 * 64;
}
__device__ __forceinline__ unsigned short half_bits(float v, bool fp16) {
    if (fp16) return __half_as_ushort(__float2half_rn(v));
    const __hip_bfloat16 b = __float2bfloat16(v);
    unsigned short u;
    __builtin_memcpy(&u, &b, 2);
    return u;
}
__device__ __forceinline__ float half_value(unsigned short u, bool fp16) {
    if (fp16) return __half2float(__ushort_as_half(u));
    __hip_bfloat16 b;
    __builtin_memcpy(&b, &u, 2);
    return __bfloat162float(b);
}
// max |U| of each square FP32 LU block (U on and above the diagonal), for the
// factorisations that do not record it themselves.
__global__ __launch_bounds__(BLOCK) void k_umax(const float *__restrict__ square, float *__restrict__ umax, int M) {
    __shared__ float red[BLOCK / 64];
    const int lane = threadIdx.x % 64, wave = wave_index();
    const float *src = square + blockIdx.x * static_cast<long long>(M) * M;
    float mx = 0;
    for (int c = wave; c < M; c += BLOCK / 64)
        for (int r = lane; r <= c; r += 64) mx = fmaxf(mx, fabsf(src[static_cast<long long>(c) * M + r]));
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    if (lane == 0) red[wave] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = 0;
        for (int w = 0; w < BLOCK / 64; ++w) m = fmaxf(m, red[w]);
        umax[blockIdx.x] = m;
    }
}
// One block per factor: the square column-major FP32 LU (k_lu's output:
// unit L strictly below the diagonal, U on and above it) into the pair layout
// in 2 bytes. FP16 scales U by sigma = 2^-e, e = floor(log2 max|U|), so that
// the largest entry lies in [1, 2) (L is unscaled: |L_ij| <= 1 after partial
// pivoting); BF16 keeps sigma = 1. The pivots are not taken from the 2-byte
// copy: rdiag[j] = 1 / (sigma U_jj) in FP32, so that a small pivot keeps its
// FP32 value instead of becoming a subnormal or zero 2-byte number; a
// reciprocal that is not finite and nonzero raises bit 16 of *flag.
// max|U| comes from the factorisation (umax[block], recorded by k_lu or
// k_umax), so every source entry is read once. Each wave converts whole
// columns: every pair (4 bytes) of the layout is written once, pads included.
__global__ __launch_bounds__(BLOCK) void k_to_pairs(const float *__restrict__ square, unsigned short *__restrict__ out,
                                                    float *__restrict__ sigma, float *__restrict__ rdiag,
                                                    const float *__restrict__ umax, int M, int fp16, int *flag) {
    const int lane = threadIdx.x % 64, wave = wave_index();
    const long long stride = pair_stride(M);
    const float *src = square + blockIdx.x * static_cast<long long>(M) * M;
    unsigned short *dst = out + blockIdx.x * stride;
    unsigned *dst32 = reinterpret_cast<unsigned *>(dst);
    const float m = fp16 ? umax[blockIdx.x] : 0.f;
    const float sc = (fp16 && m > 0) ? ldexpf(1.0f, -ilogbf(m)) : 1.0f;
    if (threadIdx.x == 0) sigma[blockIdx.x] = sc;
    bool bad = false;
    for (int j = threadIdx.x; j < M; j += BLOCK) {
        const float rd = 1.0f / (sc * src[static_cast<long long>(j) * M + j]);
        rdiag[blockIdx.x * static_cast<long long>(M) + j] = rd;
        bad = bad || !(isfinite(rd) && rd != 0.f);
    }
    if (bad) atomicOr(flag, 16);
    for (int c = wave; c < M; c += BLOCK / 64) {
        const float *col = src + static_cast<long long>(c) * M;
        // lower part: rows s0 .. M-1 in pairs; row c itself (even c) is a zero
        const int s0 = 2 * ((c + 1) / 2);
        unsigned *lo = dst32 + pair_lower(M, c) / 2;
        for (int p = lane; p < (M - s0) / 2; p += 64) {
            const int row = s0 + 2 * p;
            const unsigned a = row > c ? half_bits(col[row], fp16) : 0u;
            const unsigned b = half_bits(col[row + 1], fp16);
            lo[p] = a | (b << 16);
        }
        // upper part: rows 0 .. c scaled by sigma, padded with a zero to even length
        unsigned *up = dst32 + pair_upper(M, c) / 2;
        for (int p = lane; p < (c + 2) / 2; p += 64) {
            const int row = 2 * p;
            const unsigned a = half_bits(col[row] * sc, fp16);
            const unsigned b = row + 1 <= c ? half_bits(col[row + 1] * sc, fp16) : 0u;
            up[p] = a | (b << 16);
        }
    }
    for (long long t = pair_lower(M, M) + pair_cumulative_upper(M - 1) + threadIdx.x; t < stride; t += BLOCK)
        dst[t] = 0;
}

// Block Jacobi apply on the pair layout: lane l owns rows 128 a + 2 l and
// 128 a + 2 l + 1 of chunk a and reads both with one 4-byte load per column;
// otherwise as k_bj_apply (ring of U columns in flight, pivot values by
// readlane). With FP16 factors x is multiplied by sigma between the two
// sweeps, since U' = sigma U gives U^-1 b = U'^-1 (sigma b).
template <int M, bool FP16, int U>
__global__ __launch_bounds__(BLOCK) void k_bj_apply_pairs(const unsigned short *__restrict__ factors,
                                                          const float *__restrict__ sigma,
                                                          const float *__restrict__ rdiag,
                                                          const int *__restrict__ rowsrc, const float *__restrict__ r,
                                                          float *__restrict__ z, int nblocks) {
    static_assert(M % 2 == 0 && 128 % U == 0 && U % 2 == 0, "pair layout");
    constexpr int NA = (M + 127) / 128;
    const int lane = threadIdx.x % 64;
    // the wave index is uniform across the wave: say so, so that the block's
    // addresses stay in scalar registers
    const long long blk = static_cast<long long>(blockIdx.x) * (BLOCK / 64) +
                          __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / 64));
    if (blk >= nblocks) return;
    const unsigned *F = reinterpret_cast<const unsigned *>(factors + blk * pair_stride(M));
    float x[NA][2];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int row = 128 * a + 2 * lane + e;
            x[a][e] = row < M ? r[blk * M + rowsrc[blk * M + row]] : 0.f;
        }
    unsigned buf[U][NA];
    auto lo16 = [](unsigned w) { return static_cast<unsigned short>(w & 0xffffu); };
    auto hi16 = [](unsigned w) { return static_cast<unsigned short>(w >> 16); };
    // Small blocks (BF) as in k_bj_apply: every lane loads (lanes outside the
    // segment read its first pair, for c = M-1 the first upper pair) and the
    // sweeps mask those lanes at the use; otherwise predicated loads and zeros.
    constexpr bool BF = NA <= DGH_BJ_BRANCH_FREE_CHUNKS;
#if DGH_BJ_PAIRS_BUFFER
    // Raw buffer loads: the column's start is a scalar offset and each lane's
    // offset within the column is fixed per chunk, so no per-load address
    // arithmetic runs on the vector ALUs. Lanes outside the column's segment
    // get an offset past the block's end, for which the load returns 0. The
    // loads use the default cache policy: with the non-temporal bit the apply
    // took 2.717 ms against 2.481 ms (same box, 256^2, k = 8).
    constexpr int PAIR_LOAD_POLICY = 0;
    constexpr unsigned OOB = 0x80000000u;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<unsigned *>(F), 0, static_cast<int>(pair_stride(M) * 2), 0x00020000);
    unsigned voff[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) voff[a] = 128 * a + 2 * lane < M ? 256u * a + 4u * lane : OOB;
    // lower part of column c (rows s(c)..M-1), chunks a >= a0; for c < M the
    // segment starts in chunk a0 or a0 + 1, so later chunks need no test
    auto load_lower = [&](unsigned(&b)[NA], int c, int a0) {
        const int s0 = 2 * ((c + 1) / 2);
        const int soff = 2 * static_cast<int>(pair_lower(M, c) - s0);
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            if (a >= a0) {
                const unsigned vo = (a <= a0 + 1 && 128 * a + 2 * lane < s0) ? OOB : voff[a];
                b[a] = __builtin_amdgcn_raw_buffer_load_b32(rs, vo, soff, PAIR_LOAD_POLICY);
            }
        }
    };
    // upper part of column c (rows 0..c and the pad), chunks a <= a1; c lies
    // in chunk a1 or a1 - 1, so earlier chunks need no test
    auto load_upper = [&](unsigned(&b)[NA], int c, int a1) {
        const int soff = 2 * static_cast<int>(pair_upper(M, c));
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            if (a <= a1) {
                const unsigned vo = (a >= a1 - 1 && 128 * a + 2 * lane > c) ? OOB : voff[a];
                b[a] = __builtin_amdgcn_raw_buffer_load_b32(rs, vo, soff, PAIR_LOAD_POLICY);
            }
        }
    };
#else
    // lower part of column c from the pair holding row s(c), chunks a >= a0
    auto load_lower = [&](unsigned(&b)[NA], int c, int a0) {
        const int s0 = 2 * ((c + 1) / 2);
        const unsigned *colp = F + (pair_lower(M, c) - s0) / 2;
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            const int r0 = 128 * a + 2 * lane;
            const bool in = r0 >= s0 && r0 < M;
            if (a >= a0) {
                if constexpr (BF) b[a] = __builtin_nontemporal_load(colp + (in ? r0 : s0) / 2);
                else b[a] = in ? __builtin_nontemporal_load(colp + r0 / 2) : 0u;
            }
        }
    };
    // upper part of column c (rows 0..c and the pad), chunks a <= a1
    auto load_upper = [&](unsigned(&b)[NA], int c, int a1) {
        const unsigned *colp = F + pair_upper(M, c) / 2;
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            const int r0 = 128 * a + 2 * lane;
            if (a <= a1) {
                if constexpr (BF) b[a] = __builtin_nontemporal_load(colp + (r0 <= c ? r0 : 0) / 2);
                else b[a] = r0 <= c ? __builtin_nontemporal_load(colp + r0 / 2) : 0u;
            }
        }
    };
#endif
#pragma unroll
    for (int u = 0; u < U; ++u) load_lower(buf[u], u, 0);
#pragma unroll
    for (int ja = 0; ja < NA; ++ja) {
        const int jend = min(128, M - ja * 128);
#pragma clang loop unroll(disable)
        for (int jl = 0; jl < jend; jl += U) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (jend % U == 0 || jl + u < jend) {
                    const int j = ja * 128 + jl + u;
                    const float xj = readlane_d(x[ja][u & 1], (jl + u) >> 1);
#pragma unroll
                    for (int a = ja; a < NA; ++a) {
                        // BF: rows j+1..M-1 only, other lanes hold an unrelated pair
                        const int r0 = 128 * a + 2 * lane;
                        const float y0 = x[a][0] - half_value(lo16(buf[u][a]), FP16) * xj;
                        const float y1 = x[a][1] - half_value(hi16(buf[u][a]), FP16) * xj;
                        if (!BF || ((a > ja || r0 > j) && (a < NA - 1 || r0 < M))) x[a][0] = y0;
                        if (!BF || ((a > ja || r0 + 1 > j) && (a < NA - 1 || r0 + 1 < M))) x[a][1] = y1;
                    }
                    if (BF) load_lower(buf[u], min(j + U, M - 1), ja);  // spare in-bounds load at the end
                    else if (j + U < M) load_lower(buf[u], j + U, ja);
                }
            }
        }
    }
    if (FP16) {
        const float sc = sigma[blk];
#pragma unroll
        for (int a = 0; a < NA; ++a) x[a][0] *= sc, x[a][1] *= sc;
    }
    // FP32 reciprocal pivots of the lane's rows (1 / (sigma U_ii))
    float rd[NA][2];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int row = 128 * a + 2 * lane + e;
            rd[a][e] = row < M ? rdiag[blk * M + row] : 0.f;
        }
#pragma unroll
    for (int u = 0; u < U; ++u) load_upper(buf[u], M - 1 - u, NA - 1);
#pragma unroll
    for (int ja = NA - 1; ja >= 0; --ja) {
        const int jend = min(128, M - ja * 128);
        const int t0 = M - ja * 128 - jend;
#pragma clang loop unroll(disable)
        for (int g = 0; g < jend; g += U) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int jl = jend - 1 - g - u;
                if (jend % U == 0 || jl >= 0) {
                    const int j = ja * 128 + jl;
                    unsigned(&cur)[NA] = buf[(t0 + u) % U];
                    // jend is even, so the parity of jl is that of u + 1
                    const int e = (u & 1) ? 0 : 1;
                    const float xj = readlane_d(x[ja][e] * rd[ja][e], jl >> 1);
                    if (lane == (jl >> 1)) x[ja][e] = xj;
#pragma unroll
                    for (int a = 0; a <= ja; ++a) {
                        const int r0 = 128 * a + 2 * lane;
                        if (r0 < j) x[a][0] -= half_value(lo16(cur[a]), FP16) * xj;
                        if (r0 + 1 < j) x[a][1] -= half_value(hi16(cur[a]), FP16) * xj;
                    }
                    if (BF) load_upper(cur, max(j - U, 0), ja);  // spare in-bounds load at the end
                    else if (j - U >= 0) load_upper(cur, j - U, ja);
                }
            }
        }
    }
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int row = 128 * a + 2 * lane + e;
            if (row < M) z[blk * M + row] = x[a][e];
        }
}

// ---------------------------------------------- 2-byte factors, column pairs
// The same 2-byte factors with the entries of two columns side by side: the
// lower parts of columns (c, c+1), c even, as rows c+1..M-1 of 4-byte words
// (L_rc | L_r,c+1 << 16; row c+1 holds 0 for the unit diagonal), pair g = c/2
// at word g (M - g); the upper parts of columns (c, c-1), c = M-1, M-3, ..., as
// rows 0..c-1 of words (sigma U_rc | sigma U_r,c-1 << 16; row c-1 holds 0 for
// the diagonal, whose FP32 reciprocal is in rdiag), pair q at word
// M^2/4 + q (M - q). One column-pair step of the apply reads up to 4 M bytes
// in 64-row chunks, as the 4-byte kernels' column steps do, instead of the
// 2 M bytes of the row-pair layout (a dependency-free stream of the two
// layouts at 256^2, k = 8: 5.95 against 5.68 TB/s).
__host__ __device__ __forceinline__ long long colpair_stride(int M) {  // entries (2 bytes) per block
    return 2 * ((static_cast<long long>(M) * M / 2 + 63) / 64 * 64);
}
// One block per factor, as k_to_pairs (scale, reciprocal pivots, flag), into
// the column-pair layout; every word of the block is written once.
__global__ __launch_bounds__(BLOCK) void k_to_colpairs(const float *__restrict__ square,
                                                       unsigned short *__restrict__ out, float *__restrict__ sigma,
                                                       float *__restrict__ rdiag, const float *__restrict__ umax,
                                                       int M, int fp16, int *flag) {
    const int lane = threadIdx.x % 64, wave = wave_index();
    const float *src = square + blockIdx.x * static_cast<long long>(M) * M;
    unsigned *dst = reinterpret_cast<unsigned *>(out + blockIdx.x * colpair_stride(M));
    const float m = fp16 ? umax[blockIdx.x] : 0.f;
    const float sc = (fp16 && m > 0) ? ldexpf(1.0f, -ilogbf(m)) : 1.0f;
    if (threadIdx.x == 0) sigma[blockIdx.x] = sc;
    bool bad = false;
    for (int j = threadIdx.x; j < M; j += BLOCK) {
        const float rd = 1.0f / (sc * src[static_cast<long long>(j) * M + j]);
        rdiag[blockIdx.x * static_cast<long long>(M) + j] = rd;
        bad = bad || !(isfinite(rd) && rd != 0.f);
    }
    if (bad) atomicOr(flag, 16);
    const int half = M / 2;
    for (int g = wave; g < half; g += BLOCK / 64) {
        const int c = 2 * g, s0 = c + 1;
        const float *c0 = src + static_cast<long long>(c) * M, *c1 = c0 + M;
        unsigned *w = dst + static_cast<long long>(g) * (M - g);
        for (int r = s0 + lane; r < M; r += 64) {
            const unsigned a = half_bits(c0[r], fp16);
            const unsigned b = r > c + 1 ? half_bits(c1[r], fp16) : 0u;
            w[r - s0] = a | (b << 16);
        }
    }
    unsigned *up = dst + static_cast<long long>(M) * M / 4;
    for (int q = wave; q < half; q += BLOCK / 64) {
        const int c = M - 1 - 2 * q;
        const float *c0 = src + static_cast<long long>(c) * M, *c1 = c0 - M;
        unsigned *w = up + static_cast<long long>(q) * (M - q);
        for (int r = lane; r < c; r += 64) {
            const unsigned a = half_bits(c0[r] * sc, fp16);
            const unsigned b = r < c - 1 ? half_bits(c1[r] * sc, fp16) : 0u;
            w[r] = a | (b << 16);
        }
    }
    for (long long t = static_cast<long long>(M) * M / 2 + threadIdx.x; t < colpair_stride(M) / 2; t += BLOCK)
        dst[t] = 0;
}

// Block Jacobi apply on the column-pair layout: lane l owns rows 64 a + l; per
// column pair one 4-byte load per chunk brings both columns' entries of a row.
// The forward sweep takes x_c (readlane), applies column c to every row below,
// then takes the updated x_{c+1} and applies column c+1; the backward sweep
// likewise with the reciprocal pivots. Every row sees the same operations in
// the same order as in k_bj_apply_pairs. A ring of U column pairs is in flight
// per wave; a pair's ring slot follows its index, which is a compile-time
// constant because every 64-row chunk but the last holds 32 pairs (32 % U == 0).
template <int M, bool FP16, int U>
__global__ __launch_bounds__(BLOCK) void k_bj_apply_colpairs(const unsigned short *__restrict__ factors,
                                                             const float *__restrict__ sigma,
                                                             const float *__restrict__ rdiag,
                                                             const int *__restrict__ rowsrc,
                                                             const float *__restrict__ r, float *__restrict__ z,
                                                             int nblocks) {
    static_assert(M % 2 == 0 && 32 % U == 0, "column-pair layout");
    constexpr int NA = (M + 63) / 64, HALF = M / 2;
    constexpr int LAST = (M - 64 * (NA - 1)) / 2;  // pairs in the last chunk
    constexpr unsigned OOB = 0x80000000u;
    const int lane = threadIdx.x % 64;
    const long long blk = static_cast<long long>(blockIdx.x) * (BLOCK / 64) +
                          __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / 64));
    if (blk >= nblocks) return;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<unsigned short *>(factors + blk * colpair_stride(M)), 0, static_cast<int>(colpair_stride(M) * 2),
        0x00020000);
    float x[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const int row = 64 * a + lane;
        x[a] = row < M ? r[blk * M + rowsrc[blk * M + row]] : 0.f;
    }
    auto lo = [](unsigned w) { return half_value(static_cast<unsigned short>(w & 0xffffu), FP16); };
    auto hi = [](unsigned w) { return half_value(static_cast<unsigned short>(w >> 16), FP16); };
    unsigned buf[U][NA];
    // lower part of pair g (rows 2g+1..M-1), chunks a >= a0
    auto load_lower = [&](unsigned(&b)[NA], int g, int a0) {
        const int s0 = 2 * g + 1, soff = 4 * g * (M - g);
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            if (a >= a0) {
                const int row = 64 * a + lane;
                const unsigned vo = row >= s0 && row < M ? 4u * static_cast<unsigned>(row - s0) : OOB;
                b[a] = __builtin_amdgcn_raw_buffer_load_b32(rs, vo, soff, 0);
            }
        }
    };
    // upper part of pair q (columns c = M-1-2q and c-1, rows 0..c-1), chunks a <= a1
    auto load_upper = [&](unsigned(&b)[NA], int q, int a1) {
        const int c = M - 1 - 2 * q, soff = 4 * (M * M / 4 + q * (M - q));
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            if (a <= a1) {
                const int row = 64 * a + lane;
                const unsigned vo = row < c ? 4u * static_cast<unsigned>(row) : OOB;
                b[a] = __builtin_amdgcn_raw_buffer_load_b32(rs, vo, soff, 0);
            }
        }
    };
    // forward sweep with the unit lower factor
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (u < HALF) load_lower(buf[u], u, 0);
#pragma unroll
    for (int ja = 0; ja < NA; ++ja) {
        const int npairs = ja < NA - 1 ? 32 : LAST;
#pragma clang loop unroll(disable)
        for (int p0 = 0; p0 < npairs; p0 += U) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int p = p0 + u;
                if (npairs % U == 0 || p < npairs) {
                    const int g = 32 * ja + p;
                    unsigned(&cur)[NA] = buf[u];
                    const float xc = readlane_d(x[ja], 2 * p);
#pragma unroll
                    for (int a = ja; a < NA; ++a)
                        if (a > ja || lane > 2 * p) x[a] -= lo(cur[a]) * xc;
                    const float xc1 = readlane_d(x[ja], 2 * p + 1);
#pragma unroll
                    for (int a = ja; a < NA; ++a)
                        if (a > ja || lane > 2 * p + 1) x[a] -= hi(cur[a]) * xc1;
                    if (g + U < HALF) load_lower(cur, g + U, ja);
                }
            }
        }
    }
    if (FP16) {
        const float sc = sigma[blk];
#pragma unroll
        for (int a = 0; a < NA; ++a) x[a] *= sc;
    }
    float rd[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const int row = 64 * a + lane;
        rd[a] = row < M ? rdiag[blk * M + row] : 0.f;
    }
    // backward sweep: pair q's slot is (q + PAD) % U, so that the pairs of
    // every full chunk start at slot 0
    constexpr int PAD = (U - LAST % U) % U;
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (u < HALF) load_upper(buf[(u + PAD) % U], u, NA - 1);
#pragma unroll
    for (int ja = NA - 1; ja >= 0; --ja) {
        const int npairs = ja == NA - 1 ? LAST : 32;
        const int qbase = ja == NA - 1 ? 0 : LAST + 32 * (NA - 2 - ja);
        const int pad = ja == NA - 1 ? PAD : 0;
#pragma clang loop unroll(disable)
        for (int p0 = 0; p0 < npairs; p0 += U) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int p = p0 + u;
                if (npairs % U == 0 || p < npairs) {
                    const int q = qbase + p;
                    unsigned(&cur)[NA] = buf[(u + pad) % U];
                    const int lc = 2 * (npairs - p) - 1;  // lane of row c in chunk ja
                    const float xj = readlane_d(x[ja] * rd[ja], lc);
                    if (lane == lc) x[ja] = xj;
#pragma unroll
                    for (int a = 0; a <= ja; ++a)
                        if (a < ja || lane < lc) x[a] -= lo(cur[a]) * xj;
                    const float xj1 = readlane_d(x[ja] * rd[ja], lc - 1);
                    if (lane == lc - 1) x[ja] = xj1;
#pragma unroll
                    for (int a = 0; a <= ja; ++a)
                        if (a < ja || lane < lc - 1) x[a] -= hi(cur[a]) * xj1;
                    if (q + U < HALF) load_upper(cur, q + U, ja);
                }
            }
        }
    }
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const int row = 64 * a + lane;
        if (row < M) z[blk * M + row] = x[a];
    }
}

// ----------------------------------------------------------------- GMRES
constexpr int MAXV = 32;  // restart + 1 <= 32 basis vectors per launch

template <typename T> struct PtrsT { const T *p[MAXV]; };
using Ptrs = PtrsT<double>;

// ---- vectorised Arnoldi passes with a compile-time basis count NV
// Each thread takes W consecutive entries per step with one 16-byte load per
// vector (W = 2 doubles or 4 floats); the basis values stay in registers
// between the update of w and the next projection; the DG metric's inverse
// mass comes from an LDS copy indexed by the node (entry mod P).
template <typename TV> struct Pack;
template <> struct Pack<double> {
    using T = double2;
    static constexpr int W = 2;
    __device__ static void get(const T &a, double *o) { o[0] = a.x; o[1] = a.y; }
    __device__ static double at(const T &a, int e) { return e == 0 ? a.x : a.y; }
    __device__ static T put(const double *o) { return make_double2(o[0], o[1]); }
};
template <> struct Pack<float> {
    using T = float4;
    static constexpr int W = 4;
    __device__ static void get(const T &a, double *o) { o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w; }
    __device__ static double at(const T &a, int e) { return e == 0 ? a.x : e == 1 ? a.y : e == 2 ? a.z : a.w; }
    __device__ static T put(const double *o) {
        return make_float4(static_cast<float>(o[0]), static_cast<float>(o[1]), static_cast<float>(o[2]),
                           static_cast<float>(o[3]));
    }
};

// One packed load of a Krylov basis vector in the Gram-Schmidt passes, with the
// non-temporal bit: each pass streams up to 31 basis vectors that the next
// step reads again only after the preconditioner and the operator have moved
// many times the L2's size. Measured at 256^2, k=8 with FP16 factors: one
// inner Arnoldi step's passes 0.570 -> 0.524 ms, the update 0.924 -> 0.937
// updates/s, the same iterates (DGH_GS_NT=0 gives the plain loads). Twice as
// many reduction workgroups (2048) changed nothing.
#ifndef DGH_GS_NT
#define DGH_GS_NT 1
#endif
template <typename T> __device__ __forceinline__ T load_basis(const T *p) {
#if DGH_GS_NT
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    const u4 v = __builtin_nontemporal_load(reinterpret_cast<const u4 *>(p));
    T out;
    __builtin_memcpy(&out, &v, sizeof(T));
    return out;
#else
    return *p;
#endif
}
template <typename TV>
__device__ __forceinline__ void metric(const double *__restrict__ weights, const double *im, int P, int i, int w,
                                       double *out) {
    if (weights) {
        for (int e = 0; e < w; ++e) out[e] = weights[i + e];
    } else {
        int r = i % P;
        for (int e = 0; e < w; ++e) {
            out[e] = im[r];
            r = r + 1 == P ? 0 : r + 1;
        }
    }
}

template <int NV>
__device__ __forceinline__ void store_partials(double (&acc)[NV], int width, double *partials) {
    __shared__ double smem[BLOCK / 64][NV];
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
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

// partials[b][k] = sum_i V_k[i] * (W_i * w_i), W = inverse mass (weights ==
// nullptr) or weights[i]: the projections of gmres.cpp:39-43
template <typename TV, int NV>
__global__ __launch_bounds__(BLOCK) void k_mdot(PtrsT<TV> V, const TV *__restrict__ w,
                                                const double *__restrict__ weights, int count, int P,
                                                double *partials, const Basis *__restrict__ B) {
    using PK = Pack<TV>;
    constexpr int W = PK::W;
    __shared__ double im[MAXN * MAXN];
    if (!weights)
        for (int t = threadIdx.x; t < P; t += BLOCK) im[t] = B->inv_mass[t];
    __syncthreads();
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0;
    const int groups = count / W;
    for (int gi = blockIdx.x * BLOCK + threadIdx.x; gi < groups; gi += gridDim.x * BLOCK) {
        const int i = gi * W;
        double wv[W], mw[W];
        PK::get(*reinterpret_cast<const typename PK::T *>(w + i), wv);
        metric<TV>(weights, im, P, i, W, mw);
#pragma unroll
        for (int e = 0; e < W; ++e) wv[e] *= mw[e];
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double vk[W];
            PK::get(load_basis(reinterpret_cast<const typename PK::T *>(V.p[k] + i)), vk);
#pragma unroll
            for (int e = 0; e < W; ++e) acc[k] += vk[e] * wv[e];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < count - groups * W) {  // tail entries
        const int i = groups * W + threadIdx.x;
        double mw;
        metric<TV>(weights, im, P, i, 1, &mw);
        const double ww = mw * static_cast<double>(w[i]);
#pragma unroll
        for (int k = 0; k < NV; ++k) acc[k] += static_cast<double>(V.p[k][i]) * ww;
    }
    store_partials<NV>(acc, NV, partials);
}

// w -= sum_k V_k h_k (gmres.cpp:44-45); then mode 0: projections of the
// updated w on V (the second CGS pass), mode 1: its weighted square norm,
// mode 2: nothing
template <typename TV, int NV>
__global__ __launch_bounds__(BLOCK) void k_mupdate(PtrsT<TV> V, const double *__restrict__ h, TV *__restrict__ w,
                                                   const double *__restrict__ weights, int count, int P,
                                                   double *partials, int mode, const Basis *__restrict__ B) {
    using PK = Pack<TV>;
    constexpr int W = PK::W;
    __shared__ double im[MAXN * MAXN];
    if (!weights)
        for (int t = threadIdx.x; t < P; t += BLOCK) im[t] = B->inv_mass[t];
    __syncthreads();
    double hk[NV], acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        hk[k] = h[k];
        acc[k] = 0;
    }
    const int groups = count / W;
    for (int gi = blockIdx.x * BLOCK + threadIdx.x; gi < groups; gi += gridDim.x * BLOCK) {
        const int i = gi * W;
        typename PK::T vk[NV];  // kept in the stored type: 16 bytes per vector
        double wv[W];
#pragma unroll
        for (int k = 0; k < NV; ++k) vk[k] = load_basis(reinterpret_cast<const typename PK::T *>(V.p[k] + i));
        PK::get(*reinterpret_cast<const typename PK::T *>(w + i), wv);
#pragma unroll
        for (int e = 0; e < W; ++e) {
            double s = 0;
#pragma unroll
            for (int k = 0; k < NV; ++k) s += PK::at(vk[k], e) * hk[k];
            wv[e] = static_cast<double>(static_cast<TV>(wv[e] - s));
        }
        *reinterpret_cast<typename PK::T *>(w + i) = PK::put(wv);
        if (mode != 2) {
            double mw[W];
            metric<TV>(weights, im, P, i, W, mw);
            if (mode == 0) {
#pragma unroll
                for (int k = 0; k < NV; ++k)
#pragma unroll
                    for (int e = 0; e < W; ++e) acc[k] += PK::at(vk[k], e) * (mw[e] * wv[e]);
            } else {
#pragma unroll
                for (int e = 0; e < W; ++e) acc[0] += wv[e] * (mw[e] * wv[e]);
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < count - groups * W) {  // tail entries
        const int i = groups * W + threadIdx.x;
        double s = 0, vk[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            vk[k] = static_cast<double>(V.p[k][i]);
            s += vk[k] * hk[k];
        }
        const TV wt = static_cast<TV>(static_cast<double>(w[i]) - s);
        w[i] = wt;
        const double wi = static_cast<double>(wt);
        if (mode != 2) {
            double mw;
            metric<TV>(weights, im, P, i, 1, &mw);
            if (mode == 0) {
#pragma unroll
                for (int k = 0; k < NV; ++k) acc[k] += vk[k] * (mw * wi);
            } else {
                acc[0] += wi * (mw * wi);
            }
        }
    }
    if (mode == 2) return;
    store_partials<NV>(acc, mode == 0 ? NV : 1, partials);
}

// One Arnoldi step's scalar work on the device (gmres.cpp:44-66), for the
// graph-captured GMRES: H column j from the two CGS passes, the new norm, the
// earlier Givens rotations, the new rotation and the rotated right-hand side.
// H is (R+1) x R column-major; s holds the pass-1 / pass-2 projections at
// [0..j] / [32..32+j] and the squared norm at [64]. out (host-visible)
// receives |g_{j+1}|, the new norm and the rotation's d.
__global__ void k_arnoldi_givens(int j, int R, const double *__restrict__ s, double *H, double *cs, double *sn,
                                 double *g, double *next_out, double *out, int passes) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    auto Hm = [&](int i, int k) -> double & { return H[i + (R + 1) * k]; };
    for (int k = 0; k <= j; ++k) Hm(k, j) = (0.0 + s[k]) + (passes == 2 ? s[32 + k] : 0.0);
    const double next = sqrt(fmax(0.0, s[64]));
    Hm(j + 1, j) = next;
    *next_out = next;
    for (int i = 0; i < j; ++i) {
        const double a = Hm(i, j), b = Hm(i + 1, j);
        Hm(i, j) = cs[i] * a + sn[i] * b;
        Hm(i + 1, j) = -sn[i] * a + cs[i] * b;
    }
    const double d = hypot(Hm(j, j), Hm(j + 1, j));
    if (isfinite(d) && d > 0) {
        cs[j] = Hm(j, j) / d;
        sn[j] = Hm(j + 1, j) / d;
        Hm(j, j) = d;
        Hm(j + 1, j) = 0;
        g[j + 1] = -sn[j] * g[j];
        g[j] *= cs[j];
    }
    out[0] = fabs(g[j + 1]);
    out[1] = next;
    out[2] = d;
}

// out = x / *beta when *beta > 0 (basis_.col(j + 1) = column_ / next)
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_scale_div_dev(const T *__restrict__ x, const double *__restrict__ beta,
                                                         T *__restrict__ out, long long count) {
    const double b = *beta;
    if (!(b > 0)) return;
    for (long long i = static_cast<long long>(blockIdx.x) * BLOCK + threadIdx.x; i < count;
         i += static_cast<long long>(gridDim.x) * BLOCK)
        out[i] = static_cast<T>(static_cast<double>(x[i]) / b);
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

// partials[b] = sum a_i b_i
__global__ __launch_bounds__(BLOCK) void k_dot(const double *__restrict__ a, const double *__restrict__ b,
                                               long long count, double *partials) {
    __shared__ double smem[BLOCK / 64];
    double acc = 0;
    for (long long i = static_cast<long long>(blockIdx.x) * BLOCK + threadIdx.x; i < count;
         i += static_cast<long long>(gridDim.x) * BLOCK)
        acc += a[i] * b[i];
    const double t = block_sum(acc, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}
// out = (M lambda - k) / tau with the DG mass M(node)
template <int N>
__global__ __launch_bounds__(BLOCK) void k_mass_combine(const double *__restrict__ lambda,
                                                        const double *__restrict__ k, double tau,
                                                        double *__restrict__ out, long long count,
                                                        const Basis *__restrict__ B) {
    constexpr int P = N * N;
    for (long long i = static_cast<long long>(blockIdx.x) * BLOCK + threadIdx.x; i < count;
         i += static_cast<long long>(gridDim.x) * BLOCK)
        out[i] = (B->mass[i % P] * lambda[i] - k[i]) / tau;
}

// g = d/du of the density integral over the cells whose centre lies in x > 0,
// y > 0 (the exact quadrant for even nx, ny): the node mass on density nodes
template <int N>
__global__ __launch_bounds__(BLOCK) void k_quadrant_weights(double *__restrict__ g, const Basis *__restrict__ B,
                                                            Geo geo) {
    constexpr int P = N * N, M = 4 * P;
    const long long count = static_cast<long long>(geo.cells) * M;
    for (long long i = static_cast<long long>(blockIdx.x) * BLOCK + threadIdx.x; i < count;
         i += static_cast<long long>(gridDim.x) * BLOCK) {
        const int cell = static_cast<int>(i / M), local = static_cast<int>(i % M);
        const int cx = cell % geo.nx, cy = cell / geo.nx;
        const bool inside = 2 * cx + 1 > geo.nx && 2 * cy + 1 > geo.ny;
        g[i] = (inside && local < P) ? B->mass[local] : 0.0;
    }
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
        if (g.nq > N) k_assemble_oi<N, double, false><<<count, BLOCK, 0, s>>>(vcache, fcache, blocks, inv_step, b, g, first);
        else k_assemble<N, double, false><<<count, BLOCK, 0, s>>>(vcache, fcache, blocks, inv_step, b, g, first);
    })
}
hipError_t launch_assemble_blocks(int n, const double *vcache, const double *fcache, float *blocks,
                                  double inv_step, const Basis *b, const Geo &g, int first, int count,
                                  hipStream_t s) {
    DGH_DISPATCH_N(n, {
        if (g.nq > N) k_assemble_oi<N, float, false><<<count, BLOCK, 0, s>>>(vcache, fcache, blocks, inv_step, b, g, first);
        else k_assemble<N, float, false><<<count, BLOCK, 0, s>>>(vcache, fcache, blocks, inv_step, b, g, first);
    })
}
template <typename T>
static hipError_t assemble_colour_any(int n, const double *vcache, const double *fcache, T *blocks, double inv_step,
                                      const Basis *b, const Geo &g, int colour, hipStream_t s) {
    const int count = g.cells / 2;
    DGH_DISPATCH_N(n, {
        if (g.nq > N) k_assemble_oi<N, T, false><<<count, BLOCK, 0, s>>>(vcache, fcache, blocks, inv_step, b, g, 0, colour);
        else k_assemble<N, T, false><<<count, BLOCK, 0, s>>>(vcache, fcache, blocks, inv_step, b, g, 0, colour);
    })
}
hipError_t launch_assemble_colour(int n, const double *vcache, const double *fcache, double *blocks,
                                  double inv_step, const Basis *b, const Geo &g, int colour, hipStream_t s) {
    return assemble_colour_any(n, vcache, fcache, blocks, inv_step, b, g, colour, s);
}
hipError_t launch_assemble_colour(int n, const double *vcache, const double *fcache, float *blocks,
                                  double inv_step, const Basis *b, const Geo &g, int colour, hipStream_t s) {
    return assemble_colour_any(n, vcache, fcache, blocks, inv_step, b, g, colour, s);
}
hipError_t launch_assemble_blocks_transposed(int n, const double *vcache, const double *fcache, double *blocks,
                                             double inv_step, const Basis *b, const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, {
        if (g.nq > N) k_assemble_oi<N, double, false, true><<<g.cells, BLOCK, 0, s>>>(vcache, fcache, blocks, inv_step, b, g, 0);
        else k_assemble<N, double, false, true><<<g.cells, BLOCK, 0, s>>>(vcache, fcache, blocks, inv_step, b, g, 0);
    })
}
hipError_t launch_assemble_hdg(int n, const double *vcache, const double *fcache, double *blocks, double inv_step,
                               const Basis *b, const Geo &g, int first, int count, hipStream_t s) {
    DGH_DISPATCH_N(n, {
        if (g.nq > N) k_assemble_oi<N, double, true><<<count, BLOCK, 0, s>>>(vcache, fcache, blocks, inv_step, b, g, first);
        else k_assemble<N, double, true><<<count, BLOCK, 0, s>>>(vcache, fcache, blocks, inv_step, b, g, first);
    })
}
// Panel width and workgroup size of k_lu. A wider panel saves passes over
// the trailing matrix but its LDS panel and register tile cap the resident
// workgroups per CU. Measured at 256^2, k=8 (m = 324 = 27 x 12), FP64 / FP32
// ms per refresh: NB = 8: 306 / 211; 10: 343 / 196; 12: 224 / 149;
// 14: 252 / 192; 16: 245 / 179; 20: 322 / 204; 24: 286 / 219; 32: 1066 / 338.
#ifndef DGH_LU_NB64
#define DGH_LU_NB64 12
#endif
#ifndef DGH_LU_NB32
#define DGH_LU_NB32 12
#endif
// The k = 8 element blocks (m = 324 = 18 x 18) take 18-column panels since the
// DPP pivot search: FP64 / FP32 ms per refresh of 65,536 blocks, NB = 8: 299 /
// 207; 12: 216 / 141; 16: 227 / 176; 18: 176 / 136; 20 / 22: - / 202, 249 / -;
// 27: 216 / 144; 36: 503 / 183 (factors bit-identical at every width).
#ifndef DGH_LU_NB64_324
#define DGH_LU_NB64_324 18
#endif
#ifndef DGH_LU_NB32_324
#define DGH_LU_NB32_324 18
#endif
// Trailing columns per wavefront pass (each LDS read of L21 feeds CB
// columns), m = 324, NB = 18, ms per refresh FP64 / FP32: CB = 1: 176 / 136;
// 2: 161 / 111; 3: 165 / 104; 4: 210 / 103; 6: - / 105 (NB = 12: CB = 2 / 4:
// 204 / 115; 3 / 6: 204 / 114); factors bit-identical throughout.
#ifndef DGH_LU_CB64
#define DGH_LU_CB64 2
#endif
#ifndef DGH_LU_CB32
#define DGH_LU_CB32 4
#endif
// FP64 blocks factor faster with 8 waves per workgroup, FP32 blocks with 4
// (measured, 256^2 k=8, NB = 12: 224 vs 248 ms (256 threads) and 149 vs
// 171 ms (512 threads))
#ifndef DGH_LU_THREADS64
#define DGH_LU_THREADS64 512
#endif
#ifndef DGH_LU_THREADS32
#define DGH_LU_THREADS32 256
#endif
// k_lu_mfma (m = 324 only) is selected with DGH_LU_MFMA=1 in the environment.
// Measured at 256^2, k = 8, ms per refresh, resident trailing tile:
//   FP32 (256 threads): NB = 12: 177.5; 16: 167.7; 24: 153.7; 32: 158.5 (k_lu: 149.8)
//   FP64 (256 threads): NB = 12: 307.6; 16: 276.8; 24: 439.7           (k_lu: 223.7)
// Its trailing update is faster than k_lu's (FP32 NB = 24: 82 ms against
// 108 ms) but the unblocked panel grows with NB (FP32 panels alone: NB = 12:
// 42 ms, NB = 24: 71 ms), so k_lu stays the default.
#ifndef DGH_LU_MFMA_NB64
#define DGH_LU_MFMA_NB64 16
#endif
#ifndef DGH_LU_MFMA_NB32
#define DGH_LU_MFMA_NB32 24
#endif
#ifndef DGH_LU_MFMA_THREADS64
#define DGH_LU_MFMA_THREADS64 256
#endif
#ifndef DGH_LU_MFMA_THREADS32
#define DGH_LU_MFMA_THREADS32 256
#endif
// read at each launch, so that a caller can compare the kernels on the same blocks
static bool lu_use_mfma() {
    const char *e = std::getenv("DGH_LU_MFMA");
    return e && e[0] == '1';
}
hipError_t launch_lu(int m, double *blocks, int *rowsrc, int nblocks, int *flag, hipStream_t s) {
    if (m == 324 && lu_use_mfma()) {
        k_lu_mfma<324, DGH_LU_MFMA_NB64, double, DGH_LU_MFMA_THREADS64, true>
            <<<nblocks, DGH_LU_MFMA_THREADS64, 0, s>>>(blocks, rowsrc, nblocks, flag);
        return hipGetLastError();
    }
    if (m == 324) {
        k_lu<324, DGH_LU_NB64_324, double, DGH_LU_THREADS64, DGH_LU_CB64>
            <<<nblocks, DGH_LU_THREADS64, 0, s>>>(blocks, rowsrc, nblocks, flag, nullptr);
        return hipGetLastError();
    }
    DGH_DISPATCH_M(m, {
        k_lu<M, DGH_LU_NB64, double, DGH_LU_THREADS64>
            <<<nblocks, DGH_LU_THREADS64, 0, s>>>(blocks, rowsrc, nblocks, flag, nullptr);
    })
}
hipError_t launch_lu(int m, float *blocks, int *rowsrc, int nblocks, int *flag, hipStream_t s, float *umax) {
    if (m == 324 && lu_use_mfma()) {
        k_lu_mfma<324, DGH_LU_MFMA_NB32, float, DGH_LU_MFMA_THREADS32, true>
            <<<nblocks, DGH_LU_MFMA_THREADS32, 0, s>>>(blocks, rowsrc, nblocks, flag);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess || !umax) return e;
        k_umax<<<nblocks, BLOCK, 0, s>>>(blocks, umax, m);
        return hipGetLastError();
    }
    if (m == 324) {
        k_lu<324, DGH_LU_NB32_324, float, DGH_LU_THREADS32, DGH_LU_CB32>
            <<<nblocks, DGH_LU_THREADS32, 0, s>>>(blocks, rowsrc, nblocks, flag, umax);
        return hipGetLastError();
    }
    DGH_DISPATCH_M(m, {
        k_lu<M, DGH_LU_NB32, float, DGH_LU_THREADS32>
            <<<nblocks, DGH_LU_THREADS32, 0, s>>>(blocks, rowsrc, nblocks, flag, umax);
    })
}
hipError_t launch_bj_apply(int m, const double *factors, const int *rowsrc, const double *r, double *z,
                           int nblocks, hipStream_t s) {
    DGH_DISPATCH_M(m, {
        k_bj_apply<M, double, double, BJ_RING64><<<grid_for(nblocks, BLOCK / 64), BLOCK, 0, s>>>(factors, rowsrc, r, z, nblocks, -1, 0);
    })
}
hipError_t launch_bj_apply_colour(int m, const double *factors, const int *rowsrc, const double *r, double *z,
                                  int nblocks, int colour, int nx, hipStream_t s) {
    DGH_DISPATCH_M(m, {
        k_bj_apply<M, double, double, BJ_RING64><<<grid_for(nblocks, BLOCK / 64), BLOCK, 0, s>>>(factors, rowsrc, r, z, nblocks, colour, nx);
    })
}
hipError_t launch_bj_apply_colour(int m, const float *factors, const int *rowsrc, const float *r, float *z,
                                  int nblocks, int colour, int nx, hipStream_t s) {
    DGH_DISPATCH_M(m, {
        k_bj_apply<M, float, float, BJ_RING32><<<grid_for(nblocks, BLOCK / 64), BLOCK, 0, s>>>(factors, rowsrc, r, z, nblocks, colour, nx);
    })
}
hipError_t launch_bj_apply(int m, const float *factors, const int *rowsrc, const float *r, float *z, int nblocks,
                           hipStream_t s) {
    DGH_DISPATCH_M(m, {
        k_bj_apply<M, float, float, BJ_RING32><<<grid_for(nblocks, BLOCK / 64), BLOCK, 0, s>>>(factors, rowsrc, r, z, nblocks, -1, 0);
    })
}
// 2-byte factor layouts: column pairs (the default) or row pairs (kept for
// comparison); the caller fixes the layout per factorisation
bool two_byte_rowpairs_requested() { return std::getenv("DGH_BJ_ROWPAIRS") != nullptr; }
long long pair_factor_stride(int m, bool rowpairs) { return rowpairs ? pair_stride(m) : colpair_stride(m); }
hipError_t launch_to_pairs(int m, const float *square, void *out, float *sigma, float *rdiag, const float *umax,
                           int nblocks, int fp16, bool rowpairs, int *flag, hipStream_t s) {
    if (m % 2 != 0 || (fp16 && !umax)) return hipErrorInvalidValue;
    if (!rowpairs)
        k_to_colpairs<<<nblocks, BLOCK, 0, s>>>(square, static_cast<unsigned short *>(out), sigma, rdiag, umax, m,
                                                fp16, flag);
    else
        k_to_pairs<<<nblocks, BLOCK, 0, s>>>(square, static_cast<unsigned short *>(out), sigma, rdiag, umax, m, fp16,
                                             flag);
    return hipGetLastError();
}
// columns in flight per wave (measured at 256^2, k=8: FP16 2.67 / 2.77 / 3.18 ms
// and BF16 2.68 / 2.53 / 3.30 ms with 4 / 8 / 2 columns)
#ifndef DGH_BJ_RING_FP16
#define DGH_BJ_RING_FP16 4
#endif
#ifndef DGH_BJ_RING_BF16
#define DGH_BJ_RING_BF16 8
#endif
#ifndef DGH_BJ_RING_CP
#define DGH_BJ_RING_CP 4
#endif
hipError_t launch_bj_apply_pairs(int m, const void *factors, const float *sigma, const float *rdiag,
                                 const int *rowsrc, const float *r, float *z, int nblocks, int fp16, bool rowpairs, hipStream_t s) {
    const auto *f = static_cast<const unsigned short *>(factors);
    if (!rowpairs) {
        DGH_DISPATCH_M(m, {
            if (fp16)
                k_bj_apply_colpairs<M, true, DGH_BJ_RING_CP><<<grid_for(nblocks, BLOCK / 64), BLOCK, 0, s>>>(f, sigma, rdiag, rowsrc, r, z, nblocks);
            else
                k_bj_apply_colpairs<M, false, DGH_BJ_RING_CP><<<grid_for(nblocks, BLOCK / 64), BLOCK, 0, s>>>(f, sigma, rdiag, rowsrc, r, z, nblocks);
        })
    }
    DGH_DISPATCH_M(m, {
        if (fp16)
            k_bj_apply_pairs<M, true, DGH_BJ_RING_FP16><<<grid_for(nblocks, BLOCK / 64), BLOCK, 0, s>>>(f, sigma, rdiag, rowsrc, r, z, nblocks);
        else
            k_bj_apply_pairs<M, false, DGH_BJ_RING_BF16><<<grid_for(nblocks, BLOCK / 64), BLOCK, 0, s>>>(f, sigma, rdiag, rowsrc, r, z, nblocks);
    })
}
// In place, `chunk` blocks at a time through `scratch` (chunk * m*m entries):
// copy the square blocks aside, then write them back packed.
template <typename T>
static hipError_t pack_any(int m, T *blocks, int nblocks, T *scratch, int chunk, hipStream_t s) {
    const size_t mm = static_cast<size_t>(m) * m;
    for (int b0 = 0; b0 < nblocks; b0 += chunk) {
        const int count = min(chunk, nblocks - b0);
        hipError_t e = hipMemcpyAsync(scratch, blocks + b0 * mm, count * mm * sizeof(T), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
        k_pack<T><<<count, BLOCK, 0, s>>>(scratch, blocks + b0 * mm, m);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
hipError_t launch_pack(int m, double *blocks, int nblocks, double *scratch, int chunk, hipStream_t s) {
    return pack_any(m, blocks, nblocks, scratch, chunk, s);
}
hipError_t launch_pack(int m, float *blocks, int nblocks, float *scratch, int chunk, hipStream_t s) {
    return pack_any(m, blocks, nblocks, scratch, chunk, s);
}
hipError_t launch_convert(const double *src, float *dst, long long count, hipStream_t s) {
    k_convert<double, float><<<4096, BLOCK, 0, s>>>(src, dst, count);
    return hipGetLastError();
}
hipError_t launch_convert(const float *src, double *dst, long long count, hipStream_t s) {
    k_convert<float, double><<<4096, BLOCK, 0, s>>>(src, dst, count);
    return hipGetLastError();
}

// calls f(std::integral_constant<int, nv>) for nv in [NV, MAXV]
template <int NV, class F> static hipError_t nv_dispatch(int nv, F &&f) {
    if constexpr (NV > MAXV) {
        (void)nv;
        (void)f;
        return hipErrorInvalidValue;
    } else {
        if (nv == NV) {
            f(std::integral_constant<int, NV>{});
            return hipGetLastError();
        }
        return nv_dispatch<NV + 1>(nv, f);
    }
}

template <typename T> static bool aligned16(const T *p) { return (reinterpret_cast<std::uintptr_t>(p) & 15) == 0; }
template <typename T> static bool basis_aligned(const T *const *basis, int nvec, const T *w) {
    for (int k = 0; k < nvec; ++k)
        if (!aligned16(basis[k])) return false;
    return aligned16(w);
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
    if (nvec < 1 || count >= (1LL << 31) || n < 1 || n > MAXN || !basis_aligned(basis, nvec, w))
        return hipErrorInvalidValue;
    const auto p = make_ptrs(basis, nvec);
    return nv_dispatch<1>(nvec, [&](auto c) {
        constexpr int NV = decltype(c)::value;
        k_mdot<TV, NV><<<REDUCE_BLOCKS, BLOCK, 0, s>>>(p, w, weights, static_cast<int>(count), n * n, partials, b);
    });
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
    if (nvec < 1 || count >= (1LL << 31) || n < 1 || n > MAXN || !basis_aligned(basis, nvec, w))
        return hipErrorInvalidValue;
    const auto p = make_ptrs(basis, nvec);
    return nv_dispatch<1>(nvec, [&](auto c) {
        constexpr int NV = decltype(c)::value;
        k_mupdate<TV, NV><<<REDUCE_BLOCKS, BLOCK, 0, s>>>(p, h, w, weights, static_cast<int>(count), n * n, partials,
                                                         mode, b);
    });
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
hipError_t launch_quadrant_weights(int n, double *g, const Basis *b, const Geo &geo, hipStream_t s) {
    DGH_DISPATCH_N(n, { k_quadrant_weights<N><<<4096, BLOCK, 0, s>>>(g, b, geo); })
}
hipError_t launch_arnoldi_givens(int j, int restart, const double *scalars, double *H, double *cs, double *sn,
                                 double *g, double *next, double *out, int passes, hipStream_t s) {
    k_arnoldi_givens<<<1, 64, 0, s>>>(j, restart, scalars, H, cs, sn, g, next, out, passes);
    return hipGetLastError();
}
hipError_t launch_scale_div_dev(const double *x, const double *beta, double *out, long long count, hipStream_t s) {
    k_scale_div_dev<double><<<4096, BLOCK, 0, s>>>(x, beta, out, count);
    return hipGetLastError();
}
hipError_t launch_scale_div_dev(const float *x, const double *beta, float *out, long long count, hipStream_t s) {
    k_scale_div_dev<float><<<4096, BLOCK, 0, s>>>(x, beta, out, count);
    return hipGetLastError();
}
hipError_t launch_dot(const double *a, const double *b, long long count, double *partials, hipStream_t s) {
    k_dot<<<REDUCE_BLOCKS, BLOCK, 0, s>>>(a, b, count, partials);
    return hipGetLastError();
}
hipError_t launch_mass_combine(int n, const double *lambda, const double *k, double tau, double *out, long long count,
                               const Basis *b, hipStream_t s) {
    DGH_DISPATCH_N(n, { k_mass_combine<N><<<4096, BLOCK, 0, s>>>(lambda, k, tau, out, count, b); })
}
hipError_t launch_sub_div(const double *x, const double *y, double d, double *out, long long count, hipStream_t s) {
    k_sub_div<<<4096, BLOCK, 0, s>>>(x, y, d, out, count);
    return hipGetLastError();
}

}  // namespace dgh
