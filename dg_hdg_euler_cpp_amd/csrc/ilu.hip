// Block ILU(0) of the DG element graph (block_ilu.cpp:3-84) in red-black
// order on the device.
//
// With the cells of colour 0 ("red", x + y even) ordered before those of
// colour 1 ("black") on an even periodic grid of at least 4 x 4 cells, every
// neighbour of a cell has the other colour. ILU(0) in that order then has no
// update except on the black diagonal blocks (block_ilu.cpp:15-47 finds no
// other target), and L_bj = A_bj A_jj^-1 for every red neighbour j:
//   factorisation: A_rr = LU (red); A_bb -= sum_j A_bj A_jj^-1 A_jb; LU (black);
//   apply (block_ilu.cpp:62-84): z_r = A_rr^-1 r_r; w_b = r_b - sum_j A_bj z_j;
//     x_b = A_bb'^-1 w_b; x_r = A_rr^-1 (r_r - sum_b A_rb x_b).
// An off-diagonal block couples two elements only through their shared face:
// A_en = E_e J R_n, with R_n the restriction of n to the face (4n values), J the
// per-node 4x4 derivative of the face flux with respect to n's trace
// (face_jacobian.cpp:56-64: bl A_L + pen I for the left trace, br A_R - pen I
// for the right one) and E_e the lift into e (+shape on e's +x/+y sides,
// -shape on its -x/-y sides; discretization.cpp:61-71). So the diagonal update
// of a black cell per side is E_b J1 (R_j A_jj^-1 E_j) J2 R_b: the 4n x 4n
// middle factor S comes from a block solve of the red cell against its 4n lift
// columns, and the rest is an outer product over the face lines.
#include "dgh_internal.hpp"
#include "element_solve.hpp"

namespace dgh {
namespace {

constexpr int BLOCK = 256;
#ifndef DGH_ILU_WAVES
#define DGH_ILU_WAVES 8
#endif
#ifndef DGH_ILU_COLS
#define DGH_ILU_COLS 8
#endif
constexpr int ILU_WAVES = DGH_ILU_WAVES, ILU_COLS = DGH_ILU_COLS, ILU_RING = 2;

// physics.cpp:42-57: flux Jacobian along `axis`, row-major [v][w]
__device__ __forceinline__ void flux_jacobian(double u, double v, double H, int axis, double gamma, double a[16]) {
    const double gm = gamma - 1., s2 = u * u + v * v;
    if (axis == 0) {
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
}

// f = A d with A the flux Jacobian of physics.cpp:42-57 along `axis`
__device__ __forceinline__ void flux_jacobian_action(double u, double v, double H, int axis, double gamma,
                                                     const double d[4], double f[4]) {
    const double gm = gamma - 1., s2 = u * u + v * v;
    if (axis == 0) {
        f[0] = d[1];
        f[1] = (.5 * (gamma - 3) * u * u + .5 * gm * v * v) * d[0] + (3 - gamma) * u * d[1] - gm * v * d[2] + gm * d[3];
        f[2] = -u * v * d[0] + v * d[1] + u * d[2];
        f[3] = u * (.5 * gm * s2 - H) * d[0] + (H - gm * u * u) * d[1] - gm * u * v * d[2] + gamma * u * d[3];
    } else {
        f[0] = d[2];
        f[1] = -u * v * d[0] + v * d[1] + u * d[2];
        f[2] = (.5 * gm * u * u + .5 * (gamma - 3) * v * v) * d[0] - gm * u * d[1] + (3 - gamma) * v * d[2] + gm * d[3];
        f[3] = v * (.5 * gm * s2 - H) * d[0] - gm * u * v * d[1] + (H - gm * v * v) * d[2] + gamma * v * d[3];
    }
}

// Derivative of the face flux at node q of `face` with respect to the left
// (left = true) or the right trace, row-major [v][w].
__device__ __forceinline__ void trace_derivative(const double *__restrict__ fc, long long nodes, int face, int q,
                                                 bool left, const Geo &g, double a[16]) {
    const long long fn = static_cast<long long>(face) * g.n + q;
    flux_jacobian(fc[(left ? F_UL : F_UR) * nodes + fn], fc[(left ? F_VL : F_VR) * nodes + fn],
                  fc[(left ? F_HL : F_HR) * nodes + fn], face >= g.cells, g.gamma, a);
    const double blend = fc[(left ? F_BL : F_BR) * nodes + fn], pen = fc[F_PEN * nodes + fn];
    for (int k = 0; k < 16; ++k) a[k] *= blend;
    for (int k = 0; k < 4; ++k) a[k * 5] += left ? pen : -pen;
}

// Y[k] = A^-1 E and S[k][s] = R_s A^-1 E_s for the k-th red cell (A factorised
// in place by k_lu, compact red storage). Column c = s*4n + v*n + q of E is the
// lift of a unit face value (v, q) on side s: +right(a) (odd s) or -left(a)
// (even s) at rows (v, face_node(s, q, a)). Y[k] is column-major m x 16n;
// S[k][s] is column-major 4n x 4n with row w*n + q' (trace of variable w at
// node q') and column v*n + q. TF is the factors' precision (FP64, or FP32
// for the mixed-precision preconditioner); Y is stored in it.
template <int N, typename TF>
__global__ __launch_bounds__(ILU_WAVES * 64) void k_ilu_coupling(const TF *__restrict__ factors,
                                                                 const int *__restrict__ rowsrc,
                                                                 TF *__restrict__ Y, double *__restrict__ S,
                                                                 int col0, const Basis *__restrict__ B, Geo g) {
    constexpr int P = N * N, M = 4 * P, TS = 4 * N, T = 4 * TS, NA = (M + 63) / 64;
    __shared__ TF colbuf[ILU_WAVES][M];
    const int lane = threadIdx.x % 64, wave = wave_index();
    const long long k = blockIdx.x;
    const TF *F = factors + k * M * M;
    const int *perm = rowsrc + k * M;
    TF x[ILU_COLS][NA];
    int cols[ILU_COLS];
#pragma unroll
    for (int cc = 0; cc < ILU_COLS; ++cc) {
        const int c = col0 + wave + ILU_WAVES * cc;
        cols[cc] = c;
        const int s = c / TS, v = (c / N) % 4, q = c % N;
        const double *shape = s % 2 ? B->right : B->left;
        const double sign = s % 2 ? 1. : -1.;
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            const int row = a * 64 + lane;
            double val = 0;
            if (row < M && c < T) {
                const int src = perm[row];  // (P e)[row] = e[rowsrc[row]]
                if (src / P == v) {
                    const int node = src % P, in = node % N, jn = node / N;
                    if (s < 2 && jn == q) val = sign * shape[in];
                    if (s >= 2 && in == q) val = sign * shape[jn];
                }
            }
            x[cc][a] = static_cast<TF>(val);
        }
    }
    lu_solve_columns<M, ILU_COLS, ILU_RING>(F, x, lane);
    // restriction of each solved column to its own side; the column goes
    // through LDS so that lane r36 can sum the n values of its face line
#pragma unroll
    for (int cc = 0; cc < ILU_COLS; ++cc) {
        const int c = cols[cc];
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            const int row = a * 64 + lane;
            if (row < M) {
                colbuf[wave][row] = x[cc][a];
                if (c < T) Y[(k * T + c) * M + row] = x[cc][a];
            }
        }
        __syncthreads();
        if (c < T && lane < TS) {
            const int s = c / TS, w = lane / N, q = lane % N;
            const double *shape = s % 2 ? B->right : B->left;
            double acc = 0;
#pragma unroll
            for (int a = 0; a < N; ++a) acc += shape[a] * static_cast<double>(colbuf[wave][w * P + face_node(s, q, a, N)]);
            S[(k * 4 + s) * TS * TS + lane + static_cast<long long>(TS) * (c % TS)] = acc;
        }
        __syncthreads();
    }
}

// A_bb -= sum over the sides s of E_s J1 S_nb J2 R_s for the k-th black cell
// (compact black storage; nb = the red neighbour on side s, its facing side
// s ^ 1). J1 = derivative with respect to the neighbour's trace and J2 with
// respect to the cell's own trace at the shared face.
template <int N, typename T>
__global__ __launch_bounds__(BLOCK) void k_ilu_update(T *__restrict__ blocks, const double *__restrict__ S,
                                                      const double *__restrict__ fc, const Basis *__restrict__ B,
                                                      Geo g) {
    constexpr int P = N * N, M = 4 * P, TS = 4 * N;
    __shared__ double Tm[4][TS * TS];
    __shared__ double tmp[TS * TS];
    __shared__ double J1[4][N][16], J2[4][N][16];
    __shared__ double tl[N], tr[N];
    const long long k = blockIdx.x;
    const int cell = colour_cell(static_cast<int>(k), 1, g.nx);
    const long long nodes = 2LL * g.cells * N;
    for (int t = threadIdx.x; t < N; t += BLOCK) {
        tl[t] = B->left[t];
        tr[t] = B->right[t];
    }
    for (int t = threadIdx.x; t < 4 * N; t += BLOCK) {
        const int s = t / N, q = t % N, face = cell_face(cell, s, g);
        const bool left = s % 2;  // the cell is the left element of its +x / +y face
        trace_derivative(fc, nodes, face, q, !left, g, J1[s][q]);
        trace_derivative(fc, nodes, face, q, left, g, J2[s][q]);
    }
    __syncthreads();
    for (int s = 0; s < 4; ++s) {
        const int nb = neighbor(cell, s, g.nx, g.ny);
        const double *Sp = S + (static_cast<long long>(colour_index(nb, g.nx)) * 4 + (s ^ 1)) * TS * TS;
        // tmp[(v, q) row, (v', q') col] = sum_w1 J1[q][v][w1] S[(w1, q), (v', q')]
        for (int e = threadIdx.x; e < TS * TS; e += BLOCK) {
            const int r = e % TS, c = e / TS, v = r / N, q = r % N;
            double acc = 0;
            for (int w1 = 0; w1 < 4; ++w1) acc += J1[s][q][v * 4 + w1] * Sp[w1 * N + q + TS * c];
            tmp[e] = acc;
        }
        __syncthreads();
        // Tm[(v, q), (w, q')] = sum_v' tmp[(v, q), (v', q')] J2[q'][v'][w]
        for (int e = threadIdx.x; e < TS * TS; e += BLOCK) {
            const int r = e % TS, c = e / TS, w = c / N, qq = c % N;
            double acc = 0;
            for (int v1 = 0; v1 < 4; ++v1) acc += tmp[r + TS * (v1 * N + qq)] * J2[s][qq][v1 * 4 + w];
            Tm[s][e] = acc;
        }
        __syncthreads();
    }
    T *A = blocks + k * M * M;
    for (int col = 0; col < M; ++col) {
        const int w = col / P, cnode = col % P, ci = cnode % N, cj = cnode / N;
        for (int row = threadIdx.x; row < M; row += BLOCK) {
            const int v = row / P, rnode = row % P, ri = rnode % N, rj = rnode / N;
            double delta = 0;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                // x sides: line q = j, position a = i; y sides: q = i, a = j
                const int rq = s < 2 ? rj : ri, ra = s < 2 ? ri : rj;
                const int cq = s < 2 ? cj : ci, ca = s < 2 ? ci : cj;
                const double e = s % 2 ? tr[ra] : -tl[ra];
                const double rr = s % 2 ? tr[ca] : tl[ca];
                delta += e * Tm[s][v * N + rq + TS * (w * N + cq)] * rr;
            }
            T &entry = A[static_cast<long long>(col) * M + row];
            entry = static_cast<T>(static_cast<double>(entry) - delta);
        }
    }
}

// Back substitution of the red cells without a second block solve:
// z_r = A_rr^-1 (r_r - sum_s A_{r,b_s} x_b) = z0_r - sum_s Y_r^s (J_s R_b x_b),
// with z0_r = A_rr^-1 r_r from the forward sweep, Y_r = A_rr^-1 E_r and J_s the
// derivative of the side-s face flux with respect to the black neighbour's
// trace (given by tr). One wavefront per red cell; lane l owns rows l + 64a.
template <int N, typename TY, typename TV>
__global__ __launch_bounds__(BLOCK) void k_ilu_red_back(const TY *__restrict__ Y, const double *__restrict__ tr,
                                                        const double *__restrict__ fc, const TV *__restrict__ z0,
                                                        TV *__restrict__ z, Geo g) {
    constexpr int P = N * N, M = 4 * P, T = 16 * N, NA = (M + 63) / 64, NG = (T + 63) / 64;
    const int lane = threadIdx.x % 64;
    const long long k = static_cast<long long>(blockIdx.x) * (BLOCK / 64) + wave_index();
    if (k >= g.cells / 2) return;
    const int cell = colour_cell(static_cast<int>(k), 0, g.nx);
    const long long nodes = 2LL * g.cells * N;
    // gv[b] holds column c = 64 b + lane of J t: side s, variable v, node q
    double gv[NG];
#pragma unroll
    for (int b = 0; b < NG; ++b) {
        const int c = 64 * b + lane;
        double val = 0;
        if (c < T) {
            const int s = c / (4 * N), v = (c / N) % 4, q = c % N;
            const int nb = neighbor(cell, s, g.nx, g.ny), face = cell_face(cell, s, g);
            const bool nb_left = s % 2 == 0;
            const long long fn = static_cast<long long>(face) * N + q;
            double t[4], f[4];
            for (int w = 0; w < 4; ++w) t[w] = tr[static_cast<long long>(nb) * 16 * N + (s ^ 1) * 4 * N + w * N + q];
            flux_jacobian_action(fc[(nb_left ? F_UL : F_UR) * nodes + fn], fc[(nb_left ? F_VL : F_VR) * nodes + fn],
                                 fc[(nb_left ? F_HL : F_HR) * nodes + fn], s / 2, g.gamma, t, f);
            const double blend = fc[(nb_left ? F_BL : F_BR) * nodes + fn];
            const double pen = nb_left ? fc[F_PEN * nodes + fn] : -fc[F_PEN * nodes + fn];
            val = blend * f[v] + pen * t[v];
        }
        gv[b] = val;
    }
    double acc[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) acc[a] = 0;
    const TY *Yk = Y + k * T * M;
#pragma unroll 4
    for (int c = 0; c < T; ++c) {
        const double gc = readlane_d(gv[c / 64], c % 64);
        const TY *col = Yk + static_cast<long long>(c) * M;
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            const int row = a * 64 + lane;
            if (row < M) acc[a] += static_cast<double>(__builtin_nontemporal_load(col + row)) * gc;
        }
    }
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const int row = a * 64 + lane;
        if (row < M)
            z[static_cast<long long>(cell) * M + row] =
                static_cast<TV>(static_cast<double>(z0[static_cast<long long>(cell) * M + row]) - acc[a]);
    }
}

// out = rhs - sum_s A_{cell, nb_s} x_nb for the cells of `colour`, with x given
// by the restricted side traces tr[cell][side][v][q].
template <int N, typename TV>
__global__ __launch_bounds__(BLOCK) void k_ilu_offdiag(int colour, const double *__restrict__ tr,
                                                       const double *__restrict__ fc, const TV *__restrict__ rhs,
                                                       TV *__restrict__ out, const Basis *__restrict__ B,
                                                       Geo g) {
    constexpr int P = N * N, M = 4 * P, CPB = BLOCK / P > 0 ? BLOCK / P : 1;
    const int local = threadIdx.x / P, node = threadIdx.x % P;
    const long long k = static_cast<long long>(blockIdx.x) * CPB + local;
    if (local >= CPB || k >= g.cells / 2) return;
    const int cell = colour_cell(static_cast<int>(k), colour, g.nx);
    const int i = node % N, j = node / N;
    const long long nodes = 2LL * g.cells * N;
    double res[4];
    for (int v = 0; v < 4; ++v) res[v] = static_cast<double>(rhs[static_cast<long long>(cell) * M + v * P + node]);
#pragma unroll 1
    for (int s = 0; s < 4; ++s) {
        const int nb = neighbor(cell, s, g.nx, g.ny), face = cell_face(cell, s, g);
        const int q = s < 2 ? j : i, a = s < 2 ? i : j;
        // the cell is the left element of its +x / +y faces, so there the
        // neighbour's trace is the right one: br A_R - pen I, else bl A_L + pen I
        const bool nb_left = s % 2 == 0;
        const long long fn = static_cast<long long>(face) * N + q;
        double t[4], f[4];
        for (int w = 0; w < 4; ++w) t[w] = tr[static_cast<long long>(nb) * 16 * N + (s ^ 1) * 4 * N + w * N + q];
        flux_jacobian_action(fc[(nb_left ? F_UL : F_UR) * nodes + fn], fc[(nb_left ? F_VL : F_VR) * nodes + fn],
                             fc[(nb_left ? F_HL : F_HR) * nodes + fn], s / 2, g.gamma, t, f);
        const double blend = fc[(nb_left ? F_BL : F_BR) * nodes + fn];
        const double pen = nb_left ? fc[F_PEN * nodes + fn] : -fc[F_PEN * nodes + fn];
        const double e = nb_left ? -B->left[a] : B->right[a];
        for (int v = 0; v < 4; ++v) res[v] -= e * (blend * f[v] + pen * t[v]);
    }
    for (int v = 0; v < 4; ++v) out[static_cast<long long>(cell) * M + v * P + node] = static_cast<TV>(res[v]);
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

template <typename TF>
static hipError_t ilu_coupling_any(int n, const TF *factors_red, const int *rowsrc_red, TF *Y, double *S,
                                   const Basis *b, const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, {
        const int count = g.cells / 2, cols = 16 * N;
        for (int col0 = 0; col0 < cols; col0 += ILU_WAVES * ILU_COLS) {
            k_ilu_coupling<N, TF><<<count, ILU_WAVES * 64, 0, s>>>(factors_red, rowsrc_red, Y, S, col0, b, g);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
    })
}
hipError_t launch_ilu_coupling(int n, const double *factors_red, const int *rowsrc_red, double *Y, double *S,
                               const Basis *b, const Geo &g, hipStream_t s) {
    return ilu_coupling_any(n, factors_red, rowsrc_red, Y, S, b, g, s);
}
hipError_t launch_ilu_coupling(int n, const float *factors_red, const int *rowsrc_red, float *Y, double *S,
                               const Basis *b, const Geo &g, hipStream_t s) {
    return ilu_coupling_any(n, factors_red, rowsrc_red, Y, S, b, g, s);
}
template <typename T>
static hipError_t ilu_update_any(int n, T *blocks_black, const double *S, const double *fcache, const Basis *b,
                                 const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, { k_ilu_update<N, T><<<g.cells / 2, BLOCK, 0, s>>>(blocks_black, S, fcache, b, g); })
}
hipError_t launch_ilu_update(int n, double *blocks_black, const double *S, const double *fcache, const Basis *b,
                             const Geo &g, hipStream_t s) {
    return ilu_update_any(n, blocks_black, S, fcache, b, g, s);
}
hipError_t launch_ilu_update(int n, float *blocks_black, const double *S, const double *fcache, const Basis *b,
                             const Geo &g, hipStream_t s) {
    return ilu_update_any(n, blocks_black, S, fcache, b, g, s);
}
template <typename TY, typename TV>
static hipError_t ilu_red_back_any(int n, const TY *Y, const double *traces, const double *fcache, const TV *z0,
                                   TV *z, const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, {
        k_ilu_red_back<N, TY, TV><<<(g.cells / 2 + BLOCK / 64 - 1) / (BLOCK / 64), BLOCK, 0, s>>>(Y, traces, fcache,
                                                                                              z0, z, g);
    })
}
hipError_t launch_ilu_red_back(int n, const double *Y, const double *traces, const double *fcache, const double *z0,
                               double *z, const Geo &g, hipStream_t s) {
    return ilu_red_back_any(n, Y, traces, fcache, z0, z, g, s);
}
hipError_t launch_ilu_red_back(int n, const float *Y, const double *traces, const double *fcache, const float *z0,
                               float *z, const Geo &g, hipStream_t s) {
    return ilu_red_back_any(n, Y, traces, fcache, z0, z, g, s);
}
template <typename TV>
static hipError_t ilu_offdiag_any(int n, int colour, const double *traces, const double *fcache, const TV *rhs,
                                  TV *out, const Basis *b, const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, {
        constexpr int CPB = BLOCK / (N * N) > 0 ? BLOCK / (N * N) : 1;
        k_ilu_offdiag<N, TV><<<(g.cells / 2 + CPB - 1) / CPB, BLOCK, 0, s>>>(colour, traces, fcache, rhs, out, b, g);
    })
}
hipError_t launch_ilu_offdiag(int n, int colour, const double *traces, const double *fcache, const double *rhs,
                              double *out, const Basis *b, const Geo &g, hipStream_t s) {
    return ilu_offdiag_any(n, colour, traces, fcache, rhs, out, b, g, s);
}
hipError_t launch_ilu_offdiag(int n, int colour, const double *traces, const double *fcache, const float *rhs,
                              float *out, const Basis *b, const Geo &g, hipStream_t s) {
    return ilu_offdiag_any(n, colour, traces, fcache, rhs, out, b, g, s);
}

}  // namespace dgh
