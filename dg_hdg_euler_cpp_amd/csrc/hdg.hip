// Statically condensed HDG on the device (hybrid.cpp:31-121): element-streamed
// condensation (LU of the element block, 4*4n+1 right-hand sides, C * X),
// Schur apply, face blocks, recovery and the recovered DG residual.
//
// Element-local trace order: side*(4n) + v*n + q (hybrid.cpp:42); global
// trace index face*(4n) + v*n + q (discretization.cpp:148). X is stored per
// element column-major [m x t] (t = 16n), S_K column-major [t x t].
#include "dgh_internal.hpp"
#include "element_solve.hpp"

namespace dgh {
namespace {

constexpr int BLOCK = 256;


// FaceJacobian::hdg (face_jacobian.cpp:77-82): w ((+-1) A_phys + tau I), row-major [v][w].
__device__ void hdg_matrix(const double *__restrict__ fc, long long nodes, int face, int q, bool left,
                           const Basis *__restrict__ B, const Geo &g, double a[16]) {
    const long long fn = static_cast<long long>(face) * g.n + q;
    const double u = fc[(left ? F_UL : F_UR) * nodes + fn], v = fc[(left ? F_VL : F_VR) * nodes + fn];
    const double H = fc[(left ? F_HL : F_HR) * nodes + fn];
    const double tau = fc[(left ? F_TAUL : F_TAUR) * nodes + fn];
    const double gamma = g.gamma, gm = gamma - 1., s2 = u * u + v * v;
    if (face < g.cells) {
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
    const double weight = face_weight(face, q, B, g);
    for (int k = 0; k < 16; ++k) a[k] *= left ? 1. : -1.;
    for (int k = 0; k < 4; ++k) a[k * 5] += tau;
    for (int k = 0; k < 16; ++k) a[k] *= weight;
}

// ------------------------------------------------------- condensation solve
// [X | y] = A^-1 [B | r_K] (hybrid.cpp:59-64, dense_lu.hpp:33-37) with the
// element block already LU-factorised in place (k_lu). One 1024-thread
// workgroup per element; the right-hand sides live in registers: wave w owns
// columns w + 16c, lane l rows l + 64a. Every L/U column is read once per pass
// by all 16 waves (L1 hits) and applied to all of the pass's columns.
// B columns (hybrid.cpp:41-42): column (s, v, q) is -w tau l_s(i) at rows
// (v, face_node(s, q, i)); column t is the element's DG residual.
#ifndef DGH_SOLVE_WAVES
#define DGH_SOLVE_WAVES 8
#endif
#ifndef DGH_SOLVE_COLS
#define DGH_SOLVE_COLS 8
#endif
#ifndef DGH_SOLVE_RING
#define DGH_SOLVE_RING 2
#endif
constexpr int SOLVE_WAVES = DGH_SOLVE_WAVES, SOLVE_COLS = DGH_SOLVE_COLS;  // columns per pass: waves x cols
constexpr int SOLVE_RING = DGH_SOLVE_RING;  // factor columns in flight per wave

template <int N>
__global__ __launch_bounds__(SOLVE_WAVES * 64) void k_hdg_solve(const double *__restrict__ factors,
                                                              const int *__restrict__ rowsrc,
                                                              const double *__restrict__ fc,
                                                              const double *__restrict__ resid,
                                                              double *__restrict__ X, double *__restrict__ Y,
                                                              int first, int col0, const Basis *__restrict__ B,
                                                              Geo g) {
    constexpr int P = N * N, M = 4 * P, T = 16 * N, NA = (M + 63) / 64;
    const int lane = threadIdx.x % 64, wave = wave_index();
    const int local = blockIdx.x, cell = first + local;
    const double *F = factors + static_cast<long long>(local) * M * M;
    const int *perm = rowsrc + static_cast<long long>(local) * M;
    const long long nodes = 2LL * g.cells * N;
    double x[SOLVE_COLS][NA];
    int cols[SOLVE_COLS];
#pragma unroll
    for (int cc = 0; cc < SOLVE_COLS; ++cc) {
        const int c = col0 + wave + SOLVE_WAVES * cc;
        cols[cc] = c;
        int s = 0, v = 0, q = 0, face = 0;
        double stab = 0;
        if (c < T) {
            s = c / (4 * N);
            v = (c / N) % 4;
            q = c % N;
            face = cell_face(cell, s, g);
            const long long fn = static_cast<long long>(face) * N + q;
            stab = face_weight(face, q, B, g) * fc[(s % 2 ? F_TAUL : F_TAUR) * nodes + fn];
        }
        const double *shape = s % 2 ? B->right : B->left;
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            const int row = a * 64 + lane;
            double val = 0;
            if (row < M && c <= T) {
                const int src = perm[row];
                if (c == T) {
                    val = resid[static_cast<long long>(cell) * M + src];
                } else if (src / P == v) {
                    const int node = src % P, in = node % N, jn = node / N;
                    if (s < 2 && jn == q) val = -stab * shape[in];
                    if (s >= 2 && in == q) val = -stab * shape[jn];
                }
            }
            x[cc][a] = val;
        }
    }
    lu_solve_columns<M, SOLVE_COLS, SOLVE_RING>(F, x, lane);
#pragma unroll
    for (int cc = 0; cc < SOLVE_COLS; ++cc) {
        const int c = cols[cc];
        if (c > T) continue;
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            const int row = a * 64 + lane;
            if (row >= M) continue;
            if (c < T) X[(static_cast<long long>(cell) * T + c) * M + row] = x[cc][a];
            else Y[static_cast<long long>(cell) * M + row] = x[cc][a];
        }
    }
}

// ------------------------------------------- condensation solve on MFMA
// The same [X | y] = A^-1 [B | r_K] as k_hdg_solve, as a blocked triangular
// solve whose every product is a v_mfma_f64_16x16x4 tile product, over
// 16-row blocks (the last one padded with identity rows) with the inverses of
// the 16 x 16 diagonal blocks of L and U precomputed per element
// (k_diag_inverses). One wavefront owns 16 right-hand-side columns of one
// element and keeps its whole M x 16 tile in registers as MFMA accumulators,
// block row I in acc[I]. With the diagonal inverses stored negated every
// product accumulates:
//   forward   Y_J = (-Linv_JJ) t_J (= -X_J),  t_I += L_IJ Y_J   (I > J)
//   backward  X_J = (-Uinv_JJ) t'_J,          t'_I += U_IJ X_J  (I < J)
// where t' = -t starts as the forward's Y.
//
// Register layout. The f64 16x16x4 MFMA leaves D[l/16 + 4r][l%16] in lane l,
// register r. The accumulator rows are relabelled so that hardware row i holds
// block row tile_row(i) = 4 (i % 4) + i / 4: lane l then holds block rows
// 4 (l/16) .. 4 (l/16) + 3 of column l % 16, 32 contiguous bytes of an output
// column. Used as the B operand of k-slice r, register r of lane l is
// B[k = l/16][j = l%16], the value of block row 4 (l/16) + r, so the A operand
// of slice r is A[i = l%16][k = l/16] = L_IJ[tile_row(l%16)][4 (l/16) + r].
__device__ __forceinline__ int tile_row(int i) { return 4 * (i % 4) + i / 4; }

typedef double hdg_d4 __attribute__((ext_vector_type(4)));

// Negated inverses of the unit-lower and upper 16 x 16 diagonal blocks of the
// factors, padded with identity rows past row M, stored per (element, block,
// L|U) as 256 doubles in MFMA A-operand order (slice r, lane l). Lanes 0-15 of
// a 32-lane half invert L (column j = lane), lanes 16-31 invert U, by
// substitution against e_j.
template <int N>
__global__ __launch_bounds__(BLOCK) void k_diag_inverses(const double *__restrict__ factors,
                                                         double *__restrict__ dinv, int count) {
    constexpr int M = 4 * N * N, NT = (M + 15) / 16;
    const long long tid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long task = tid / 32;
    if (task >= static_cast<long long>(count) * NT) return;
    const int j = static_cast<int>(tid % 16), upper = static_cast<int>((tid / 16) % 2);
    const int local = static_cast<int>(task / NT), J = static_cast<int>(task % NT);
    const double *F = factors + static_cast<long long>(local) * M * M;
    const int base = 16 * J, valid = min(16, M - base);
    auto at = [&](int i, int k) {  // block entry (i, k), identity past the valid rows
        return (i < valid && k < valid) ? F[static_cast<long long>(base + k) * M + base + i] : (i == k ? 1.0 : 0.0);
    };
    double x[16];
    if (!upper) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            double s = i == j ? 1.0 : 0.0;
#pragma unroll
            for (int k = 0; k < i; ++k) s -= at(i, k) * x[k];
            x[i] = s;
        }
    } else {
#pragma unroll
        for (int i = 15; i >= 0; --i) {
            double s = i == j ? 1.0 : 0.0;
#pragma unroll
            for (int k = i + 1; k < 16; ++k) s -= at(i, k) * x[k];
            x[i] = s / at(i, i);
        }
    }
    double *D = dinv + (task * 2 + upper) * 256;
#pragma unroll
    for (int p = 0; p < 16; ++p) D[(j % 4) * 64 + (j / 4) * 16 + tile_row(p)] = -x[p];
}

// Logical workgroup index that keeps consecutive groups on one XCD: the
// hardware hands workgroup b to XCD b % 8, so logical groups are dealt out
// XCD by XCD (a bijection on 0..nblocks-1).
__device__ __forceinline__ int xcd_grouped(int b, int nblocks) {
    const int q = nblocks / 8, rem = nblocks % 8, xcd = b % 8, idx = b / 8;
    return (xcd < rem ? xcd * (q + 1) : rem * (q + 1) + (xcd - rem) * q) + idx;
}

#ifndef DGH_SOLVE_MFMA_WAVES
#define DGH_SOLVE_MFMA_WAVES 8
#endif
#ifndef DGH_SOLVE_AHEAD
#define DGH_SOLVE_AHEAD 2  // L / U tiles loaded ahead of their MFMAs
#endif
constexpr int MFMA_SOLVE_WAVES = DGH_SOLVE_MFMA_WAVES;

template <int N>
__global__ __launch_bounds__(MFMA_SOLVE_WAVES * 64) void k_hdg_solve_mfma(
    const double *__restrict__ factors, const int *__restrict__ rowsrc, const double *__restrict__ dinv,
    const double *__restrict__ fc, const double *__restrict__ resid, double *__restrict__ X, double *__restrict__ Y,
    int first, int count, const Basis *__restrict__ B, Geo g) {
    constexpr int P = N * N, M = 4 * P, T = 16 * N, NT = (M + 15) / 16, CT = (T + 1 + 15) / 16;
    constexpr int AHEAD = DGH_SOLVE_AHEAD, RING = AHEAD + 1;
    const int lane = threadIdx.x % 64, wave = wave_index();
    const long long tile = static_cast<long long>(xcd_grouped(blockIdx.x, gridDim.x)) * MFMA_SOLVE_WAVES + wave;
    if (tile >= static_cast<long long>(count) * CT) return;  // whole wavefront
    const int local = static_cast<int>(tile / CT), ct = static_cast<int>(tile % CT), cell = first + local;
    const double *F = factors + static_cast<long long>(local) * M * M;
    const int *perm = rowsrc + static_cast<long long>(local) * M;
    const double *Dv = dinv + static_cast<long long>(local) * NT * 512;
    const long long nodes = 2LL * g.cells * N;
    const int g4 = lane / 16, arow = tile_row(lane % 16);

    // this lane's right-hand-side column c (B columns, hybrid.cpp:41-42; column T is r_K)
    const int c = ct * 16 + lane % 16;
    int s = 0, v = 0, q = 0;
    double stab = 0;
    if (c < T) {
        s = c / (4 * N);
        v = (c / N) % 4;
        q = c % N;
        const int face = cell_face(cell, s, g);
        const long long fn = static_cast<long long>(face) * N + q;
        stab = face_weight(face, q, B, g) * fc[(s % 2 ? F_TAUL : F_TAUR) * nodes + fn];
    }
    const double *shape = s % 2 ? B->right : B->left;
    hdg_d4 acc[NT];
#pragma unroll
    for (int I = 0; I < NT; ++I) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * I + 4 * g4 + r;
            double val = 0;
            if (row < M && c <= T) {
                const int src = perm[row];
                if (c == T) {
                    val = resid[static_cast<long long>(cell) * M + src];
                } else if (src / P == v) {
                    const int node = src % P, in = node % N, jn = node / N;
                    if (s < 2 && jn == q) val = -stab * shape[in];
                    if (s >= 2 && in == q) val = -stab * shape[jn];
                }
            }
            acc[I][r] = val;
        }
    }
    double ring[RING][4];
    hdg_d4 cur = acc[0];  // pending block row J, the B operand of the diagonal product
    // forward substitution over block columns J
#pragma clang loop unroll(disable)
    for (int J = 0; J < NT; ++J) {
        // Tile L(I, J) for I > J; the loads of tiles I <= J (not needed) are
        // clamped to tile J + 1, so that every position issues the same loads
        // and the waits before the MFMAs count exactly AHEAD tiles in flight.
        // Rows past M read row M - 1: they only reach the padding rows. For
        // J = NT - 1 no tile is needed and the columns are clamped to M - 1.
        const double *cb = F + static_cast<long long>(16 * J + 4 * g4) * M;
        const int cols_left = M - 1 - 16 * J - 4 * g4;  // last in-range k-slice of this lane
        auto load_l = [&](int I, double(&op)[4]) {
            const int row = min(16 * min(max(I, J + 1), NT - 1) + arow, M - 1);
#pragma unroll
            for (int r = 0; r < 4; ++r) op[r] = cb[min(r, cols_left) * M + row];
        };
#pragma unroll
        for (int I = 0; I < AHEAD && I < NT; ++I) load_l(I, ring[I % RING]);
        const double *Li = Dv + (2 * J) * 256;
        hdg_d4 y = {0, 0, 0, 0};
#pragma unroll
        for (int r = 0; r < 4; ++r) y = __builtin_amdgcn_mfma_f64_16x16x4f64(Li[r * 64 + lane], cur[r], y, 0, 0, 0);
#pragma unroll
        for (int I = 0; I < NT; ++I) {
            if (I + AHEAD < NT) load_l(I + AHEAD, ring[(I + AHEAD) % RING]);
            if (I == J) acc[I] = y;
            if (I > J) {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    acc[I] = __builtin_amdgcn_mfma_f64_16x16x4f64(ring[I % RING][r], y[r], acc[I], 0, 0, 0);
                if (I == J + 1) cur = acc[I];
            }
        }
    }
    // backward substitution over block columns J, from the last. The padding
    // rows are cleared, so the U entries read past column M meet zeros.
#pragma unroll
    for (int r = 0; r < 4; ++r)
        if (16 * (NT - 1) + 4 * g4 + r >= M) acc[NT - 1][r] = 0.0;
    cur = acc[NT - 1];
#pragma clang loop unroll(disable)
    for (int J = NT - 1; J >= 0; --J) {
        // Tile U(I, J) for I < J, loads of tiles I >= J clamped to tile J - 1
        // (to tile 0 for J = 0) as in the forward sweep. Columns past M read
        // column M - 1; they meet the zero padding rows of X_J.
        const double *cb = F + static_cast<long long>(16 * J + 4 * g4) * M;
        const int cols_left = M - 1 - 16 * J - 4 * g4;  // last in-range k-slice of this lane
        auto load_u = [&](int I, double(&op)[4]) {
            const int row = 16 * max(min(I, J - 1), 0) + arow;
#pragma unroll
            for (int r = 0; r < 4; ++r) op[r] = cb[min(r, cols_left) * M + row];
        };
#pragma unroll
        for (int t = 0; t < AHEAD && t < NT; ++t) load_u(NT - 1 - t, ring[t % RING]);
        const double *Ui = Dv + (2 * J + 1) * 256;
        hdg_d4 x = {0, 0, 0, 0};
#pragma unroll
        for (int r = 0; r < 4; ++r) x = __builtin_amdgcn_mfma_f64_16x16x4f64(Ui[r * 64 + lane], cur[r], x, 0, 0, 0);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int I = NT - 1 - t;  // block rows from the last one up
            if (t + AHEAD < NT) load_u(I - AHEAD, ring[(t + AHEAD) % RING]);
            if (I == J) acc[I] = x;
            if (I < J) {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    acc[I] = __builtin_amdgcn_mfma_f64_16x16x4f64(ring[t % RING][r], x[r], acc[I], 0, 0, 0);
                if (I == J - 1) cur = acc[I];
            }
        }
    }
    if (c > T) return;
    double *out = c < T ? X + (static_cast<long long>(cell) * T + c) * M : Y + static_cast<long long>(cell) * M;
#pragma unroll
    for (int I = 0; I < NT; ++I) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * I + 4 * g4 + r;
            if (row < M) out[row] = acc[I][r];
        }
    }
}

// face_jacobian.cpp:88-110 + hybrid.cpp:65-70: C [X | y] per side (restriction
// of every column to the side, then the 4x4 HDG derivative per face node);
// S_K = -C X (column-major t x t), face_result = C y.
template <int N>
__global__ __launch_bounds__(BLOCK) void k_hdg_cx(const double *__restrict__ X, const double *__restrict__ Y,
                                                  const double *__restrict__ fc, double *__restrict__ S,
                                                  double *__restrict__ fres, const Basis *__restrict__ B, Geo g) {
    constexpr int P = N * N, M = 4 * P, T = 16 * N;
    const int cell = blockIdx.x;
    const long long nodes = 2LL * g.cells * N;
    // work item: (column c in 0..T, side s, q)
    for (int item = threadIdx.x; item < (T + 1) * 4 * N; item += BLOCK) {
        const int c = item / (4 * N), s = (item / N) % 4, q = item % N;
        const double *col = c < T ? X + (static_cast<long long>(cell) * T + c) * M : Y + static_cast<long long>(cell) * M;
        const double *shape = s % 2 ? B->right : B->left;
        double sample[4];
        for (int w = 0; w < 4; ++w) {
            double acc = 0;
            for (int i = 0; i < N; ++i) acc += col[w * P + face_node(s, q, i, N)] * shape[i];
            sample[w] = acc;
        }
        double h[16];
        hdg_matrix(fc, nodes, cell_face(cell, s, g), q, s % 2 != 0, B, g, h);
        for (int v = 0; v < 4; ++v) {
            double out = sample[0] * h[v * 4 + 0];
            for (int w = 1; w < 4; ++w) out += sample[w] * h[v * 4 + w];
            const int r = s * 4 * N + v * N + q;
            if (c < T) S[(static_cast<long long>(cell) * T + c) * T + r] = -out;
            else fres[static_cast<long long>(cell) * T + r] = out;
        }
    }
}

// Sum of the two element contributions of every face (cells in ascending
// order, as the reference's cell loop scatters them), plus `base`:
// out[f] = base[f] + loc[cL][(2a+1) part] + loc[cR][(2a) part].
__global__ void k_hdg_face_sum(const double *__restrict__ base, double base_scale, const double *__restrict__ diag_x,
                               const double *__restrict__ loc, double *__restrict__ out, int n, Geo g) {
    const int fw = 4 * n, T = 16 * n;
    const long long total = 2LL * g.cells * fw;
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    const int face = static_cast<int>(gid / fw), k = static_cast<int>(gid % fw);
    const int axis = face >= g.cells, cl = face - axis * g.cells;
    const int cr = neighbor(cl, 2 * axis + 1, g.nx, g.ny);
    double acc = base ? base_scale * base[gid] : 0.0;
    if (diag_x) acc = diag_x[gid];
    const double vl = loc[static_cast<long long>(cl) * T + (2 * axis + 1) * fw + k];
    const double vr = loc[static_cast<long long>(cr) * T + (2 * axis) * fw + k];
    if (cl < cr) { acc += vl; acc += vr; }
    else { acc += vr; acc += vl; }
    out[gid] = acc;
}

// Diagonal of the skeleton block, D_f = -w (a + b) (face_jacobian.cpp:34), times x.
__global__ void k_hdg_diag_x(const double *__restrict__ fc, const double *__restrict__ x, double *__restrict__ out,
                             const Basis *__restrict__ B, Geo g) {
    const int n = g.n, fw = 4 * n;
    const long long nodes = 2LL * g.cells * n, total = 2LL * g.cells * fw;
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    const int face = static_cast<int>(gid / fw), q = static_cast<int>(gid % n);
    const long long fn = static_cast<long long>(face) * n + q;
    const double d = -face_weight(face, q, B, g) * (fc[F_TAUL * nodes + fn] + fc[F_TAUR * nodes + fn]);
    out[gid] = d * x[gid];
}

// hybrid.cpp:73-82: per element S_K gather(x) (one wavefront per element,
// S_K columns streamed once, x gathered into registers and broadcast).
template <int N>
__global__ __launch_bounds__(BLOCK) void k_hdg_schur(const double *__restrict__ S, const double *__restrict__ x,
                                                     double *__restrict__ loc, Geo g) {
    constexpr int T = 16 * N, NA = (T + 63) / 64, FW = 4 * N;
    const int lane = threadIdx.x % 64;
    const long long cell = static_cast<long long>(blockIdx.x) * (BLOCK / 64) + wave_index();
    if (cell >= g.cells) return;
    double xg[NA], acc[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const int r = a * 64 + lane;
        acc[a] = 0;
        xg[a] = 0;
        if (r < T) {
            const int face = cell_face(static_cast<int>(cell), r / FW, g);
            xg[a] = x[static_cast<long long>(face) * FW + r % FW];
        }
    }
    const double *Sk = S + cell * T * T;
#pragma unroll
    for (int ca = 0; ca < NA; ++ca) {
        const int cend = min(64, T - ca * 64);
        for (int cl = 0; cl < cend; ++cl) {
            const int c = ca * 64 + cl;
            const double xc = readlane_d(xg[ca], cl);
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                const int r = a * 64 + lane;
                if (r < T) acc[a] += __builtin_nontemporal_load(Sk + static_cast<long long>(c) * T + r) * xc;
            }
        }
    }
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const int r = a * 64 + lane;
        if (r < T) loc[cell * T + r] = acc[a];
    }
}

// hybrid.cpp:83-96 for row == col: face block D_f + S_cL[(2a+1), (2a+1)] + S_cR[(2a), (2a)].
template <int N>
__global__ __launch_bounds__(BLOCK) void k_hdg_face_blocks(const double *__restrict__ S,
                                                           const double *__restrict__ fc,
                                                           double *__restrict__ blocks,
                                                           const Basis *__restrict__ B, Geo g) {
    constexpr int T = 16 * N, FW = 4 * N;
    const int face = blockIdx.x;
    const int axis = face >= g.cells, cl = face - axis * g.cells;
    const int cr = neighbor(cl, 2 * axis + 1, g.nx, g.ny);
    const long long nodes = 2LL * g.cells * N;
    const double *SL = S + static_cast<long long>(cl) * T * T, *SR = S + static_cast<long long>(cr) * T * T;
    const int ol = (2 * axis + 1) * FW, orr = (2 * axis) * FW;
    for (int e = threadIdx.x; e < FW * FW; e += BLOCK) {
        const int i = e % FW, j = e / FW;
        double val = 0;
        if (i == j) {
            const int q = i % N;
            const long long fn = static_cast<long long>(face) * N + q;
            val = -face_weight(face, q, B, g) * (fc[F_TAUL * nodes + fn] + fc[F_TAUR * nodes + fn]);
        }
        val += SL[static_cast<long long>(ol + j) * T + ol + i];
        val += SR[static_cast<long long>(orr + j) * T + orr + i];
        blocks[static_cast<long long>(face) * FW * FW + e] = val;
    }
}

// hybrid.cpp:97-105: du = -X gather(dlambda) - y, one wavefront per element.
template <int N>
__global__ __launch_bounds__(BLOCK) void k_hdg_recover(const double *__restrict__ X, const double *__restrict__ Y,
                                                       const double *__restrict__ dl, double *__restrict__ du,
                                                       Geo g) {
    constexpr int P = N * N, M = 4 * P, T = 16 * N, FW = 4 * N, NA = (M + 63) / 64, NT = (T + 63) / 64;
    const int lane = threadIdx.x % 64;
    const long long cell = static_cast<long long>(blockIdx.x) * (BLOCK / 64) + wave_index();
    if (cell >= g.cells) return;
    double xg[NT], acc[NA];
#pragma unroll
    for (int a = 0; a < NT; ++a) {
        const int r = a * 64 + lane;
        xg[a] = 0;
        if (r < T) xg[a] = dl[static_cast<long long>(cell_face(static_cast<int>(cell), r / FW, g)) * FW + r % FW];
    }
#pragma unroll
    for (int a = 0; a < NA; ++a) acc[a] = 0;
    const double *Xk = X + cell * T * M;
#pragma unroll
    for (int ca = 0; ca < NT; ++ca) {
        const int cend = min(64, T - ca * 64);
        for (int cl = 0; cl < cend; ++cl) {
            const int c = ca * 64 + cl;
            const double xc = readlane_d(xg[ca], cl);
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                const int r = a * 64 + lane;
                if (r < M) acc[a] += Xk[static_cast<long long>(c) * M + r] * xc;
            }
        }
    }
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const int r = a * 64 + lane;
        if (r < M) du[cell * M + r] = -acc[a] - Y[cell * M + r];
    }
}

// hybrid.cpp:106-121: lift of w tau e / D_f on every side of every element.
template <int N>
__global__ __launch_bounds__(BLOCK) void k_hdg_recovered_residual(const double *__restrict__ e,
                                                                  const double *__restrict__ fc,
                                                                  double *__restrict__ out,
                                                                  const Basis *__restrict__ B, Geo g) {
    constexpr int P = N * N, M = 4 * P, FW = 4 * N;
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= static_cast<long long>(g.cells) * P) return;
    const int cell = static_cast<int>(gid / P), node = static_cast<int>(gid % P), i = node % N, j = node / N;
    const long long nodes = 2LL * g.cells * N;
    auto phi = [&](int side, int q, int v) {
        const int face = cell_face(cell, side, g);
        const long long fn = static_cast<long long>(face) * N + q;
        const double a = fc[F_TAUL * nodes + fn], b = fc[F_TAUR * nodes + fn];
        const double w = face_weight(face, q, B, g);
        const double d = -w * (a + b);
        const double tau = side % 2 ? a : b;
        return w * tau * e[static_cast<long long>(face) * FW + v * N + q] / d;
    };
    for (int v = 0; v < 4; ++v) {
        double r = 0;
        r += (1 * B->left[i]) * phi(0, j, v);
        r += (1 * B->right[i]) * phi(1, j, v);
        r += (1 * phi(2, i, v)) * B->left[j];
        r += (1 * phi(3, i, v)) * B->right[j];
        out[static_cast<long long>(cell) * M + v * P + node] = r;
    }
}

// 1 / face_weight per trace entry (the HDG Krylov metric, time_integrator.cpp:27).
__global__ void k_inverse_trace_mass(double *__restrict__ out, const Basis *__restrict__ B, Geo g) {
    const int n = g.n, fw = 4 * n;
    const long long total = 2LL * g.cells * fw;
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    const int face = static_cast<int>(gid / fw), q = static_cast<int>(gid % n);
    out[gid] = 1.0 / face_weight(face, q, B, g);
}

unsigned grid_for(long long threads, int block) { return static_cast<unsigned>((threads + block - 1) / block); }

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

namespace {
hipError_t launch_hdg_solve_mfma(int n, const double *factors, const int *rowsrc, double *dinv,
                                 const double *fcache, const double *resid, double *X, double *Y, int first,
                                 int count, const Basis *b, const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, {
        constexpr int M = 4 * N * N, NT = (M + 15) / 16, CT = (16 * N + 1 + 15) / 16;
        k_diag_inverses<N><<<grid_for(32LL * count * NT, BLOCK), BLOCK, 0, s>>>(factors, dinv, count);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        k_hdg_solve_mfma<N><<<grid_for(static_cast<long long>(count) * CT, MFMA_SOLVE_WAVES), MFMA_SOLVE_WAVES * 64,
                              0, s>>>(factors, rowsrc, dinv, fcache, resid, X, Y, first, count, b, g);
    })
}
}  // namespace
size_t hdg_solve_workspace(int n, int count) {
    const int M = 4 * n * n, NT = (M + 15) / 16;
    return static_cast<size_t>(count) * NT * 512;
}
hipError_t launch_hdg_solve(int n, const double *factors, const int *rowsrc, const double *fcache,
                            const double *resid, double *X, double *Y, int first, int count, const Basis *b,
                            const Geo &g, hipStream_t s, double *workspace) {
    if (workspace)
        return launch_hdg_solve_mfma(n, factors, rowsrc, workspace, fcache, resid, X, Y, first, count, b, g, s);
    DGH_DISPATCH_N(n, {
        const int cols = 16 * N + 1;
        for (int col0 = 0; col0 < cols; col0 += SOLVE_WAVES * SOLVE_COLS) {
            k_hdg_solve<N><<<count, SOLVE_WAVES * 64, 0, s>>>(factors, rowsrc, fcache, resid, X, Y, first, col0, b, g);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
    })
}
hipError_t launch_hdg_cx(int n, const double *X, const double *Y, const double *fcache, double *S, double *fres,
                         const Basis *b, const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, { k_hdg_cx<N><<<g.cells, BLOCK, 0, s>>>(X, Y, fcache, S, fres, b, g); })
}
hipError_t launch_hdg_face_sum(int n, const double *base, double base_scale, const double *diag_x, const double *loc,
                               double *out, const Geo &g, hipStream_t s) {
    k_hdg_face_sum<<<grid_for(2LL * g.cells * 4 * n, BLOCK), BLOCK, 0, s>>>(base, base_scale, diag_x, loc, out, n, g);
    return hipGetLastError();
}
hipError_t launch_hdg_diag_x(int n, const double *fcache, const double *x, double *out, const Basis *b, const Geo &g,
                             hipStream_t s) {
    k_hdg_diag_x<<<grid_for(2LL * g.cells * 4 * n, BLOCK), BLOCK, 0, s>>>(fcache, x, out, b, g);
    return hipGetLastError();
}
hipError_t launch_hdg_schur(int n, const double *S, const double *x, double *loc, const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, { k_hdg_schur<N><<<grid_for(g.cells, BLOCK / 64), BLOCK, 0, s>>>(S, x, loc, g); })
}
hipError_t launch_hdg_face_blocks(int n, const double *S, const double *fcache, double *blocks, const Basis *b,
                                  const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, { k_hdg_face_blocks<N><<<2 * g.cells, BLOCK, 0, s>>>(S, fcache, blocks, b, g); })
}
hipError_t launch_hdg_recover(int n, const double *X, const double *Y, const double *dl, double *du, const Geo &g,
                              hipStream_t s) {
    DGH_DISPATCH_N(n, { k_hdg_recover<N><<<grid_for(g.cells, BLOCK / 64), BLOCK, 0, s>>>(X, Y, dl, du, g); })
}
hipError_t launch_hdg_recovered_residual(int n, const double *e, const double *fcache, double *out, const Basis *b,
                                         const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, {
        k_hdg_recovered_residual<N><<<grid_for(static_cast<long long>(g.cells) * N * N, BLOCK), BLOCK, 0, s>>>(
            e, fcache, out, b, g);
    })
}
hipError_t launch_inverse_trace_mass(int n, double *out, const Basis *b, const Geo &g, hipStream_t s) {
    k_inverse_trace_mass<<<grid_for(2LL * g.cells * 4 * n, BLOCK), BLOCK, 0, s>>>(out, b, g);
    return hipGetLastError();
}

}  // namespace dgh
