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
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
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
    const long long cell = static_cast<long long>(blockIdx.x) * (BLOCK / 64) + threadIdx.x / 64;
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
    const long long cell = static_cast<long long>(blockIdx.x) * (BLOCK / 64) + threadIdx.x / 64;
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

hipError_t launch_hdg_solve(int n, const double *factors, const int *rowsrc, const double *fcache,
                            const double *resid, double *X, double *Y, int first, int count, const Basis *b,
                            const Geo &g, hipStream_t s) {
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
