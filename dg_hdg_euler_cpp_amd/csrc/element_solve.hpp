// Device helpers shared by the kernels that solve with a factorised element
// block against many right-hand sides at once (HDG condensation, block
// ILU(0) couplings): mesh topology, cross-lane broadcast and the column-
// streamed forward/backward substitution.
#pragma once
#include "dgh_internal.hpp"

namespace dgh {
namespace {

__device__ __forceinline__ int neighbor(int cell, int side, int nx, int ny) {
    int x = cell % nx, y = cell / nx;
    if (side < 2) x = (x + (side == 0 ? nx - 1 : 1)) % nx;
    else y = (y + (side == 2 ? ny - 1 : 1)) % ny;
    return x + nx * y;
}
// face on side s of a cell: its own +x / +y face for s = 1 / 3, the -x / -y
// neighbour's for s = 0 / 2 (discretization.cpp:32-34)
__device__ __forceinline__ int cell_face(int cell, int side, const Geo &g) {
    return side == 0 ? neighbor(cell, 0, g.nx, g.ny)
         : side == 1 ? cell
         : side == 2 ? g.cells + neighbor(cell, 2, g.nx, g.ny)
                     : g.cells + cell;
}
__device__ __forceinline__ int face_node(int side, int q, int i, int n) { return side < 2 ? i + n * q : q + n * i; }
__device__ __forceinline__ double readlane_d(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float readlane_d(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double face_weight(int face, int q, const Basis *__restrict__ B, const Geo &g) {
    return .5 * (face < g.cells ? g.hy : g.hx) * B->w[q];
}

// x <- U^-1 L^-1 x for COLS right-hand sides held in registers (lane l owns
// rows l + 64a; the rows are already permuted by the pivots), with the M x M
// column-major LU factors F of one element block (dense_lu.hpp:33-37), in the
// factors' precision T.
template <int M, int COLS, int RING, typename T>
__device__ __forceinline__ void lu_solve_columns(const T *__restrict__ F, T (&x)[COLS][(M + 63) / 64], int lane) {
    constexpr int NA = (M + 63) / 64;
    // The factor columns stream through a ring of RING buffers so the
    // loads of the next columns are in flight while a column is applied; they
    // are read once per pass, hence non-temporal.
    T buf[RING][NA];
    auto load_lower = [&](T(&b)[NA], int c, int a0) {  // rows > c, chunks >= a0
        const T *colp = F + static_cast<long long>(c) * M;
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            const int row = a * 64 + lane;
            if (a >= a0) b[a] = (row > c && row < M) ? __builtin_nontemporal_load(colp + row) : T(0);
        }
    };
    auto load_upper = [&](T(&b)[NA], int c, int a1) {  // rows <= c, chunks <= a1
        const T *colp = F + static_cast<long long>(c) * M;
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            const int row = a * 64 + lane;
            if (a <= a1) b[a] = row <= c ? __builtin_nontemporal_load(colp + row) : T(0);
        }
    };
    // forward substitution with the unit lower L (column j in ring slot j % RING)
#pragma unroll
    for (int u = 0; u < RING; ++u)
        if (u < M) load_lower(buf[u], u, 0);
#pragma unroll
    for (int ja = 0; ja < NA; ++ja) {
        const int jend = min(64, M - ja * 64);
#pragma clang loop unroll(disable)
        for (int jl = 0; jl < jend; jl += RING) {
#pragma unroll
            for (int u = 0; u < RING; ++u) {
                if (jl + u < jend) {
                    const int j = ja * 64 + jl + u;
#pragma unroll
                    for (int cc = 0; cc < COLS; ++cc) {
                        const T xj = readlane_d(x[cc][ja], jl + u);
#pragma unroll
                        for (int a = ja; a < NA; ++a) x[cc][a] -= buf[u][a] * xj;
                    }
                    if (j + RING < M) load_lower(buf[u], j + RING, ja);
                }
            }
        }
    }
    // backward substitution with U (column j in ring slot (M - 1 - j) % RING)
#pragma unroll
    for (int u = 0; u < RING; ++u)
        if (M - 1 - u >= 0) load_upper(buf[u], M - 1 - u, NA - 1);
#pragma unroll
    for (int ja = NA - 1; ja >= 0; --ja) {
        const int jend = min(64, M - ja * 64);
        const int t0 = M - ja * 64 - jend;
#pragma clang loop unroll(disable)
        for (int g0 = 0; g0 < jend; g0 += RING) {
#pragma unroll
            for (int u = 0; u < RING; ++u) {
                const int jl = jend - 1 - g0 - u;
                if (jl >= 0) {
                    const int j = ja * 64 + jl;
                    T(&cur)[NA] = buf[(t0 + u) % RING];
                    const T diag = readlane_d(cur[ja], jl);
#pragma unroll
                    for (int cc = 0; cc < COLS; ++cc) {
                        const T xj = readlane_d(x[cc][ja], jl) / diag;
                        if (lane == jl) x[cc][ja] = xj;
#pragma unroll
                        for (int a = 0; a <= ja; ++a) {
                            const int row = a * 64 + lane;
                            if (row < j) x[cc][a] -= cur[a] * xj;
                        }
                    }
                    if (j - RING >= 0) load_upper(cur, j - RING, ja);
                }
            }
        }
    }
}

}  // namespace
}  // namespace dgh
