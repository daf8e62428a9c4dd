// 1D periodic Euler DG on the device: the reference's dg_hdg_1d solver
// (discretization.cpp, volume_jacobian.cpp, dg_operator.cpp,
// time_integrator.cpp) with GMRES(restart) + block Jacobi, and the extern "C"
// entry points of include/dgh/dgh1d_device.h.
//
// The problem is small (cells * 3(k+1) unknowns, 1152 at the reference's
// 128-cell k=2 protocol), so launches, not bytes, bound it. The element
// kernels give one thread a whole cell, and GMRES runs as one 512-thread
// workgroup that keeps the Hessenberg matrix, the Givens rotations and every
// reduction on chip and applies the block Jacobi and the Jacobian itself:
// one launch per linear solve instead of about ten per Krylov iteration.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dgh/dgh1d_device.h"
#include "dgh_host.hpp"
#include "dgh_internal.hpp"

namespace dgh {
namespace {

constexpr int NV1 = 3;            // density, momentum, energy (dg_hdg_1d types.hpp:8)
constexpr int CELL_BLOCK = 64;    // threads per block of the per-cell kernels
constexpr int SOLVE_THREADS = 512;  // 8 waves: 256 registers per lane for the inlined cell work
constexpr int SOLVE_WAVES = SOLVE_THREADS / 64;
constexpr int DOT_CHUNKS = SOLVE_THREADS / 32;  // threads per basis vector in the CGS2 dot products
constexpr int MAX_RESTART = 31;

// Reference-element tables of the 1D basis (nodal_basis.hpp:34-65): the
// volume rule is the nodal one (Q = n) or the Q-point over-integration rule.
struct Line {
    int n, Q, eq, cells;
    double h, gamma, safety;
    double amplitude, entropy;  // initial Riemann invariant amplitude, p = entropy rho^gamma
    double wd[MAXQ * MAXN];  // weighted derivative WD(q, i) at [q + Q*i]
    double vq[MAXQ * MAXN];  // basis values V(q, k) at [q + Q*k] (identity when Q = n)
    double w[MAXN], left[MAXN], right[MAXN], mass[MAXN];
    double ev[MAXEQ * MAXN], ex[MAXEQ], ew[MAXEQ], proj[MAXN * MAXEQ];
};

// ---------------------------------------------------------------- physics
// dg_hdg_1d physics.cpp:4-39
__device__ __forceinline__ double pressure1(const double u[3], double g) {
    return (g - 1) * (u[2] - .5 * u[1] * u[1] / u[0]);
}
__device__ __forceinline__ void flux1(const double u[3], double g, double f[3]) {
    const double v = u[1] / u[0], p = pressure1(u, g);
    f[0] = u[1];
    f[1] = u[1] * v + p;
    f[2] = (u[2] + p) * v;
}
// column-major 3x3, J(v, w) at [v + 3w]
__device__ __forceinline__ void jacobian1(const double u[3], double g, double a[9]) {
    const double v = u[1] / u[0], H = (u[2] + pressure1(u, g)) / u[0];
    a[0] = 0, a[3] = 1, a[6] = 0;
    a[1] = .5 * (g - 3) * v * v, a[4] = (3 - g) * v, a[7] = g - 1;
    a[2] = v * (.5 * (g - 1) * v * v - H), a[5] = H - (g - 1) * v * v, a[8] = g * v;
}
__device__ __forceinline__ double spectral1(const double u[3], double g) {
    return fabs(u[1] / u[0]) + sqrt(g * pressure1(u, g) / u[0]);
}
__device__ __forceinline__ bool physical1(const double z[3], double g, double floor) {
    return z[0] > floor && pressure1(z, g) > floor;
}
// Einfeldt HLL speeds with Roe averages; false for an invalid face state
__device__ bool einfeldt1(const double l[3], const double r[3], double g, double safety, double &a, double &b) {
    if (!(l[0] > 0 && r[0] > 0 && pressure1(l, g) > 0 && pressure1(r, g) > 0)) return false;
    const double sl = sqrt(l[0]), sr = sqrt(r[0]);
    const double ul = l[1] / l[0], ur = r[1] / r[0];
    const double cl = sqrt(g * pressure1(l, g) / l[0]), cr = sqrt(g * pressure1(r, g) / r[0]);
    const double v = (sl * ul + sr * ur) / (sl + sr);
    const double H = (sl * (l[2] + pressure1(l, g)) / l[0] + sr * (r[2] + pressure1(r, g)) / r[0]) / (sl + sr);
    const double c = sqrt((g - 1) * (H - .5 * v * v));
    a = safety * fmax(fmax(0.0, -ul + cl), -v + c);
    b = safety * fmax(fmax(0.0, ur + cr), v + c);
    return true;
}
// simple_wave.cpp:3-25: the simple wave through its characteristic foot, with
// J(xi) = A sin(pi xi) the invariant u - 2c/(gamma-1) at t = 0 and
// p = K rho^gamma: the reference's wave is A = 1, K = 1; the `density` case
// (no reference counterpart) A = -0.4, K = 1/gamma, i.e. rho = 1 + 0.2 sin(pi x)
// and u = 2 - rho at t = 0 for gamma = 3.
__device__ void simple_wave(double x, double t, double g, double A, double K, double out[3]) {
    const double pi = 3.14159265358979323846, a = (3 - g) / 2, b = (g + 1) / 4, bs = b * fabs(A);
    double lo = x - t * (a + bs), hi = x - t * (a - bs), xi = x - t * a;
    for (int i = 0; i < 80; ++i) {
        const double f = xi + t * (a + b * (A * sin(pi * xi))) - x;
        if (fabs(f) < 2e-15) break;
        if (f > 0) hi = xi;
        else lo = xi;
        const double z = xi - f / (1 + t * (b * A) * pi * cos(pi * xi));
        xi = z > lo && z < hi ? z : .5 * (lo + hi);
    }
    const double jm = A * sin(pi * xi), v = .5 * (2 + jm), c = (g - 1) * (2 - jm) / 4;
    const double rho = pow(c * c / (g * K), 1 / (g - 1)), p = K * pow(rho, g);
    out[0] = rho;
    out[1] = rho * v;
    out[2] = p / (g - 1) + .5 * rho * v * v;
}

// ------------------------------------------------------------ cell access
template <int N, typename T>
__device__ __forceinline__ void sample1(const T *u, int cell, const double *shape, double s[3]) {
    const T *c = u + static_cast<long long>(cell) * 3 * N;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        double acc = 0;
#pragma unroll
        for (int i = 0; i < N; ++i) acc += shape[i] * static_cast<double>(c[v * N + i]);
        s[v] = acc;
    }
}
// state at volume point q (discretization.cpp:22-30)
template <int N, typename T>
__device__ __forceinline__ void point1(const T *u, int cell, int q, const Line &L, double s[3]) {
    const T *c = u + static_cast<long long>(cell) * 3 * N;
    if (L.Q == N) {
#pragma unroll
        for (int v = 0; v < 3; ++v) s[v] = static_cast<double>(c[v * N + q]);
        return;
    }
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        double acc = 0;
#pragma unroll
        for (int i = 0; i < N; ++i) acc += L.vq[q + L.Q * i] * static_cast<double>(c[v * N + i]);
        s[v] = acc;
    }
}
__device__ __forceinline__ int prev_cell(int e, int cells) { return e == 0 ? cells - 1 : e - 1; }
__device__ __forceinline__ int next_cell(int e, int cells) { return e + 1 == cells ? 0 : e + 1; }

// ---------------------------------------------------------------- kernels
// discretization.cpp:31-73: L2 projection of the simple wave on the 20-point
// rule, then the convex scaling toward the cell mean that keeps every node,
// volume point, error point and face value physical.
template <int N>
__global__ __launch_bounds__(CELL_BLOCK) void k1_initial(double *u, int *flag, const Line *__restrict__ LP) {
    const Line &L = *LP;
    const int e = blockIdx.x * CELL_BLOCK + threadIdx.x;
    if (e >= L.cells) return;
    const double g = L.gamma;
    double nodal[N][3];
#pragma unroll
    for (int i = 0; i < N; ++i) nodal[i][0] = nodal[i][1] = nodal[i][2] = 0;
    for (int q = 0; q < L.eq; ++q) {
        double s[3];
        simple_wave(-1 + L.h * (e + .5) + .5 * L.h * L.ex[q], 0.0, g, L.amplitude, L.entropy, s);
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
            for (int v = 0; v < 3; ++v) nodal[i][v] += L.proj[i + N * q] * s[v];
    }
    double mean[3];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        double acc = 0;
#pragma unroll
        for (int i = 0; i < N; ++i) acc += nodal[i][v] * L.w[i];
        mean[v] = .5 * acc;
    }
    if (!(mean[0] > 0 && pressure1(mean, g) > 0)) atomicOr(flag, 1);
    double theta = 1;
    auto check = [&](const double z[3]) {
        double d[3];
        for (int v = 0; v < 3; ++v) d[v] = z[v] - mean[v];
        auto ok = [&](double f) {
            double y[3];
            for (int v = 0; v < 3; ++v) y[v] = mean[v] + f * d[v];
            return physical1(y, g, 1e-12);
        };
        if (!ok(theta)) {
            double low = 0, high = theta;
            for (int k = 0; k < 60; ++k) {
                const double mid = .5 * (low + high);
                if (ok(mid)) low = mid;
                else high = mid;
            }
            theta = .99 * low;
        }
    };
    for (int i = 0; i < N; ++i) check(nodal[i]);
    auto mapped = [&](const double *map, int rows) {
        for (int q = 0; q < rows; ++q) {
            double z[3] = {0, 0, 0};
            for (int v = 0; v < 3; ++v)
                for (int i = 0; i < N; ++i) z[v] += map[q + rows * i] * nodal[i][v];
            check(z);
        }
    };
    mapped(L.vq, L.Q);
    mapped(L.ev, L.eq);
    double l[3] = {0, 0, 0}, r[3] = {0, 0, 0};
    for (int v = 0; v < 3; ++v)
        for (int i = 0; i < N; ++i) {
            l[v] += L.left[i] * nodal[i][v];
            r[v] += L.right[i] * nodal[i][v];
        }
    check(l);
    check(r);
    double *c = u + static_cast<long long>(e) * 3 * N;
#pragma unroll
    for (int v = 0; v < 3; ++v)
#pragma unroll
        for (int i = 0; i < N; ++i) c[v * N + i] = mean[v] + theta * (nodal[i][v] - mean[v]);
}

// discretization.cpp:75-92, first loop: Einfeldt speeds of face f (between
// cell f and f + 1)
template <int N>
__global__ __launch_bounds__(CELL_BLOCK) void k1_speeds(const double *u, double *speeds, int *flag,
                                                        const Line *__restrict__ LP) {
    const Line &L = *LP;
    const int f = blockIdx.x * CELL_BLOCK + threadIdx.x;
    if (f >= L.cells) return;
    double l[3], r[3], a = 0, b = 0;
    sample1<N>(u, f, L.right, l);
    sample1<N>(u, next_cell(f, L.cells), L.left, r);
    if (!einfeldt1(l, r, L.gamma, L.safety, a, b)) atomicOr(flag, 2);
    speeds[2 * f] = a;
    speeds[2 * f + 1] = b;
}
// second loop: rate = max_e alpha_e / h; positive doubles order like their bits
template <int N>
__global__ __launch_bounds__(CELL_BLOCK) void k1_rate(const double *u, const double *speeds,
                                                      unsigned long long *rate_bits, int *flag,
                                                      const Line *__restrict__ LP) {
    const Line &L = *LP;
    const int e = blockIdx.x * CELL_BLOCK + threadIdx.x;
    if (e >= L.cells) return;
    const double *sl = speeds + 2 * prev_cell(e, L.cells), *sr = speeds + 2 * e;
    double alpha = fmax(fmax(sl[0], sl[1]), fmax(sr[0], sr[1]));
    for (int q = 0; q < L.Q; ++q) {
        double z[3];
        point1<N>(u, e, q, L, z);
        alpha = fmax(alpha, spectral1(z, L.gamma));
    }
    const double rate = alpha / L.h;
    if (!(rate >= 0) || !isfinite(rate)) atomicOr(flag, 4);
    else atomicMax(rate_bits, static_cast<unsigned long long>(__double_as_longlong(rate)));
}

// HLL flux of face f from its two traces (discretization.cpp:118-121)
__device__ __forceinline__ void hll1(const double l[3], const double r[3], const double *s, double g, double f[3]) {
    double fl[3], fr[3];
    flux1(l, g, fl);
    flux1(r, g, fr);
    for (int v = 0; v < 3; ++v) f[v] = (s[1] * fl[v] + s[0] * fr[v] + s[0] * s[1] * (l[v] - r[v])) / (s[0] + s[1]);
}
// Problem::dg_residual (discretization.cpp:102-127) of one cell: mass term,
// volume term and the fluxes of its two faces (each face flux is computed by
// both of its cells with the same operations)
template <int N>
__global__ __launch_bounds__(CELL_BLOCK) void k1_residual(const double *u, const double *old, const double *speeds,
                                                          double *r, double tau, const Line *__restrict__ LP) {
    const Line &L = *LP;
    const int e = blockIdx.x * CELL_BLOCK + threadIdx.x;
    if (e >= L.cells) return;
    const long long base = static_cast<long long>(e) * 3 * N;
    double res[3][N];
#pragma unroll
    for (int v = 0; v < 3; ++v)
#pragma unroll
        for (int i = 0; i < N; ++i)
            res[v][i] = L.mass[i] * (u[base + v * N + i] - old[base + v * N + i]) / tau;
    if (N > 1) {
        double acc[3][N];
#pragma unroll
        for (int v = 0; v < 3; ++v)
#pragma unroll
            for (int i = 0; i < N; ++i) acc[v][i] = 0;
        for (int q = 0; q < L.Q; ++q) {
            double z[3], f[3];
            point1<N>(u, e, q, L, z);
            flux1(z, L.gamma, f);
#pragma unroll
            for (int v = 0; v < 3; ++v)
#pragma unroll
                for (int i = 0; i < N; ++i) acc[v][i] += L.wd[q + L.Q * i] * f[v];
        }
#pragma unroll
        for (int v = 0; v < 3; ++v)
#pragma unroll
            for (int i = 0; i < N; ++i) res[v][i] -= acc[v][i];
    }
    const int p = prev_cell(e, L.cells), nx = next_cell(e, L.cells);
    double tl[3], tr[3], fm[3], fp[3];
    sample1<N>(u, p, L.right, tl);
    sample1<N>(u, e, L.left, tr);
    hll1(tl, tr, speeds + 2 * p, L.gamma, fm);
    sample1<N>(u, e, L.right, tl);
    sample1<N>(u, nx, L.left, tr);
    hll1(tl, tr, speeds + 2 * e, L.gamma, fp);
#pragma unroll
    for (int v = 0; v < 3; ++v)
#pragma unroll
        for (int i = 0; i < N; ++i) {
            double acc = res[v][i];
            acc -= L.left[i] * fm[v];
            acc += L.right[i] * fp[v];
            r[base + v * N + i] = acc;
        }
}

// discretization.cpp:128-144: density and pressure above 1e-13 at the nodes,
// the volume points and both face values, all values finite
template <int N>
__global__ __launch_bounds__(CELL_BLOCK) void k1_admissible(const double *u, int *flag, const Line *__restrict__ LP) {
    const Line &L = *LP;
    const int e = blockIdx.x * CELL_BLOCK + threadIdx.x;
    if (e >= L.cells) return;
    const double *c = u + static_cast<long long>(e) * 3 * N;
    bool ok = true;
    for (int k = 0; k < 3 * N; ++k) ok = ok && isfinite(c[k]);
    for (int i = 0; i < N && ok; ++i) {
        const double z[3] = {c[i], c[N + i], c[2 * N + i]};
        ok = physical1(z, L.gamma, 1e-13);
    }
    for (int q = 0; q < L.Q && ok; ++q) {
        double z[3];
        point1<N>(u, e, q, L, z);
        ok = physical1(z, L.gamma, 1e-13);
    }
    if (ok) {
        double l[3], r[3];
        sample1<N>(u, e, L.left, l);
        sample1<N>(u, e, L.right, r);
        ok = physical1(l, L.gamma, 1e-13) && physical1(r, L.gamma, 1e-13);
    }
    if (!ok) atomicOr(flag, 1);
}

// DGOperator::refresh (dg_operator.cpp:7-23 + volume_jacobian.cpp:9-37):
// F'(U) at the volume points of cell e and the two HLL derivatives of face e,
// left = b/(a+b) (F'(U_L) + a I), right = a/(a+b) (F'(U_R) - b I)
template <int N>
__global__ __launch_bounds__(CELL_BLOCK) void k1_refresh(const double *u, const double *speeds, double *coef,
                                                         double *fl, double *fr, const Line *__restrict__ LP) {
    const Line &L = *LP;
    const int e = blockIdx.x * CELL_BLOCK + threadIdx.x;
    if (e >= L.cells) return;
    if (N > 1)
        for (int q = 0; q < L.Q; ++q) {
            double z[3], J[9];
            point1<N>(u, e, q, L, z);
            jacobian1(z, L.gamma, J);
            for (int k = 0; k < 9; ++k) coef[(static_cast<long long>(e) * L.Q + q) * 9 + k] = J[k];
        }
    const double a = speeds[2 * e], b = speeds[2 * e + 1];
    double l[3], r[3], JL[9], JR[9];
    sample1<N>(u, e, L.right, l);
    sample1<N>(u, next_cell(e, L.cells), L.left, r);
    jacobian1(l, L.gamma, JL);
    jacobian1(r, L.gamma, JR);
    for (int k = 0; k < 9; ++k) {
        const double id = k % 4 == 0 ? 1.0 : 0.0;
        fl[9 * e + k] = b / (a + b) * (JL[k] + a * id);
        fr[9 * e + k] = a / (a + b) * (JR[k] - b * id);
    }
}

// DGOperator::apply (dg_operator.cpp:24-49 + volume_jacobian.cpp:38-104) for
// cell e: y = M x / tau - WD^T (F'(U) x at the points) + the face lifts
template <int N, typename TX>
__device__ __forceinline__ void apply_cell(int e, const TX *x, double *y, const double *coef, const double *fl,
                                           const double *fr, double inv_step, const Line &L) {
    const long long base = static_cast<long long>(e) * 3 * N;
    double out[3][N];
#pragma unroll
    for (int v = 0; v < 3; ++v)
#pragma unroll
        for (int i = 0; i < N; ++i) out[v][i] = L.mass[i] * static_cast<double>(x[base + v * N + i]) * inv_step;
    if (N > 1) {
        double acc[3][N];
#pragma unroll
        for (int v = 0; v < 3; ++v)
#pragma unroll
            for (int i = 0; i < N; ++i) acc[v][i] = 0;
        for (int q = 0; q < L.Q; ++q) {
            double d[3], fd[3];
            point1<N>(x, e, q, L, d);
            const double *J = coef + (static_cast<long long>(e) * L.Q + q) * 9;
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                double s = 0;
#pragma unroll
                for (int w = 0; w < 3; ++w) s += J[v + 3 * w] * d[w];
                fd[v] = s;
            }
#pragma unroll
            for (int v = 0; v < 3; ++v)
#pragma unroll
                for (int i = 0; i < N; ++i) acc[v][i] += L.wd[q + L.Q * i] * fd[v];
        }
#pragma unroll
        for (int v = 0; v < 3; ++v)
#pragma unroll
            for (int i = 0; i < N; ++i) out[v][i] -= acc[v][i];
    }
    const int p = prev_cell(e, L.cells), nx = next_cell(e, L.cells);
    double flux[2][3];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const int face = side == 0 ? p : e;
        double l[3], r[3];
        sample1<N>(x, face, L.right, l);
        sample1<N>(x, side == 0 ? e : nx, L.left, r);
        const double *A = fl + 9 * face, *B = fr + 9 * face;
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            double a = 0, b = 0;
#pragma unroll
            for (int w = 0; w < 3; ++w) {
                a += A[v + 3 * w] * l[w];
                b += B[v + 3 * w] * r[w];
            }
            flux[side][v] = a + b;
        }
    }
#pragma unroll
    for (int v = 0; v < 3; ++v)
#pragma unroll
        for (int i = 0; i < N; ++i) {
            double acc = out[v][i];
            acc -= L.left[i] * flux[0][v];
            acc += L.right[i] * flux[1][v];
            y[base + v * N + i] = acc;
        }
}
template <int N>
__global__ __launch_bounds__(CELL_BLOCK) void k1_apply(const double *x, double *y, const double *coef,
                                                       const double *fl, const double *fr, double inv_step,
                                                       const Line *__restrict__ LP) {
    const int e = blockIdx.x * CELL_BLOCK + threadIdx.x;
    if (e < LP->cells) apply_cell<N>(e, x, y, coef, fl, fr, inv_step, *LP);
}

// permuted unit-lower + upper solve with one cell's packed-in-place LU
template <int M>
__device__ __forceinline__ void lu_solve_cell(const double *A, const int *perm, const double *rhs, double *out) {
    double y[M];
#pragma unroll
    for (int i = 0; i < M; ++i) y[i] = rhs[perm[i]];
#pragma unroll
    for (int j = 0; j < M; ++j) {
        const double yj = y[j];
#pragma unroll
        for (int i = j + 1; i < M; ++i) y[i] -= A[i + M * j] * yj;
    }
#pragma unroll
    for (int j = M - 1; j >= 0; --j) {
        y[j] /= A[j + M * j];
        const double yj = y[j];
#pragma unroll
        for (int i = 0; i < j; ++i) y[i] -= A[i + M * j] * yj;
    }
#pragma unroll
    for (int i = 0; i < M; ++i) out[i] = y[i];
}
// in-place partial-pivot LU of an m x m column-major block (dense_lu.hpp:10-46);
// false when singular or not finite
__device__ bool lu_in_place(double *A, int *perm, int m) {
    auto at = [&](int i, int j) -> double & { return A[i + m * j]; };
    for (int i = 0; i < m; ++i) perm[i] = i;
    for (int k = 0; k < m; ++k) {
        int piv = k;
        double best = fabs(at(k, k));
        for (int i = k + 1; i < m; ++i)
            if (fabs(at(i, k)) > best) {
                best = fabs(at(i, k));
                piv = i;
            }
        if (piv != k) {
            for (int j = 0; j < m; ++j) {
                const double t = at(k, j);
                at(k, j) = at(piv, j);
                at(piv, j) = t;
            }
            const int t = perm[k];
            perm[k] = perm[piv];
            perm[piv] = t;
        }
        const double pivot = at(k, k);
        if (pivot != 0)
            for (int i = k + 1; i < m; ++i) at(i, k) /= pivot;
        for (int j = k + 1; j < m; ++j) {
            const double ukj = at(k, j);
            for (int i = k + 1; i < m; ++i) at(i, j) -= at(i, k) * ukj;
        }
    }
    bool ok = true;
    for (int k = 0; k < m * m; ++k) ok = ok && isfinite(A[k]);
    for (int k = 0; k < m; ++k) ok = ok && fabs(at(k, k)) > 2.2250738585072014e-308;
    return ok;
}
// DGOperator::block(row, row) (dg_operator.cpp:50-79, volume_jacobian.cpp:
// 105-132), column-major 3n x 3n with index v*n + i, then (factorize != 0) the
// partial-pivot LU of dense_lu.hpp:10-46 in place: max |a| pivot, first row on
// ties, row interchanges across the full row, unit-lower L; rowsrc holds
// (P b)[i] = b[rowsrc[i]].
template <int N>
__global__ __launch_bounds__(CELL_BLOCK) void k1_block(const double *coef, const double *fl, const double *fr,
                                                       double inv_step, double *blocks, int *rowsrc, int *flag,
                                                       int factorize, const Line *__restrict__ LP) {
    constexpr int m = 3 * N;
    const Line &L = *LP;
    const int row = blockIdx.x * CELL_BLOCK + threadIdx.x;
    if (row >= L.cells) return;
    double *A = blocks + static_cast<long long>(row) * m * m;
    auto at = [&](int i, int j) -> double & { return A[i + m * j]; };
    for (int w = 0; w < 3; ++w)
        for (int j = 0; j < N; ++j)
            for (int v = 0; v < 3; ++v)
                for (int i = 0; i < N; ++i) {
                    double value = 0;
                    if (N > 1) {
                        if (L.Q == N) {
                            value = -coef[(static_cast<long long>(row) * N + j) * 9 + v + 3 * w] * L.wd[j + N * i];
                        } else {
                            double s = 0;
                            for (int q = 0; q < L.Q; ++q)
                                s += L.wd[q + L.Q * i] *
                                     (coef[(static_cast<long long>(row) * L.Q + q) * 9 + v + 3 * w] * L.vq[q + L.Q * j]);
                            value = -s;
                        }
                    }
                    at(v * N + i, w * N + j) = value;
                }
    for (int k = 0; k < m; ++k) at(k, k) += L.mass[k % N] * inv_step;
    // left face (row - 1 | row): the cell is its right element, trial left, R
    // derivative, test -left; right face (row | row + 1): trial right, L
    // derivative, test +right
    const int p = prev_cell(row, L.cells);
    for (int side = 0; side < 2; ++side) {
        const double *test = side == 0 ? L.left : L.right, *trial = side == 0 ? L.left : L.right;
        const double sign = side == 0 ? -1.0 : 1.0;
        const double *d = side == 0 ? fr + 9 * p : fl + 9 * row;
        for (int w = 0; w < 3; ++w)
            for (int j = 0; j < N; ++j)
                for (int v = 0; v < 3; ++v)
                    for (int i = 0; i < N; ++i) at(v * N + i, w * N + j) += (sign * trial[j] * d[v + 3 * w]) * test[i];
    }
    if (!factorize) return;
    if (!lu_in_place(A, rowsrc + static_cast<long long>(row) * m, m)) atomicOr(flag, 8);
}

// BlockJacobi::apply of one block (block_jacobi.hpp:24-39): permute, unit
// lower solve, upper solve
template <int N>
__device__ __forceinline__ void bj_cell(int e, const double *r, double *z, const double *factors, const int *rowsrc) {
    constexpr int m = 3 * N;
    lu_solve_cell<m>(factors + static_cast<long long>(e) * m * m, rowsrc + static_cast<long long>(e) * m,
                     r + static_cast<long long>(e) * m, z + static_cast<long long>(e) * m);
}
template <int N>
__global__ __launch_bounds__(CELL_BLOCK) void k1_bj_apply(const double *r, double *z, const double *factors,
                                                          const int *rowsrc, int cells) {
    const int e = blockIdx.x * CELL_BLOCK + threadIdx.x;
    if (e < cells) bj_cell<N>(e, r, z, factors, rowsrc);
}

// ------------------------------------------------------ condensed HDG
// hybrid.cpp:42-81 (1D) for cell e: A = volume block + M / tau + the HDG face
// terms shape_i shape_j d_s(v, w) with d_s = sign_s F'(U_s) + stab_s I (side 0:
// sign -1, stab b of face e-1; side 1: sign +1, stab a of face e); B(v n + i,
// 3 s + v) = -shape_i stab_s; C(3 s + w, v n + i) = shape_i d_s(w, v). The
// partial-pivot LU of A (kept in factors) gives X = A^-1 B (m x 6), y = A^-1 r;
// S = -C X (6 x 6) and C y (the cell's part of the Schur right-hand side).
template <int N>
__device__ __forceinline__ void hdg_side(const double *u, int e, int side, const double *speeds, const Line &L,
                                         double d[9], double &stab, const double *&shape) {
    const int face = side == 0 ? prev_cell(e, L.cells) : e;
    shape = side == 0 ? L.left : L.right;
    stab = side == 0 ? speeds[2 * face + 1] : speeds[2 * face];
    const double sign = side == 0 ? -1.0 : 1.0;
    double z[3], J[9];
    sample1<N>(u, e, shape, z);
    jacobian1(z, L.gamma, J);
    for (int k = 0; k < 9; ++k) d[k] = sign * J[k] + stab * (k % 4 == 0 ? 1.0 : 0.0);
}
template <int N>
__global__ __launch_bounds__(CELL_BLOCK) void k1_condense(const double *u, const double *r, const double *speeds,
                                                          const double *coef, double inv_step, double *factors,
                                                          int *rowsrc, double *X, double *Y, double *S, double *CY,
                                                          int *flag, const Line *__restrict__ LP) {
    constexpr int m = 3 * N;
    const Line &L = *LP;
    const int e = blockIdx.x * CELL_BLOCK + threadIdx.x;
    if (e >= L.cells) return;
    double *A = factors + static_cast<long long>(e) * m * m;
    auto at = [&](int i, int j) -> double & { return A[i + m * j]; };
    // volume block (volume_jacobian.cpp:105-132), as k1_block without its faces
    for (int w = 0; w < 3; ++w)
        for (int j = 0; j < N; ++j)
            for (int v = 0; v < 3; ++v)
                for (int i = 0; i < N; ++i) {
                    double value = 0;
                    if (N > 1) {
                        if (L.Q == N) {
                            value = -coef[(static_cast<long long>(e) * N + j) * 9 + v + 3 * w] * L.wd[j + N * i];
                        } else {
                            double s = 0;
                            for (int q = 0; q < L.Q; ++q)
                                s += L.wd[q + L.Q * i] *
                                     (coef[(static_cast<long long>(e) * L.Q + q) * 9 + v + 3 * w] * L.vq[q + L.Q * j]);
                            value = -s;
                        }
                    }
                    at(v * N + i, w * N + j) = value;
                }
    for (int k = 0; k < m; ++k) at(k, k) += L.mass[k % N] * inv_step;
    double d[2][9], stab[2];
    const double *shape[2];
    for (int side = 0; side < 2; ++side) {
        hdg_side<N>(u, e, side, speeds, L, d[side], stab[side], shape[side]);
        for (int v = 0; v < 3; ++v)
            for (int i = 0; i < N; ++i)
                for (int w = 0; w < 3; ++w)
                    for (int j = 0; j < N; ++j)
                        at(v * N + i, w * N + j) += shape[side][i] * shape[side][j] * d[side][v + 3 * w];
    }
    int *perm = rowsrc + static_cast<long long>(e) * m;
    bool ok = true;
    for (int k = 0; k < m * m; ++k) ok = ok && isfinite(A[k]);
    if (!ok) {
        atomicOr(flag, 16);
        return;
    }
    if (!lu_in_place(A, perm, m)) {
        atomicOr(flag, 32);
        return;
    }
    double *Xe = X + static_cast<long long>(e) * m * 6, *Ye = Y + static_cast<long long>(e) * m;
    double col[m];
    for (int side = 0; side < 2; ++side)
        for (int v = 0; v < 3; ++v) {
            for (int k = 0; k < m; ++k) col[k] = 0;
            for (int i = 0; i < N; ++i) col[v * N + i] = -shape[side][i] * stab[side];
            lu_solve_cell<m>(A, perm, col, Xe + m * (side * 3 + v));
        }
    lu_solve_cell<m>(A, perm, r + static_cast<long long>(e) * m, Ye);
    // C(3 s + w, v N + i) = shape_s[i] d_s(w, v)
    auto Cm = [&](int a, int k) {
        const int side = a / 3, w = a % 3, v = k / N, i = k % N;
        return shape[side][i] * d[side][w + 3 * v];
    };
    double *Se = S + static_cast<long long>(e) * 36, *CYe = CY + static_cast<long long>(e) * 6;
    for (int j = 0; j < 6; ++j)
        for (int a = 0; a < 6; ++a) {
            double acc = 0;
            for (int k = 0; k < m; ++k) acc += Cm(a, k) * Xe[k + m * j];
            Se[a + 6 * j] = -acc;
        }
    for (int a = 0; a < 6; ++a) {
        double acc = 0;
        for (int k = 0; k < m; ++k) acc += Cm(a, k) * Ye[k];
        CYe[a] = acc;
    }
}
// Schur right-hand side of face f: the C y of its two cells in the cell order
// of hybrid.cpp:76-80 (face cells-1 gets cell 0's part first)
__global__ void k1_hdg_rhs(const double *CY, double *rhs, int cells) {
    const int f = blockIdx.x * CELL_BLOCK + threadIdx.x;
    if (f >= cells) return;
    const int nx = f + 1 == cells ? 0 : f + 1;
    for (int w = 0; w < 3; ++w) {
        const double own = CY[6 * f + 3 + w], next = CY[6 * nx + w];
        rhs[3 * f + w] = f + 1 == cells ? (0.0 + next) + own : (0.0 + own) + next;
    }
}
// HybridOperator::apply (hybrid.cpp:82-101) for face f: -(a + b) x_f plus the
// S of its two cells on their [x_{c-1}; x_c], in the reference's cell order
__device__ __forceinline__ void schur_face(int f, const double *x, double *y, const double *S, const double *speeds,
                                           int cells) {
    const int nx = f + 1 == cells ? 0 : f + 1;
    double acc[3];
    for (int w = 0; w < 3; ++w) acc[w] = -(speeds[2 * f] + speeds[2 * f + 1]) * x[3 * f + w];
    auto add = [&](int cell, int side) {
        const int p = cell == 0 ? cells - 1 : cell - 1;
        const double g[6] = {x[3 * p], x[3 * p + 1], x[3 * p + 2], x[3 * cell], x[3 * cell + 1], x[3 * cell + 2]};
        const double *Sc = S + static_cast<long long>(cell) * 36;
        for (int w = 0; w < 3; ++w) {
            double s = 0;
            for (int k = 0; k < 6; ++k) s += Sc[side * 3 + w + 6 * k] * g[k];
            acc[w] += s;
        }
    };
    if (f + 1 == cells) {
        add(nx, 0);
        add(f, 1);
    } else {
        add(f, 1);
        add(nx, 0);
    }
    for (int w = 0; w < 3; ++w) y[3 * f + w] = acc[w];
}
__global__ void k1_schur_apply(const double *x, double *y, const double *S, const double *speeds, int cells) {
    const int f = blockIdx.x * CELL_BLOCK + threadIdx.x;
    if (f < cells) schur_face(f, x, y, S, speeds, cells);
}
// HybridOperator::block(f, f) (hybrid.cpp:102-115): -(a + b) I + S_f(side 1,
// side 1) + S_{f+1}(side 0, side 0); written to blocks (3x3 column-major) and,
// when factorize, replaced by its partial-pivot LU
__global__ void k1_face_blocks(const double *S, const double *speeds, double *blocks, int *perm, int *flag,
                               int factorize, int cells) {
    const int f = blockIdx.x * CELL_BLOCK + threadIdx.x;
    if (f >= cells) return;
    const int nx = f + 1 == cells ? 0 : f + 1;
    double *B = blocks + 9 * f;
    for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a) B[a + 3 * b] = a == b ? -(speeds[2 * f] + speeds[2 * f + 1]) : 0.0;
    for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a) B[a + 3 * b] += S[36LL * f + 3 + a + 6 * (3 + b)];
    for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a) B[a + 3 * b] += S[36LL * nx + a + 6 * b];
    if (factorize && !lu_in_place(B, perm + 3 * f, 3)) atomicOr(flag, 8);
}
__device__ __forceinline__ void face_bj(int f, const double *r, double *z, const double *fac, const int *perm) {
    lu_solve_cell<3>(fac + 9 * f, perm + 3 * f, r + 3 * f, z + 3 * f);
}
__global__ void k1_face_bj_apply(const double *r, double *z, const double *fac, const int *perm, int cells) {
    const int f = blockIdx.x * CELL_BLOCK + threadIdx.x;
    if (f < cells) face_bj(f, r, z, fac, perm);
}
// HybridOperator::recover (hybrid.cpp:128-137): du = -y - X [dl_{e-1}; dl_e]
template <int N>
__global__ __launch_bounds__(CELL_BLOCK) void k1_recover(const double *dl, double *du, const double *X,
                                                         const double *Y, int cells) {
    constexpr int m = 3 * N;
    const int e = blockIdx.x * CELL_BLOCK + threadIdx.x;
    if (e >= cells) return;
    const int p = e == 0 ? cells - 1 : e - 1;
    const double g[6] = {dl[3 * p], dl[3 * p + 1], dl[3 * p + 2], dl[3 * e], dl[3 * e + 1], dl[3 * e + 2]};
    const double *Xe = X + static_cast<long long>(e) * m * 6, *Ye = Y + static_cast<long long>(e) * m;
    for (int k = 0; k < m; ++k) {
        double acc = 0;
        for (int j = 0; j < 6; ++j) acc += Xe[k + m * j] * g[j];
        du[static_cast<long long>(e) * m + k] = -Ye[k] - acc;
    }
}
// HybridOperator::recovered_residual (hybrid.cpp:116-127): with e = S dl - rhs
// (in err), each side's e divided by -(a + b) of its face, lifted by -B
template <int N>
__global__ __launch_bounds__(CELL_BLOCK) void k1_recovered(const double *err, const double *rhs, double *out,
                                                           const double *speeds, const Line *__restrict__ LP) {
    const Line &L = *LP;
    const int e = blockIdx.x * CELL_BLOCK + threadIdx.x;
    if (e >= L.cells) return;
    const int p = prev_cell(e, L.cells);
    double g[6];
    for (int side = 0; side < 2; ++side) {
        const int f = side == 0 ? p : e;
        for (int v = 0; v < 3; ++v)
            g[side * 3 + v] = (err[3 * f + v] - rhs[3 * f + v]) / -(speeds[2 * f] + speeds[2 * f + 1]);
    }
    const double st[2] = {speeds[2 * p + 1], speeds[2 * e]};
    for (int v = 0; v < 3; ++v)
        for (int i = 0; i < N; ++i) {
            double acc = 0;
            for (int j = 0; j < 6; ++j) {
                const int side = j / 3;
                const double b = j % 3 == v ? -(side == 0 ? L.left[i] : L.right[i]) * st[side] : 0.0;
                acc += b * g[j];
            }
            out[static_cast<long long>(e) * 3 * N + v * N + i] = -acc;
        }
}

// ------------------------------------------------------------ reductions
// Sum over the workgroup's threads of count (<= 32) per-thread values: a butterfly
// within each wave, then the 16 wave sums in wave order. Deterministic.
__device__ void block_sums(double *part, int count, double (*red)[32], double *out) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < count; ++k) {
        double v = part[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < count) {
        double s = 0;
        for (int w = 0; w < SOLVE_WAVES; ++w) s += red[w][threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}
// mode 0: sum a^2 w; mode 1: sum a^2 / w (the mass-weighted norms)
__global__ __launch_bounds__(SOLVE_THREADS) void k1_norm2(const double *a, const double *w, long long D, int mode,
                                                          double *out) {
    __shared__ double red[SOLVE_WAVES][32], res[32];
    double part[1] = {0};
    for (long long k = threadIdx.x; k < D; k += SOLVE_THREADS)
        part[0] += mode == 0 ? a[k] * a[k] * w[k] : a[k] * a[k] / w[k];
    block_sums(part, 1, red, res);
    if (threadIdx.x == 0) out[0] = res[0];
}
// Problem::integral (discretization.cpp:148-155) and Problem::error
// (discretization.cpp:156-168): per-cell contributions, block sums
template <int N>
__global__ __launch_bounds__(SOLVE_THREADS) void k1_integrals(const double *u, double t, int error, double *out,
                                                              const Line *__restrict__ LP) {
    const Line &L = *LP;
    __shared__ double red[SOLVE_WAVES][32], res[32];
    double part[3] = {0, 0, 0};
    for (int e = threadIdx.x; e < L.cells; e += SOLVE_THREADS) {
        const double *c = u + static_cast<long long>(e) * 3 * N;
        if (!error) {
            for (int v = 0; v < 3; ++v)
                for (int i = 0; i < N; ++i) part[v] += .5 * L.h * L.w[i] * c[v * N + i];
        } else {
            for (int q = 0; q < L.eq; ++q) {
                double z[3] = {0, 0, 0}, x[3];
                for (int v = 0; v < 3; ++v)
                    for (int i = 0; i < N; ++i) z[v] += L.ev[q + L.eq * i] * c[v * N + i];
                simple_wave(-1 + L.h * (e + .5) + .5 * L.h * L.ex[q], t, L.gamma, L.amplitude, L.entropy, x);
                for (int v = 0; v < 3; ++v) part[v] += .5 * L.h * L.ew[q] * (z[v] - x[v]) * (z[v] - x[v]);
            }
        }
    }
    block_sums(part, 3, red, res);
    if (threadIdx.x < 3) out[threadIdx.x] = res[threadIdx.x];
}
__global__ void k1_axpby(double *z, double a, const double *x, double b, const double *y, long long D) {
    const long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < D) z[k] = a * x[k] + b * y[k];
}
// z = x + c y
__global__ void k1_axpy(double *z, const double *x, double c, const double *y, long long D) {
    const long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < D) z[k] = x[k] + c * y[k];
}
// z = (x - y) / c
__global__ void k1_diffdiv(double *z, const double *x, const double *y, double c, long long D) {
    const long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < D) z[k] = (x[k] - y[k]) / c;
}

// ----------------------------------------------------------------- GMRES
// GMRES::solve (gmres.cpp:17-86) with right block-Jacobi preconditioning and
// the M^-1 metric, as one workgroup: vectors are spread over the 512
// threads, cell work (the preconditioner and J) gives a thread a cell, the
// weighted CGS2 dot products reduce on chip, and thread 0 applies the Givens
// rotations. status: 0 converged, 1 Arnoldi breakdown, 2 nonfinite
// residual, 3 iteration limit.
struct SolveArgs {
    const double *coef, *fl, *fr, *factors, *inv_mass;
    const int *rowsrc;
    const double *rhs;
    double *x, *V, *w, *z, *res, *result;
    double inv_step, tol;
    long long D;
    int restart, limit;
    // condensed HDG: the trace system with the cells' S, the face speeds and
    // the face-block LUs (weights inv_mass = 1)
    const double *S, *speeds;
};
template <int N, bool HDG>
__global__ __launch_bounds__(SOLVE_THREADS) void k1_gmres(SolveArgs a, const Line *__restrict__ LP) {
    const Line &L = *LP;
    __shared__ double red[SOLVE_WAVES][32], dots[32];
    __shared__ double H[(MAX_RESTART + 1) * MAX_RESTART], cs[MAX_RESTART], sn[MAX_RESTART], g[MAX_RESTART + 1];
    __shared__ double coefficients[MAX_RESTART];
    __shared__ int control[2];  // count, stop
    const int t = threadIdx.x, R = a.restart, cells = L.cells;
    const long long D = a.D;
    auto Vk = [&](int k) { return a.V + static_cast<long long>(k) * D; };
    auto norm = [&](const double *v) {
        double part[1] = {0};
        for (long long i = t; i < D; i += SOLVE_THREADS) part[0] += v[i] * (a.inv_mass[i] * v[i]);
        block_sums(part, 1, red, dots);
        return sqrt(fmax(0.0, dots[0]));
    };
    auto precondition = [&](const double *in, double *out) {
        for (int e = t; e < cells; e += SOLVE_THREADS) {
            if (HDG) face_bj(e, in, out, a.factors, a.rowsrc);
            else bj_cell<N>(e, in, out, a.factors, a.rowsrc);
        }
        __syncthreads();
    };
    auto jacobian = [&](const double *in, double *out) {
        for (int e = t; e < cells; e += SOLVE_THREADS) {
            if (HDG) schur_face(e, in, out, a.S, a.speeds, cells);
            else apply_cell<N>(e, in, out, a.coef, a.fl, a.fr, a.inv_step, L);
        }
        __syncthreads();
    };
    for (long long i = t; i < D; i += SOLVE_THREADS) {
        a.x[i] = 0;
        a.res[i] = a.rhs[i];
    }
    __syncthreads();
    const double rhs_norm = norm(a.rhs);
    if (rhs_norm == 0) {
        if (t == 0) a.result[0] = 0, a.result[1] = 0, a.result[2] = 0;
        return;
    }
    int iterations = 0;
    while (iterations < a.limit) {
        const double beta = norm(a.res);
        if (beta <= a.tol * rhs_norm) {
            if (t == 0) a.result[0] = iterations, a.result[1] = beta / rhs_norm, a.result[2] = 0;
            return;
        }
        for (long long i = t; i < D; i += SOLVE_THREADS) Vk(0)[i] = a.res[i] / beta;
        for (int k = t; k < (MAX_RESTART + 1) * MAX_RESTART; k += SOLVE_THREADS) H[k] = 0;
        if (t <= MAX_RESTART) g[t] = t == 0 ? beta : 0.0;
        if (t == 0) control[0] = 0, control[1] = 0;
        __syncthreads();
        for (int j = 0; j < R && iterations < a.limit; ++j) {
            precondition(Vk(j), a.z);
            jacobian(a.z, a.w);
            for (int pass = 0; pass < 2; ++pass) {
                // (V_k, M^-1 w) for k <= j: DOT_CHUNKS consecutive threads per
                // vector, each a strided partial sum, then a butterfly over them
                const int slot = t / DOT_CHUNKS, chunk = t % DOT_CHUNKS;
                double s = 0;
                if (slot <= j)
                    for (long long i = chunk; i < D; i += DOT_CHUNKS) s += Vk(slot)[i] * (a.inv_mass[i] * a.w[i]);
#pragma unroll
                for (int off = DOT_CHUNKS / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
                if (chunk == 0 && slot <= j) dots[slot] = s;
                __syncthreads();
                if (t <= j) H[t + (R + 1) * j] += dots[t];
                for (long long i = t; i < D; i += SOLVE_THREADS) {
                    double s = 0;
                    for (int k = 0; k <= j; ++k) s += Vk(k)[i] * dots[k];
                    a.w[i] -= s;
                }
                __syncthreads();
            }
            const double next = norm(a.w);
            if (next > 0)
                for (long long i = t; i < D; i += SOLVE_THREADS) Vk(j + 1)[i] = a.w[i] / next;
            if (t == 0) {
                double *Hj = H + (R + 1) * j;
                Hj[j + 1] = next;
                for (int i = 0; i < j; ++i) {
                    const double p = Hj[i], q = Hj[i + 1];
                    Hj[i] = cs[i] * p + sn[i] * q;
                    Hj[i + 1] = -sn[i] * p + cs[i] * q;
                }
                const double d = hypot(Hj[j], Hj[j + 1]);
                if (!(isfinite(d) && d > 0)) {
                    control[1] = 2;  // breakdown
                } else {
                    cs[j] = Hj[j] / d;
                    sn[j] = Hj[j + 1] / d;
                    Hj[j] = d;
                    Hj[j + 1] = 0;
                    g[j + 1] = -sn[j] * g[j];
                    g[j] *= cs[j];
                    control[0] = j + 1;
                    if (fabs(g[j + 1]) <= a.tol * rhs_norm || next <= 2.2250738585072014e-308) control[1] = 1;
                }
            }
            __syncthreads();
            if (control[1] == 2) {
                if (t == 0) a.result[0] = iterations, a.result[1] = -1, a.result[2] = 1;
                return;
            }
            ++iterations;
            if (control[1] == 1) break;
        }
        const int count = control[0];
        if (t == 0)
            for (int i = count - 1; i >= 0; --i) {
                double value = g[i];
                for (int j = i + 1; j < count; ++j) value -= H[i + (R + 1) * j] * coefficients[j];
                coefficients[i] = value / H[i + (R + 1) * i];
            }
        __syncthreads();
        for (long long i = t; i < D; i += SOLVE_THREADS) {
            double s = 0;
            for (int k = 0; k < count; ++k) s += Vk(k)[i] * coefficients[k];
            a.w[i] = s;
        }
        __syncthreads();
        precondition(a.w, a.z);
        for (long long i = t; i < D; i += SOLVE_THREADS) a.x[i] += a.z[i];
        __syncthreads();
        jacobian(a.x, a.w);
        for (long long i = t; i < D; i += SOLVE_THREADS) a.res[i] = a.rhs[i] - a.w[i];
        __syncthreads();
        const double relative = norm(a.res) / rhs_norm;
        if (!isfinite(relative)) {
            if (t == 0) a.result[0] = iterations, a.result[1] = relative, a.result[2] = 2;
            return;
        }
        if (relative <= a.tol) {
            if (t == 0) a.result[0] = iterations, a.result[1] = relative, a.result[2] = 0;
            return;
        }
    }
    if (t == 0) a.result[0] = iterations, a.result[1] = -1, a.result[2] = 3;
}

// ------------------------------------------------------------ host side
void require1(bool condition, const char *message) {
    if (!condition) throw std::runtime_error(message);
}
void hip_check1(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}
#define HIPCHECK1(call) hip_check1((call), #call)

// device buffer that frees itself (also when a constructor throws half way)
template <class T> struct Buf {
    T *p = nullptr;
    size_t n = 0;
    Buf() = default;
    Buf(const Buf &) = delete;
    Buf &operator=(const Buf &) = delete;
    Buf(Buf &&o) noexcept : p(o.p), n(o.n) {
        o.p = nullptr;
        o.n = 0;
    }
    Buf &operator=(Buf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~Buf() { release(); }
    void alloc(size_t count) {
        if (p && n >= count) return;
        release();
        HIPCHECK1(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

#define DGH1_DISPATCH(n, ...)                     \
    switch (n) {                                  \
    case 1: { constexpr int N = 1; __VA_ARGS__; } break; \
    case 2: { constexpr int N = 2; __VA_ARGS__; } break; \
    case 3: { constexpr int N = 3; __VA_ARGS__; } break; \
    case 4: { constexpr int N = 4; __VA_ARGS__; } break; \
    case 5: { constexpr int N = 5; __VA_ARGS__; } break; \
    case 6: { constexpr int N = 6; __VA_ARGS__; } break; \
    case 7: { constexpr int N = 7; __VA_ARGS__; } break; \
    case 8: { constexpr int N = 8; __VA_ARGS__; } break; \
    case 9: { constexpr int N = 9; __VA_ARGS__; } break; \
    default: throw std::runtime_error("Invalid mesh or degree"); \
    }

using Clock = std::chrono::steady_clock;
double seconds_since(Clock::time_point s) { return std::chrono::duration<double>(Clock::now() - s).count(); }

}  // namespace
}  // namespace dgh

using namespace dgh;

struct dgh1_ctx {
    dgh1_config cfg{};
    int n = 0, Q = 0, cells = 0, m = 0;
    long long D = 0;
    double h = 0, tau = 0;
    hipStream_t stream = nullptr;
    Line hline{};
    Buf<Line> line;
    Buf<double> mass, inv_mass, old, speeds, coef, fl, fr, factors, scratch, V, w, z, res, result, sums;
    Buf<double> state, trial, residual, trial_residual, direction, rhs, step_state, stage_base, slopes;
    Buf<int> rowsrc, flag, face_perm;
    Buf<unsigned long long> rate_bits;
    // condensed HDG: X, y, S, C y per cell; face-block LUs; trace vectors
    Buf<double> hX, hY, hS, hCY, face_lu, trhs, tdir, terr, ones;
    bool frozen = false, refreshed = false, factored = false, hdg_ready = false;
    std::string error;

    explicit dgh1_ctx(const dgh1_config &c) : cfg(c) {
        require1(c.cells >= 3 && c.degree >= 0 && c.degree <= 8 && c.quadrature >= 0, "Invalid mesh or degree");
        require1(c.quadrature == 0 || c.quadrature >= c.degree + 1, "Quadrature must be >= degree+1");
        require1(c.quadrature <= MAXQ, "Over-integration above 16 points is not supported on the device");
        require1(c.restart > 0 && c.max_linear > 0 && c.max_newton > 0 && c.steps >= 0, "Invalid iteration limit");
        require1(c.restart <= MAX_RESTART, "GMRES restart above 31 is not supported on the device");
        n = c.degree + 1;
        Q = c.quadrature ? c.quadrature : n;
        cells = c.cells;
        m = 3 * n;
        D = static_cast<long long>(cells) * m;
        h = 2.0 / cells;
        const Basis b = make_basis(n, 1.0, 1.0, Q > n ? Q : 0);
        hline.n = n;
        hline.Q = Q;
        hline.eq = std::max(20, 2 * n + 3);
        hline.cells = cells;
        hline.h = h;
        hline.gamma = 3.0;  // the simple wave's gas (simple_wave.hpp:6)
        hline.safety = c.safety;
        hline.amplitude = c.testcase == 1 ? -0.4 : 1.0;
        hline.entropy = c.testcase == 1 ? 1 / hline.gamma : 1.0;
        for (int q = 0; q < Q; ++q)
            for (int i = 0; i < n; ++i) {
                hline.wd[q + Q * i] = Q > n ? b.dq[q + Q * i] : b.wd[q + n * i];
                hline.vq[q + Q * i] = Q > n ? b.vq[q + Q * i] : (q == i ? 1.0 : 0.0);
            }
        for (int i = 0; i < n; ++i) {
            hline.w[i] = b.w[i];
            hline.left[i] = b.left[i];
            hline.right[i] = b.right[i];
            hline.mass[i] = .5 * h * b.w[i];
        }
        for (int q = 0; q < hline.eq; ++q) {
            hline.ex[q] = b.ex[q];
            hline.ew[q] = b.ew[q];
            for (int i = 0; i < n; ++i) {
                hline.ev[q + hline.eq * i] = b.ev[q + hline.eq * i];
                hline.proj[i + n * q] = b.proj[i + n * q];
            }
        }
        line.alloc(1);
        HIPCHECK1(hipMemcpy(line.p, &hline, sizeof(Line), hipMemcpyHostToDevice));
        std::vector<double> hm(D), him(D);
        for (long long k = 0; k < D; ++k) {
            hm[k] = hline.mass[k % n];
            him[k] = 1.0 / hm[k];
        }
        for (auto *b2 : {&mass, &inv_mass, &old, &scratch, &w, &z, &res, &state, &trial, &residual, &trial_residual,
                         &direction, &rhs, &step_state, &stage_base})
            b2->alloc(D);
        HIPCHECK1(hipMemcpy(mass.p, hm.data(), D * sizeof(double), hipMemcpyHostToDevice));
        HIPCHECK1(hipMemcpy(inv_mass.p, him.data(), D * sizeof(double), hipMemcpyHostToDevice));
        speeds.alloc(2 * cells);
        coef.alloc(static_cast<size_t>(cells) * Q * 9);
        fl.alloc(9 * cells);
        fr.alloc(9 * cells);
        factors.alloc(static_cast<size_t>(cells) * m * m);
        rowsrc.alloc(static_cast<size_t>(cells) * m);
        V.alloc(static_cast<size_t>(std::min<long long>(D, c.restart) + 1) * D);
        slopes.alloc(4 * D);
        result.alloc(4);
        sums.alloc(4);
        face_perm.alloc(3 * cells);
        for (auto *b2 : {&trhs, &tdir, &terr, &ones}) b2->alloc(3 * cells);
        const std::vector<double> one(3 * cells, 1.0);
        HIPCHECK1(hipMemcpy(ones.p, one.data(), 3 * cells * sizeof(double), hipMemcpyHostToDevice));
        HIPCHECK1(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));  // last: nothing throws after it
        flag.alloc(1);
        rate_bits.alloc(1);
    }
    ~dgh1_ctx() {
        for (auto *b : {&mass, &inv_mass, &old, &speeds, &coef, &fl, &fr, &factors, &scratch, &V, &w, &z, &res,
                        &result, &sums, &state, &trial, &residual, &trial_residual, &direction, &rhs, &step_state,
                        &stage_base, &slopes, &hX, &hY, &hS, &hCY, &face_lu, &trhs, &tdir, &terr, &ones})
            b->release();
        rowsrc.release();
        face_perm.release();
        flag.release();
        rate_bits.release();
        line.release();
        if (stream) (void)hipStreamDestroy(stream);
    }
    unsigned cell_grid() const { return static_cast<unsigned>((cells + CELL_BLOCK - 1) / CELL_BLOCK); }
    unsigned vec_grid() const { return static_cast<unsigned>((D + 255) / 256); }
    void sync() { HIPCHECK1(hipStreamSynchronize(stream)); }
    int read_flag() {
        int f = 0;
        HIPCHECK1(hipMemcpyAsync(&f, flag.p, sizeof(int), hipMemcpyDeviceToHost, stream));
        sync();
        return f;
    }
    void clear_flag() { HIPCHECK1(hipMemsetAsync(flag.p, 0, sizeof(int), stream)); }
    void copy(double *dst, const double *src) {
        HIPCHECK1(hipMemcpyAsync(dst, src, D * sizeof(double), hipMemcpyDeviceToDevice, stream));
    }

    // ---- Problem
    void initial(double *u) {
        clear_flag();
        DGH1_DISPATCH(n, { k1_initial<N><<<cell_grid(), CELL_BLOCK, 0, stream>>>(u, flag.p, line.p); })
        HIPCHECK1(hipGetLastError());
        require1(!(read_flag() & 1), "Nonphysical cell average");
        require1(admissible(u), "Invalid projected initial state");
    }
    long long trace_size() const { return 3LL * cells; }
    double freeze(const double *u, double *host_speeds = nullptr) {
        copy(old.p, u);
        clear_flag();
        HIPCHECK1(hipMemsetAsync(rate_bits.p, 0, sizeof(unsigned long long), stream));
        DGH1_DISPATCH(n, {
            k1_speeds<N><<<cell_grid(), CELL_BLOCK, 0, stream>>>(u, speeds.p, flag.p, line.p);
            k1_rate<N><<<cell_grid(), CELL_BLOCK, 0, stream>>>(u, speeds.p, rate_bits.p, flag.p, line.p);
        })
        HIPCHECK1(hipGetLastError());
        unsigned long long bits = 0;
        HIPCHECK1(hipMemcpyAsync(&bits, rate_bits.p, sizeof(bits), hipMemcpyDeviceToHost, stream));
        const int f = read_flag();
        require1(!(f & 2), "Invalid frozen face state");
        double rate;
        std::memcpy(&rate, &bits, sizeof(rate));
        require1(!(f & 4) && std::isfinite(rate) && rate > 0, "Invalid CFL rate");
        if (host_speeds)
            HIPCHECK1(hipMemcpy(host_speeds, speeds.p, 2 * cells * sizeof(double), hipMemcpyDeviceToHost));
        frozen = true;
        refreshed = factored = hdg_ready = false;
        return rate;
    }
    void set_step(double t, const double *base) {
        tau = t;
        copy(old.p, base);
        refreshed = factored = hdg_ready = false;
    }
    void residual_of(const double *u, double *r) {
        require1(frozen && tau > 0, "Freeze the state and set the step before the residual");
        DGH1_DISPATCH(n, { k1_residual<N><<<cell_grid(), CELL_BLOCK, 0, stream>>>(u, old.p, speeds.p, r, tau, line.p); })
        HIPCHECK1(hipGetLastError());
    }
    bool admissible(const double *u) {
        clear_flag();
        DGH1_DISPATCH(n, { k1_admissible<N><<<cell_grid(), CELL_BLOCK, 0, stream>>>(u, flag.p, line.p); })
        HIPCHECK1(hipGetLastError());
        return read_flag() == 0;
    }
    double norm2(const double *a, int mode) {  // mode 0: sum a^2 M; 1: sum a^2 / M
        k1_norm2<<<1, SOLVE_THREADS, 0, stream>>>(a, mass.p, D, mode, sums.p);
        HIPCHECK1(hipGetLastError());
        double s = 0;
        HIPCHECK1(hipMemcpyAsync(&s, sums.p, sizeof(double), hipMemcpyDeviceToHost, stream));
        sync();
        return s;
    }
    double dual_norm(const double *r) { return std::sqrt(norm2(r, 1)); }   // time_integrator.cpp:38-40
    double state_norm(const double *x) { return std::sqrt(norm2(x, 0)); }  // time_integrator.cpp:41-43
    void integrals(const double *u, double t, int error, double *out3) {
        DGH1_DISPATCH(n, { k1_integrals<N><<<1, SOLVE_THREADS, 0, stream>>>(u, t, error, sums.p, line.p); })
        HIPCHECK1(hipGetLastError());
        HIPCHECK1(hipMemcpyAsync(out3, sums.p, 3 * sizeof(double), hipMemcpyDeviceToHost, stream));
        sync();
        if (error)
            for (int v = 0; v < 3; ++v) out3[v] = std::sqrt(out3[v]);
    }

    // ---- DG operator and block Jacobi
    void refresh(const double *u) {
        require1(tau > 0, "Set the time step before refreshing the volume Jacobian");
        DGH1_DISPATCH(n, { k1_refresh<N><<<cell_grid(), CELL_BLOCK, 0, stream>>>(u, speeds.p, coef.p, fl.p, fr.p, line.p); })
        HIPCHECK1(hipGetLastError());
        refreshed = true;
        factored = false;
    }
    void apply(const double *x, double *y) {
        require1(refreshed, "Refresh the DG Jacobian before applying it");
        DGH1_DISPATCH(n, { k1_apply<N><<<cell_grid(), CELL_BLOCK, 0, stream>>>(x, y, coef.p, fl.p, fr.p, 1 / tau, line.p); })
        HIPCHECK1(hipGetLastError());
    }
    void block(int row, double *host_out) {
        require1(refreshed, "Refresh the DG Jacobian before assembling a block");
        require1(row >= 0 && row < cells, "Block row out of range");
        Buf<double> tmp;
        tmp.alloc(static_cast<size_t>(cells) * m * m);
        DGH1_DISPATCH(n, {
            k1_block<N><<<cell_grid(), CELL_BLOCK, 0, stream>>>(coef.p, fl.p, fr.p, 1 / tau, tmp.p, rowsrc.p, flag.p, 0, line.p);
        })
        HIPCHECK1(hipGetLastError());
        HIPCHECK1(hipMemcpyAsync(host_out, tmp.p + static_cast<size_t>(row) * m * m, m * m * sizeof(double),
                                 hipMemcpyDeviceToHost, stream));
        sync();
        tmp.release();
    }
    void bj_refresh() {
        require1(refreshed, "Refresh the DG Jacobian before assembling a block");
        clear_flag();
        DGH1_DISPATCH(n, {
            k1_block<N><<<cell_grid(), CELL_BLOCK, 0, stream>>>(coef.p, fl.p, fr.p, 1 / tau, factors.p, rowsrc.p, flag.p, 1, line.p);
        })
        HIPCHECK1(hipGetLastError());
        require1(!(read_flag() & 8), "Singular block Jacobi factor");
        factored = true;
    }
    void bj_apply(const double *r, double *zz) {
        require1(factored, "Refresh the block Jacobi preconditioner first");
        DGH1_DISPATCH(n, { k1_bj_apply<N><<<cell_grid(), CELL_BLOCK, 0, stream>>>(r, zz, factors.p, rowsrc.p, cells); })
        HIPCHECK1(hipGetLastError());
    }
    // GMRES::solve; leaves rhs - J x in res (its residual())
    void gmres(const double *b, double *x, int limit, double tol, int &its, double &rel) {
        require1(factored, "Refresh the block Jacobi preconditioner first");
        const int restart = static_cast<int>(std::min<long long>(D, cfg.restart));
        SolveArgs a{coef.p, fl.p, fr.p, factors.p, inv_mass.p, rowsrc.p, b, x, V.p, w.p, z.p, res.p, result.p,
                    1 / tau, tol, D, restart, limit, nullptr, nullptr};
        DGH1_DISPATCH(n, { k1_gmres<N, false><<<1, SOLVE_THREADS, 0, stream>>>(a, line.p); })
        HIPCHECK1(hipGetLastError());
        double r[3];
        HIPCHECK1(hipMemcpyAsync(r, result.p, 3 * sizeof(double), hipMemcpyDeviceToHost, stream));
        sync();
        its = static_cast<int>(r[0]);
        rel = r[1];
        const int status = static_cast<int>(r[2]);
        require1(status != 1, "GMRES Arnoldi breakdown");
        require1(status != 2, "Nonfinite GMRES residual");
        require1(status != 3, "GMRES iteration limit reached");
    }

    // ---- condensed HDG (hybrid.cpp, 1D)
    void hdg_refresh(const double *u, const double *r) {
        require1(frozen && tau > 0, "Set the time step before refreshing the volume Jacobian");
        refresh(u);  // the volume Jacobian F'(U) at the points
        hX.alloc(static_cast<size_t>(cells) * m * 6);
        hY.alloc(static_cast<size_t>(cells) * m);
        hS.alloc(36 * static_cast<size_t>(cells));
        hCY.alloc(6 * static_cast<size_t>(cells));
        face_lu.alloc(9 * static_cast<size_t>(cells));
        clear_flag();
        DGH1_DISPATCH(n, {
            k1_condense<N><<<cell_grid(), CELL_BLOCK, 0, stream>>>(u, r, speeds.p, coef.p, 1 / tau, factors.p,
                                                                   rowsrc.p, hX.p, hY.p, hS.p, hCY.p, flag.p, line.p);
        })
        k1_hdg_rhs<<<cell_grid(), CELL_BLOCK, 0, stream>>>(hCY.p, trhs.p, cells);
        k1_face_blocks<<<cell_grid(), CELL_BLOCK, 0, stream>>>(hS.p, speeds.p, face_lu.p, face_perm.p, flag.p, 1, cells);
        HIPCHECK1(hipGetLastError());
        const int f = read_flag();
        require1(!(f & 16), "Nonfinite block LU input");
        require1(!(f & 32), "Singular or nonfinite LU factors");
        require1(!(f & 8), "Singular block Jacobi factor");
        factored = false;  // the cell LUs of the condensation reuse the block-Jacobi storage
        hdg_ready = true;
    }
    void hdg_apply(const double *x, double *y) {
        require1(hdg_ready, "Refresh the hybrid operator first");
        k1_schur_apply<<<cell_grid(), CELL_BLOCK, 0, stream>>>(x, y, hS.p, speeds.p, cells);
        HIPCHECK1(hipGetLastError());
    }
    void hdg_block(int face, double *host_out) {
        require1(hdg_ready, "Refresh the hybrid operator first");
        require1(face >= 0 && face < cells, "Block row out of range");
        Buf<double> tmp;
        tmp.alloc(9 * static_cast<size_t>(cells));
        k1_face_blocks<<<cell_grid(), CELL_BLOCK, 0, stream>>>(hS.p, speeds.p, tmp.p, face_perm.p, flag.p, 0, cells);
        HIPCHECK1(hipGetLastError());
        HIPCHECK1(hipMemcpyAsync(host_out, tmp.p + 9 * face, 9 * sizeof(double), hipMemcpyDeviceToHost, stream));
        sync();
        tmp.release();
    }
    void hdg_bj_apply(const double *r, double *zz) {
        require1(hdg_ready, "Refresh the hybrid operator first");
        k1_face_bj_apply<<<cell_grid(), CELL_BLOCK, 0, stream>>>(r, zz, face_lu.p, face_perm.p, cells);
        HIPCHECK1(hipGetLastError());
    }
    void hdg_recover(const double *dl, double *du) {
        require1(hdg_ready, "Refresh the hybrid operator first");
        DGH1_DISPATCH(n, { k1_recover<N><<<cell_grid(), CELL_BLOCK, 0, stream>>>(dl, du, hX.p, hY.p, cells); })
        HIPCHECK1(hipGetLastError());
    }
    void hdg_recovered(const double *dl, double *out) {
        hdg_apply(dl, terr.p);
        DGH1_DISPATCH(n, { k1_recovered<N><<<cell_grid(), CELL_BLOCK, 0, stream>>>(terr.p, trhs.p, out, speeds.p, line.p); })
        HIPCHECK1(hipGetLastError());
    }
    // GMRES::solve on the trace system, unweighted (time_integrator.cpp:36)
    void hdg_gmres(const double *b, double *x, int limit, double tol, int &its, double &rel) {
        require1(hdg_ready, "Refresh the hybrid operator first");
        const long long T = trace_size();
        const int restart = static_cast<int>(std::min<long long>(T, cfg.restart));
        SolveArgs a{nullptr, nullptr, nullptr, face_lu.p, ones.p, face_perm.p, b, x, V.p, w.p, z.p, res.p,
                    result.p, 1 / tau, tol, T, restart, limit, hS.p, speeds.p};
        DGH1_DISPATCH(n, { k1_gmres<N, true><<<1, SOLVE_THREADS, 0, stream>>>(a, line.p); })
        HIPCHECK1(hipGetLastError());
        double r[3];
        HIPCHECK1(hipMemcpyAsync(r, result.p, 3 * sizeof(double), hipMemcpyDeviceToHost, stream));
        sync();
        its = static_cast<int>(r[0]);
        rel = r[1];
        const int status = static_cast<int>(r[2]);
        require1(status != 1, "GMRES Arnoldi breakdown");
        require1(status != 2, "Nonfinite GMRES residual");
        require1(status != 3, "GMRES iteration limit reached");
    }

    // ---- TimeIntegrator::advance (time_integrator.cpp:45-207), method dg or hdg
    // simple_wave.hpp:8: characteristics of the initial invariant cross at 1 / (b |A| pi)
    double shock_time() const { return 4 / ((hline.gamma + 1) * std::acos(-1.0) * std::fabs(hline.amplitude)); }
    void advance(double *solution, dgh_run_stats &stats, std::vector<dgh_update_stats> &updates) {
        const dgh1_config &c = cfg;
        const bool hdg = c.method == DGH_METHOD_HDG;  // dg and both run DG
        require1(admissible(solution), "Invalid initial solution");
        const double shock = shock_time();
        require1(c.steps > 0 || c.final_time < shock, "Final time reaches simple-wave shock");
        constexpr double alpha = .24169426078821, beta = alpha / 4., eta = .12915286960590;
        constexpr double A[4][4] = {{alpha, 0, 0, 0}, {-alpha, alpha, 0, 0}, {0, 1 - alpha, alpha, 0},
                                    {beta, eta, .5 - beta - eta - alpha, alpha}};
        constexpr double B[4] = {0, 1. / 6, 1. / 6, 2. / 3}, stage_c[4] = {alpha, 0, 1, .5};
        const int stages = c.integrator == 0 ? 1 : 4;
        stats = dgh_run_stats{};
        updates.clear();
        freeze(solution);
        double initial_integral[3];
        integrals(solution, 0, 0, initial_integral);
        const auto begin = Clock::now();
        auto slope = [&](int s) { return slopes.p + static_cast<size_t>(s) * D; };
        while (c.steps > 0 ? stats.steps < c.steps : stats.final_time < c.final_time) {
            const double rate = freeze(solution);
            double step = c.cfl / rate;
            if (c.steps == 0) {
                const double remaining = c.final_time - stats.final_time;
                step = std::min(step, remaining);
                if (remaining - step <=
                    16 * std::numeric_limits<double>::epsilon() * std::max(remaining, c.final_time))
                    step = remaining;
            }
            require1(std::isfinite(step) && step > 0 && stats.final_time + step > stats.final_time, "Invalid CFL step");
            require1(stats.final_time + step < shock, "Step reaches simple-wave shock");
            copy(step_state.p, solution);
            copy(state.p, solution);
            for (int stage = 0; stage < stages; ++stage) {
                const double diagonal = stages == 1 ? 1. : A[stage][stage];
                copy(stage_base.p, step_state.p);
                for (int j = 0; j < stage; ++j)
                    k1_axpy<<<vec_grid(), 256, 0, stream>>>(stage_base.p, stage_base.p, step * A[stage][j], slope(j), D);
                if (stage > 0) freeze(state.p);
                set_step(diagonal * step, stage_base.p);
                const double old_norm = state_norm(stage_base.p);
                const int limit = c.newton_steps > 0 ? c.newton_steps : c.max_newton;
                double initial_rms = -1.;
                bool converged = false, residual_ready = false;
                for (int it = 0; it <= limit; ++it) {
                    const auto update_start = Clock::now();
                    if (!residual_ready) residual_of(state.p, residual.p);
                    const double norm = dual_norm(residual.p), rms = norm / std::sqrt(2.);
                    if (initial_rms < 0) initial_rms = rms;
                    if (c.newton_steps == 0 && rms <= c.nonlinear_tolerance * std::max(1., initial_rms)) {
                        converged = true;
                        break;
                    }
                    if (it == limit) break;
                    const double allowance = 128 * std::numeric_limits<double>::epsilon() *
                                             ((state_norm(state.p) + old_norm) / tau + norm);
                    const double target = std::max(10 * c.linear_tolerance * norm, allowance);
                    int its = 0;
                    double rel = 0;
                    const double *equation_error = res.p;  // DG: -(rhs - J x), same norm
                    if (!hdg) {
                        refresh(state.p);
                        k1_axpby<<<vec_grid(), 256, 0, stream>>>(rhs.p, -1.0, residual.p, 0.0, residual.p, D);
                        bj_refresh();
                        gmres(rhs.p, direction.p, c.max_linear, c.linear_tolerance, its, rel);
                    } else {
                        // time_integrator.cpp:125-153: projected condensation, then up to four
                        // Schur solves, each tightening ltol by 100 until the recovered DG
                        // residual meets the target
                        hdg_refresh(state.p, residual.p);
                        double tolerance = c.linear_tolerance;
                        bool valid = false;
                        for (int attempt = 0; attempt < 4; ++attempt) {
                            int cur = 0;
                            hdg_gmres(trhs.p, tdir.p, c.max_linear, tolerance, cur, rel);
                            its += cur;
                            hdg_recover(tdir.p, direction.p);
                            hdg_recovered(tdir.p, scratch.p);
                            valid = dual_norm(scratch.p) <= target;
                            if (valid) break;
                            tolerance *= .01;
                        }
                        require1(valid, "Recovered DG residual criterion failed after Schur refinement");
                        equation_error = scratch.p;
                    }
                    const double error_norm = dual_norm(equation_error);
                    require1(error_norm <= target, "DG Newton linear residual too large");
                    double damping = 1.;
                    bool accepted = false;
                    for (int search = 0; search < 30; ++search) {
                        k1_axpy<<<vec_grid(), 256, 0, stream>>>(trial.p, state.p, damping, direction.p, D);
                        if (admissible(trial.p)) {
                            residual_of(trial.p, trial_residual.p);
                            if (dual_norm(trial_residual.p) <= (1 - 1e-4 * damping) * norm + allowance) {
                                accepted = true;
                                break;
                            }
                        }
                        damping *= .5;
                    }
                    require1(accepted, "Newton line search failed");
                    std::swap(state, trial);
                    std::swap(residual, trial_residual);
                    residual_ready = true;
                    const double duration = seconds_since(update_start);
                    stats.max_linear_residual = std::max(stats.max_linear_residual, rel);
                    const double dgrel = norm > 0 ? error_norm / norm : 0.;
                    stats.max_dg_linear_residual = std::max(stats.max_dg_linear_residual, dgrel);
                    stats.krylov_iterations += its;
                    if (stats.newton_iterations >= c.warmup_updates) {
                        stats.measured_newton_seconds += duration;
                        ++stats.measured_updates;
                    }
                    ++stats.newton_iterations;
                    dgh_update_stats u{};
                    u.step = stats.steps;
                    u.stage = stage;
                    u.iteration = it;
                    u.krylov_iterations = its;
                    u.time = stats.final_time + step;
                    u.stage_time = stats.final_time + (stages == 1 ? 1. : stage_c[stage]) * step;
                    u.dt = step;
                    u.cfl = step * rate;
                    u.residual_physical_rms = rms;
                    u.damping = damping;
                    u.linear_relative_residual = rel;
                    u.dg_linear_relative_residual = dgrel;
                    u.iteration_s = duration;
                    u.inner_iterations = 0;
                    u.outer_iterations = 0;
                    u.precision_error = -1;
                    updates.push_back(u);
                }
                require1(converged || c.newton_steps > 0, "Newton iteration limit reached");
                k1_diffdiv<<<vec_grid(), 256, 0, stream>>>(slope(stage), state.p, stage_base.p, tau, D);
            }
            if (stages == 1) copy(solution, state.p);
            else {
                copy(solution, step_state.p);
                for (int stage = 0; stage < stages; ++stage)
                    k1_axpy<<<vec_grid(), 256, 0, stream>>>(solution, solution, step * B[stage], slope(stage), D);
                require1(admissible(solution), "DIRK3 step produced an inadmissible state");
            }
            stats.final_time =
                c.steps == 0 && step >= c.final_time - stats.final_time ? c.final_time : stats.final_time + step;
            stats.minimum_step = stats.steps == 0 ? step : std::min(stats.minimum_step, step);
            stats.maximum_step = std::max(stats.maximum_step, step);
            ++stats.steps;
            double now[3];
            integrals(solution, 0, 0, now);
            double drift = 0;
            for (int v = 0; v < 3; ++v) drift = std::max(drift, std::abs(now[v] - initial_integral[v]));
            stats.conservation_drift = std::max(stats.conservation_drift, drift);
        }
        sync();
        stats.seconds = seconds_since(begin);
    }
};

namespace {
thread_local std::string g_create_error1;
template <class F> int run1(dgh1_ctx *ctx, F &&f) {
    if (!ctx) return DGH_ERROR;
    try {
        f();
        return DGH_OK;
    } catch (const std::exception &e) {
        ctx->error = e.what();
        return DGH_ERROR;
    }
}
}  // namespace

extern "C" {
int dgh1_parse_config(int argc, const char *const *argv, dgh1_config *out) {
    try {
        dgh1_config c{};
        c.cells = 32, c.degree = 3, c.quadrature = 0;
        c.restart = 30, c.max_linear = 1500, c.max_newton = 30, c.steps = 0;
        c.newton_steps = 0, c.warmup_updates = 0;
        c.cfl = 2.0, c.final_time = 0.2, c.safety = 1.1;
        c.nonlinear_tolerance = 1e-12, c.linear_tolerance = 1e-12;
        c.integrator = 1;
        c.method = DGH_METHOD_BOTH;
        c.timing_only = 0;
        std::snprintf(c.prefix, sizeof(c.prefix), "%s", "output/run");
        bool time_seen = false, steps_seen = false;
        for (int i = 0; i < argc; ++i) {
            const std::string o = argv[i];
            auto value = [&]() {
                require1(i + 1 < argc, "Missing option value");
                return std::string(argv[++i]);
            };
            auto integer = [&]() {
                const std::string t = value();
                size_t end = 0;
                const int v = std::stoi(t, &end);
                require1(end == t.size(), "Invalid integer");
                return v;
            };
            auto real = [&]() {
                const std::string t = value();
                size_t end = 0;
                const double v = std::stod(t, &end);
                require1(end == t.size(), "Invalid number");
                return v;
            };
            if (o == "--cells") c.cells = integer();
            else if (o == "--newton-steps") c.newton_steps = integer();
            else if (o == "--warmup-updates") c.warmup_updates = integer();
            else if (o == "--integrator") {
                const std::string v = value();
                require1(v == "be" || v == "ssp-dirk3", "Invalid integrator");
                c.integrator = v == "be" ? 0 : 1;
            } else if (o == "--timing-only") c.timing_only = 1;
            else if (o == "--check-update") {
            } else if (o == "--k") c.degree = integer();
            else if (o == "--quad") c.quadrature = integer();
            else if (o == "--cfl") c.cfl = real();
            else if (o == "--T") {
                c.final_time = real();
                time_seen = true;
            } else if (o == "--steps") {
                c.steps = integer();
                steps_seen = true;
            } else if (o == "--safety") c.safety = real();
            else if (o == "--method") {
                const std::string v = value();
                require1(v == "dg" || v == "hdg" || v == "both", "Method must be dg, hdg, or both");
                c.method = v == "dg" ? DGH_METHOD_DG : v == "hdg" ? DGH_METHOD_HDG : DGH_METHOD_BOTH;
            } else if (o == "--prefix") std::snprintf(c.prefix, sizeof(c.prefix), "%s", value().c_str());
            else if (o == "--ntol") c.nonlinear_tolerance = real();
            else if (o == "--ltol") c.linear_tolerance = real();
            else if (o == "--restart") c.restart = integer();
            else if (o == "--max-linear") c.max_linear = integer();
            else if (o == "--max-newton") c.max_newton = integer();
            else if (o == "--linear") {
                const std::string v = value();
                require1(v == "direct" || v == "gmres", "Invalid linear solver");
                require1(v == "gmres", "The 1D device path solves with GMRES: use --linear gmres");
            } else if (o == "--ordering") {
                const std::string v = value();
                require1(v == "natural" || v == "colamd", "Invalid ordering");
            } else if (o == "--precond") {
                const std::string v = value();
                require1(v == "block-jacobi" || v == "block-ilu0", "Invalid block preconditioner");
                require1(v == "block-jacobi", "The 1D device path preconditions with block Jacobi");
            } else if (o == "--device") {
                (void)value();
            } else if (o == "--testcase") {
                const std::string v = value();
                require1(v == "simple" || v == "density", "Test case must be simple or density");
                c.testcase = v == "density" ? 1 : 0;
            } else
                throw std::runtime_error("Unknown option: " + o);
        }
        require1(!(time_seen && steps_seen), "Choose either --T or --steps");
        require1(!steps_seen || c.steps > 0, "steps must be positive");
        require1(c.cells >= 3 && c.degree >= 0 && c.degree <= 8 && c.quadrature >= 0, "Invalid mesh or degree");
        require1(c.quadrature == 0 || c.quadrature >= c.degree + 1, "Quadrature must be >= degree+1");
        require1(c.restart > 0 && c.max_linear > 0 && c.max_newton > 0 && c.steps >= 0, "Invalid iteration limit");
        require1(c.newton_steps >= 0 && c.warmup_updates >= 0, "Invalid timing update count");
        require1(std::isfinite(c.cfl) && c.cfl > 0 && std::isfinite(c.final_time) && c.final_time >= 0,
                 "CFL must be positive and final time nonnegative");
        require1(std::isfinite(c.safety) && c.safety >= 1 && std::isfinite(c.nonlinear_tolerance) &&
                     c.nonlinear_tolerance > 0 && std::isfinite(c.linear_tolerance) && c.linear_tolerance > 0 &&
                     c.linear_tolerance < 1,
                 "Invalid solver tolerance or wave-speed safety");
        *out = c;
        return DGH_OK;
    } catch (const std::exception &e) {
        g_create_error1 = e.what();
        return DGH_ERROR;
    }
}
dgh1_ctx *dgh1_create(const dgh1_config *config) {
    try {
        require1(config != nullptr, "Null configuration");
        return new dgh1_ctx(*config);
    } catch (const std::exception &e) {
        g_create_error1 = e.what();
        return nullptr;
    }
}
const char *dgh1_create_error(void) { return g_create_error1.c_str(); }
void dgh1_destroy(dgh1_ctx *ctx) { delete ctx; }
const char *dgh1_last_error(const dgh1_ctx *ctx) { return ctx ? ctx->error.c_str() : "Null context"; }
long long dgh1_size(const dgh1_ctx *ctx) { return ctx ? ctx->D : 0; }
int dgh1_nodes(const dgh1_ctx *ctx, double *x) {
    if (!ctx) return DGH_ERROR;
    std::vector<double> nodes, weights;
    gauss_rule(ctx->n, nodes, weights);
    for (int i = 0; i < ctx->n; ++i) x[i] = nodes[i];
    return DGH_OK;
}
int dgh1_weights(const dgh1_ctx *ctx, double *w) {
    if (!ctx) return DGH_ERROR;
    for (int i = 0; i < ctx->n; ++i) w[i] = ctx->hline.w[i];
    return DGH_OK;
}
int dgh1_initial(dgh1_ctx *ctx, double *u) {
    return run1(ctx, [&] {
        ctx->initial(u);
        ctx->sync();
    });
}
int dgh1_freeze(dgh1_ctx *ctx, const double *u, double *rate, double *speeds) {
    return run1(ctx, [&] { *rate = ctx->freeze(u, speeds); });
}
int dgh1_set_step(dgh1_ctx *ctx, double tau, const double *old) {
    return run1(ctx, [&] {
        require1(std::isfinite(tau) && tau > 0, "Invalid step size");
        ctx->set_step(tau, old);
        ctx->sync();
    });
}
int dgh1_residual(dgh1_ctx *ctx, const double *u, double *r) {
    return run1(ctx, [&] {
        ctx->residual_of(u, r);
        ctx->sync();
    });
}
int dgh1_admissible(dgh1_ctx *ctx, const double *u, int *ok) {
    return run1(ctx, [&] { *ok = ctx->admissible(u) ? 1 : 0; });
}
int dgh1_refresh(dgh1_ctx *ctx, const double *u) {
    return run1(ctx, [&] {
        ctx->refresh(u);
        ctx->sync();
    });
}
int dgh1_apply(dgh1_ctx *ctx, const double *x, double *y) {
    return run1(ctx, [&] {
        ctx->apply(x, y);
        ctx->sync();
    });
}
int dgh1_block(dgh1_ctx *ctx, int row, double *out) { return run1(ctx, [&] { ctx->block(row, out); }); }
int dgh1_bj_refresh(dgh1_ctx *ctx) { return run1(ctx, [&] { ctx->bj_refresh(); }); }
int dgh1_bj_apply(dgh1_ctx *ctx, const double *r, double *z) {
    return run1(ctx, [&] {
        ctx->bj_apply(r, z);
        ctx->sync();
    });
}
int dgh1_gmres(dgh1_ctx *ctx, const double *rhs, double *x, int max_iterations, double tolerance, int *iterations,
               double *relative_residual) {
    return run1(ctx, [&] {
        int its = 0;
        double rel = 0;
        ctx->gmres(rhs, x, max_iterations, tolerance, its, rel);
        *iterations = its;
        *relative_residual = rel;
    });
}
int dgh1_error(dgh1_ctx *ctx, const double *u, double t, double *error3) {
    return run1(ctx, [&] {
        require1(t < ctx->shock_time(), "Simple wave reaches shock time");
        ctx->integrals(u, t, 1, error3);
    });
}
int dgh1_integral(dgh1_ctx *ctx, const double *u, double *integral3) {
    return run1(ctx, [&] { ctx->integrals(u, 0, 0, integral3); });
}
int dgh1_hdg_refresh(dgh1_ctx *ctx, const double *u, const double *r) {
    return run1(ctx, [&] { ctx->hdg_refresh(u, r); });
}
int dgh1_hdg_rhs(dgh1_ctx *ctx, double *rhs) {
    return run1(ctx, [&] {
        require1(ctx->hdg_ready, "Refresh the hybrid operator first");
        HIPCHECK1(hipMemcpyAsync(rhs, ctx->trhs.p, ctx->trace_size() * sizeof(double), hipMemcpyDeviceToDevice,
                                 ctx->stream));
        ctx->sync();
    });
}
int dgh1_hdg_apply(dgh1_ctx *ctx, const double *x, double *y) {
    return run1(ctx, [&] {
        ctx->hdg_apply(x, y);
        ctx->sync();
    });
}
int dgh1_hdg_block(dgh1_ctx *ctx, int face, double *out) { return run1(ctx, [&] { ctx->hdg_block(face, out); }); }
int dgh1_hdg_bj_apply(dgh1_ctx *ctx, const double *r, double *z) {
    return run1(ctx, [&] {
        ctx->hdg_bj_apply(r, z);
        ctx->sync();
    });
}
int dgh1_hdg_recover(dgh1_ctx *ctx, const double *dl, double *du) {
    return run1(ctx, [&] {
        ctx->hdg_recover(dl, du);
        ctx->sync();
    });
}
int dgh1_hdg_recovered_residual(dgh1_ctx *ctx, const double *dl, double *out) {
    return run1(ctx, [&] {
        ctx->hdg_recovered(dl, out);
        ctx->sync();
    });
}
int dgh1_hdg_gmres(dgh1_ctx *ctx, const double *rhs, double *x, int max_iterations, double tolerance,
                   int *iterations, double *relative_residual) {
    return run1(ctx, [&] {
        int its = 0;
        double rel = 0;
        ctx->hdg_gmres(rhs, x, max_iterations, tolerance, its, rel);
        *iterations = its;
        *relative_residual = rel;
    });
}
int dgh1_advance(dgh1_ctx *ctx, double *u, dgh_run_stats *run, dgh_update_stats *updates, int capacity,
                 int *count) {
    return run1(ctx, [&] {
        std::vector<dgh_update_stats> all;
        dgh_run_stats stats{};
        ctx->advance(u, stats, all);
        if (run) *run = stats;
        if (count) *count = static_cast<int>(all.size());
        if (updates)
            for (int k = 0; k < std::min<int>(capacity, static_cast<int>(all.size())); ++k) updates[k] = all[k];
    });
}
}
