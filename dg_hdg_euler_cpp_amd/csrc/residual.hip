// Element and face kernels of the DG residual, its linearisation (Jv and the
// adjoint J^T w), the frozen HLL speeds, admissibility and initial states.
//
// Layout (reference types.hpp:22-31): state index cell*4n^2 + v*n^2 + (i + n*j),
// i = x node (fastest). Face f < C is the +x face of cell f, f >= C the +y face
// of cell f-C (discretization.cpp:32-34). Per-element side traces are stored as
// traces[cell][side][v][q] (side 0 -x, 1 +x, 2 -y, 3 +y), face quantities as
// [face][v][q] (the reference's lambda layout, discretization.cpp:148).
#include "dgh_internal.hpp"

namespace dgh {
namespace {

constexpr int BLOCK = 256;

__device__ __forceinline__ int neighbor(int cell, int side, int nx, int ny) {
    int x = cell % nx, y = cell / nx;
    if (side < 2) x = (x + (side == 0 ? nx - 1 : 1)) % nx;
    else y = (y + (side == 2 ? ny - 1 : 1)) % ny;
    return x + nx * y;
}

struct Prim { double inv, u, v, p, H; };

// physics.cpp:7-14
__device__ __forceinline__ Prim primitive(const double s[4], double gamma) {
    Prim p;
    p.inv = 1 / s[0];
    p.u = s[1] * p.inv;
    p.v = s[2] * p.inv;
    p.p = (gamma - 1) * (s[3] - .5 * (s[1] * p.u + s[2] * p.v));
    p.H = (s[3] + p.p) * p.inv;
    return p;
}
// physics.cpp:15-22
__device__ __forceinline__ void flux(const double s[4], const Prim &p, int axis, double f[4]) {
    const double vel = axis == 0 ? p.u : p.v;
    f[0] = s[1 + axis];
    f[1] = s[1] * vel;
    f[2] = s[2] * vel;
    f[1 + axis] += p.p;
    f[3] = (s[3] + p.p) * vel;
}
// physics.cpp:4-6
__device__ __forceinline__ double pressure(const double s[4], double gamma) {
    return (gamma - 1) * (s[3] - .5 * (s[1] * s[1] + s[2] * s[2]) / s[0]);
}
// volume_jacobian.cpp:43-58 / face_jacobian.cpp:38-55: dF_axis from (u, v, H)
__device__ __forceinline__ void normal_action(double u, double v, double H, const double d[4], int axis,
                                              double gm, double f[4]) {
    const double dp = gm * (d[3] - u * d[1] - v * d[2] + .5 * (u * u + v * v) * d[0]);
    if (axis == 0) {
        f[0] = d[1];
        f[1] = 2 * u * d[1] - u * u * d[0] + dp;
        f[2] = u * d[2] + v * d[1] - u * v * d[0];
        f[3] = H * (d[1] - u * d[0]) + u * (d[3] + dp);
    } else {
        f[0] = d[2];
        f[1] = u * d[2] + v * d[1] - u * v * d[0];
        f[2] = 2 * v * d[2] - v * v * d[0] + dp;
        f[3] = H * (d[2] - v * d[0]) + v * (d[3] + dp);
    }
}
// Transpose of normal_action: g = (dF/dU)^T f, with dF/dU the cached Jacobian
// of physics.cpp:42-57.
__device__ __forceinline__ void normal_action_t(double u, double v, double H, const double f[4], int axis,
                                                double gm, double gamma, double g[4]) {
    const double s2 = u * u + v * v;
    if (axis == 0) {
        g[0] = (.5 * (gamma - 3) * u * u + .5 * gm * v * v) * f[1] - u * v * f[2] + u * (.5 * gm * s2 - H) * f[3];
        g[1] = f[0] + (3 - gamma) * u * f[1] + v * f[2] + (H - gm * u * u) * f[3];
        g[2] = -gm * v * f[1] + u * f[2] - gm * u * v * f[3];
        g[3] = gm * f[1] + gamma * u * f[3];
    } else {
        g[0] = -u * v * f[1] + (.5 * gm * u * u + .5 * (gamma - 3) * v * v) * f[2] + v * (.5 * gm * s2 - H) * f[3];
        g[1] = v * f[1] - gm * u * f[2] - gm * u * v * f[3];
        g[2] = f[0] + u * f[1] + (3 - gamma) * v * f[2] + (H - gm * v * v) * f[3];
        g[3] = gm * f[2] + gamma * v * f[3];
    }
}

template <int N> struct Shape {
    static constexpr int P = N * N, M = 4 * P;
    static constexpr int CPB = BLOCK / P > 0 ? BLOCK / P : 1;  // cells per 256-thread block
};

// ---------------------------------------------------------------- traces
// Restriction of every element to its four sides (discretization.cpp:50-60):
// x sides sum over i of line(i, q) * shape_i, y sides over j of line(q, j).
// One thread per (element, variable, q) reads line q in both directions and
// writes the four sides' value at q; no LDS, so every thread's 2n loads are
// in flight together (the LDS-staged form measured 10% slower at n = 9).
template <int N, typename TX>
__global__ __launch_bounds__(256) void k_restrict(const TX *__restrict__ x, double *__restrict__ tr, Geo g,
                                                  const Basis *__restrict__ B) {
    constexpr int M = 4 * N * N, P = N * N;
    const long long gid = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    if (gid >= static_cast<long long>(g.cells) * 4 * N) return;
    const long long cell = gid / (4 * N);
    const int v = static_cast<int>((gid / N) % 4), q = static_cast<int>(gid % N);
    const TX *sheet = x + cell * M + v * P;
    double row[N], col[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        row[i] = static_cast<double>(sheet[i + N * q]);
        col[i] = static_cast<double>(sheet[q + N * i]);
    }
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        s0 += row[i] * B->left[i];
        s1 += row[i] * B->right[i];
        s2 += col[i] * B->left[i];
        s3 += col[i] * B->right[i];
    }
    double *out = tr + cell * 16 * N + v * N + q;
    out[0] = s0;
    out[4 * N] = s1;
    out[8 * N] = s2;
    out[12 * N] = s3;
}

// ----------------------------------------------------------- frozen speeds
// discretization.cpp:163-173 + physics.cpp:58-70 (Einfeldt / Roe HLL speeds).
template <int N>
__global__ void k_freeze_speeds(const double *__restrict__ tr, double *__restrict__ speeds, int *flag, Geo g) {
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= 2LL * g.cells * N) return;
    const int face = static_cast<int>(gid / N), q = static_cast<int>(gid % N);
    const int axis = face >= g.cells, cell = face - axis * g.cells;
    const int next = neighbor(cell, 2 * axis + 1, g.nx, g.ny);
    double l[4], r[4];
    for (int v = 0; v < 4; ++v) {
        l[v] = tr[static_cast<long long>(cell) * 16 * N + (2 * axis + 1) * 4 * N + v * N + q];
        r[v] = tr[static_cast<long long>(next) * 16 * N + (2 * axis) * 4 * N + v * N + q];
    }
    const Prim pl = primitive(l, g.gamma), pr = primitive(r, g.gamma);
    if (!(l[0] > 0 && r[0] > 0 && pl.p > 0 && pr.p > 0)) {
        atomicOr(flag, 1);
        speeds[2 * gid] = 0;
        speeds[2 * gid + 1] = 0;
        return;
    }
    const double sl = sqrt(l[0]), sr = sqrt(r[0]);
    const double vu = (sl * pl.u + sr * pr.u) / (sl + sr), vv = (sl * pl.v + sr * pr.v) / (sl + sr);
    const double H = (sl * pl.H + sr * pr.H) / (sl + sr);
    const double sound = sqrt((g.gamma - 1) * (H - .5 * (vu * vu + vv * vv)));
    const double cl = sqrt(g.gamma * pl.p * pl.inv), cr = sqrt(g.gamma * pr.p * pr.inv);
    const double vl = axis ? pl.v : pl.u, vr = axis ? pr.v : pr.u, vroe = axis ? vv : vu;
    speeds[2 * gid] = g.safety * fmax(fmax(0., -vl + cl), -vroe + sound);
    speeds[2 * gid + 1] = g.safety * fmax(fmax(0., vr + cr), vroe + sound);
}

// discretization.cpp:174-194: rate = max_e(alpha_x/hx + alpha_y/hy), alpha from
// the element's face speeds and |u|+c at its volume points.
template <int N>
__global__ __launch_bounds__(BLOCK) void k_freeze_rate(const double *__restrict__ u,
                                                       const double *__restrict__ speeds,
                                                       unsigned long long *rate_bits, int *flag, Geo g) {
    using S = Shape<N>;
    __shared__ unsigned long long alpha[S::CPB][2];
    const int local = threadIdx.x / S::P, node = threadIdx.x % S::P;
    const long long cell = static_cast<long long>(blockIdx.x) * S::CPB + local;
    const bool active = local < S::CPB && cell < g.cells;
    for (int k = threadIdx.x; k < S::CPB * 2; k += BLOCK) alpha[k / 2][k % 2] = 0;
    __syncthreads();
    if (active) {
        double s[4];
        for (int v = 0; v < 4; ++v) s[v] = u[cell * S::M + v * S::P + node];
        const Prim p = primitive(s, g.gamma);
        if (!(s[0] > 0 && p.p > 0)) atomicOr(flag, 2);
        const double sound = sqrt(g.gamma * p.p * p.inv);
        atomicMax(&alpha[local][0], static_cast<unsigned long long>(__double_as_longlong(fabs(p.u) + sound)));
        atomicMax(&alpha[local][1], static_cast<unsigned long long>(__double_as_longlong(fabs(p.v) + sound)));
        for (int t = node; t < 4 * N; t += S::P) {
            const int side = t / N, q = t % N, c = static_cast<int>(cell);
            const int face = side == 0 ? neighbor(c, 0, g.nx, g.ny)
                           : side == 1 ? c
                           : side == 2 ? g.cells + neighbor(c, 2, g.nx, g.ny)
                                       : g.cells + c;
            const double a = speeds[2LL * (static_cast<long long>(face) * N + q)];
            const double b = speeds[2LL * (static_cast<long long>(face) * N + q) + 1];
            atomicMax(&alpha[local][side / 2], static_cast<unsigned long long>(__double_as_longlong(fmax(a, b))));
        }
    }
    __syncthreads();
    if (active && node == 0) {
        const double ax = __longlong_as_double(static_cast<long long>(alpha[local][0]));
        const double ay = __longlong_as_double(static_cast<long long>(alpha[local][1]));
        const double rate = ax / g.hx + ay / g.hy;
        atomicMax(rate_bits, static_cast<unsigned long long>(__double_as_longlong(rate)));
    }
}

// ------------------------------------------------------------ face fluxes
// discretization.cpp:133-135: weighted HLL flux per face node.
template <int N>
__global__ void k_residual_flux(const double *__restrict__ tr, const double *__restrict__ speeds,
                                double *__restrict__ out, const Basis *__restrict__ B, Geo g) {
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= 2LL * g.cells * N) return;
    const int face = static_cast<int>(gid / N), q = static_cast<int>(gid % N);
    const int axis = face >= g.cells, cell = face - axis * g.cells;
    const int next = neighbor(cell, 2 * axis + 1, g.nx, g.ny);
    double l[4], r[4], fl[4], fr[4];
    for (int v = 0; v < 4; ++v) {
        l[v] = tr[static_cast<long long>(cell) * 16 * N + (2 * axis + 1) * 4 * N + v * N + q];
        r[v] = tr[static_cast<long long>(next) * 16 * N + (2 * axis) * 4 * N + v * N + q];
    }
    flux(l, primitive(l, g.gamma), axis, fl);
    flux(r, primitive(r, g.gamma), axis, fr);
    const double a = speeds[2 * gid], b = speeds[2 * gid + 1];
    const double weight = .5 * (axis == 0 ? g.hy : g.hx) * B->w[q];
    const double scale = weight / (a + b);
    for (int v = 0; v < 4; ++v) out[static_cast<long long>(face) * 4 * N + v * N + q] =
        scale * (b * fl[v] + a * fr[v] + a * b * (l[v] - r[v]));
}

// Lift of the four face quantities of `cell` into node (i, j), variable v
// (discretization.cpp:61-71, 183-184): +shape on the element's own +x/+y faces,
// -shape on the faces it shares with its -x/-y neighbours.
template <int N>
__device__ __forceinline__ double lift_faces(const double *__restrict__ fq, const Basis *__restrict__ B, int cell,
                                             int i, int j, int v, const Geo &g, double acc) {
    const long long f0 = neighbor(cell, 0, g.nx, g.ny), f1 = cell;
    const long long f2 = g.cells + neighbor(cell, 2, g.nx, g.ny), f3 = g.cells + cell;
    acc += (-1 * B->left[i]) * fq[f0 * 4 * N + v * N + j];
    acc += (1 * B->right[i]) * fq[f1 * 4 * N + v * N + j];
    acc += (-1 * fq[f2 * 4 * N + v * N + i]) * B->left[j];
    acc += (1 * fq[f3 * 4 * N + v * N + i]) * B->right[j];
    return acc;
}

// The element kernels read the reference-element tables from LDS (staged once
// per block) and load their 16 face values at the start, so that neither waits
// on a dependent global load after the block's barrier (measured: the element
// Jv kernel 160 -> 129 us at 256^2, k=8).
template <int N> struct ElementTables {
    double wd[N * N], w[N], left[N], right[N];
};
template <int N>
__device__ __forceinline__ void stage_tables(const Basis *__restrict__ B, ElementTables<N> &t) {
    for (int k = threadIdx.x; k < N * N + 3 * N; k += BLOCK) {
        if (k < N * N) t.wd[k] = B->wd[k];
        else if (k < N * N + N) t.w[k - N * N] = B->w[k - N * N];
        else if (k < N * N + 2 * N) t.left[k - N * N - N] = B->left[k - N * N - N];
        else t.right[k - N * N - 2 * N] = B->right[k - N * N - 2 * N];
    }
}
// lift_faces with preloaded face values and staged tables (same operation order)
template <int N>
__device__ __forceinline__ double lift_staged(const double (&f)[4], const ElementTables<N> &t, int i, int j,
                                              double acc) {
    acc += (-1 * t.left[i]) * f[0];
    acc += (1 * t.right[i]) * f[1];
    acc += (-1 * f[2]) * t.left[j];
    acc += (1 * f[3]) * t.right[j];
    return acc;
}
// - (hy/2) w_j sum_q d(q,i) X(q,j) - (hx/2) w_i sum_q Y(i,q) d(q,j) for one variable
template <int N>
__device__ __forceinline__ double weak_volume(const double *X, const double *Y, const ElementTables<N> &t, int i,
                                              int j, const Geo &g, double acc) {
    const double sx = .5 * g.hy, sy = .5 * g.hx;
    double wx = 0, wy = 0;
#pragma unroll
    for (int q = 0; q < N; ++q) wx += t.wd[q + N * i] * X[q + N * j];
#pragma unroll
    for (int q = 0; q < N; ++q) wy += Y[i + N * q] * t.wd[q + N * j];
    acc -= (sx * t.w[j]) * wx;
    acc -= sy * t.w[i] * wy;
    return acc;
}

// Stage the block's face values once, coalesced, as [local cell][slot][v][q]
// (4N values per slot): slot s reads 4N consecutive values of face face_of(c, s)
// at offset off(s) * 4N in a face-major array of stride 4N * per_face. One
// copy per cell replaces the 16 redundant per-node loads and their registers
// (measured: element Jv 214 -> 199 us at 256^2, k=8; 117 -> 60 VGPRs).
template <int N, class FaceOf>
__device__ __forceinline__ void stage_faces(const double *__restrict__ src, int per_face, double *dst,
                                            const Geo &g, FaceOf face_of) {
    using S = Shape<N>;
    for (int t = threadIdx.x; t < S::CPB * 16 * N; t += BLOCK) {
        const long long c = static_cast<long long>(blockIdx.x) * S::CPB + t / (16 * N);
        if (c >= g.cells) continue;
        const int slot = (t / (4 * N)) % 4, vq = t % (4 * N);
        int off = 0;
        const long long face = face_of(static_cast<int>(c), slot, off);
        dst[t] = src[(face * per_face + off) * 4 * N + vq];
    }
}
// the DG face fluxes of a cell's sides 0..3 in a [face][v][q] array
struct SideFaces {
    Geo g;
    __device__ long long operator()(int c, int side, int &off) const {
        off = 0;
        return side == 0 ? neighbor(c, 0, g.nx, g.ny)
             : side == 1 ? c
             : side == 2 ? g.cells + neighbor(c, 2, g.nx, g.ny)
                         : g.cells + static_cast<long long>(c);
    }
};
// the face adjoint values of a cell in a [face][2][v][q] array: the left-side
// half of its +x and +y faces, the right-side half of its -x and -y faces
struct AdjointSides {
    Geo g;
    __device__ long long operator()(int c, int slot, int &off) const {
        off = slot == 1 || slot == 3 ? 1 : 0;
        return slot == 0 ? static_cast<long long>(c)
             : slot == 1 ? neighbor(c, 0, g.nx, g.ny)
             : slot == 2 ? g.cells + static_cast<long long>(c)
                         : g.cells + neighbor(c, 2, g.nx, g.ny);
    }
};

// ------------------------------------------------------- element residual
// discretization.cpp:111-140 + tensor_product.cpp:23-38:
// r = M(u-u_b)/tau - (hy/2) w_j sum_q d(q,i) Fx(q,j) - (hx/2) w_i sum_q Fy(i,q) d(q,j) + lifts.
template <int N>
__global__ __launch_bounds__(BLOCK) void k_element_residual(const double *__restrict__ u,
                                                            const double *__restrict__ old,
                                                            const double *__restrict__ fq, double *__restrict__ r,
                                                            double tau, const Basis *__restrict__ B, Geo g) {
    using S = Shape<N>;
    __shared__ double sfx[S::CPB * S::M], sfy[S::CPB * S::M], sfq[S::CPB * 16 * N];
    __shared__ ElementTables<N> tab;
    stage_tables<N>(B, tab);
    stage_faces<N>(fq, 1, sfq, g, SideFaces{g});
    const int local = threadIdx.x / S::P, node = threadIdx.x % S::P;
    const long long cell = static_cast<long long>(blockIdx.x) * S::CPB + local;
    const bool active = local < S::CPB && cell < g.cells;
    const int i = node % N, j = node / N;
    double res[4];
    if (active) {
        double s[4];
        for (int v = 0; v < 4; ++v) {
            s[v] = u[cell * S::M + v * S::P + node];
            res[v] = B->mass[node] * (s[v] - old[cell * S::M + v * S::P + node]) / tau;
        }
        if (N > 1) {
            const Prim p = primitive(s, g.gamma);
            double fx[4], fy[4];
            flux(s, p, 0, fx);
            flux(s, p, 1, fy);
            for (int v = 0; v < 4; ++v) {
                sfx[local * S::M + v * S::P + node] = fx[v];
                sfy[local * S::M + v * S::P + node] = fy[v];
            }
        }
    }
    __syncthreads();
    if (!active) return;
    for (int v = 0; v < 4; ++v) {
        if (N > 1) res[v] = weak_volume<N>(sfx + local * S::M + v * S::P, sfy + local * S::M + v * S::P, tab, i, j, g,
                                           res[v]);
        const double *f = sfq + local * 16 * N + v * N;
        const double fv[4] = {f[j], f[4 * N + j], f[8 * N + i], f[12 * N + i]};
        r[cell * S::M + v * S::P + node] = lift_staged<N>(fv, tab, i, j, res[v]);
    }
}

// discretization.cpp:211-233: density and pressure > 1e-13 and finite at every
// node and at the four restricted face values of every element.
template <int N>
__global__ __launch_bounds__(BLOCK) void k_admissible(const double *__restrict__ u, int *flag,
                                                      const Basis *__restrict__ B, Geo g) {
    using S = Shape<N>;
    __shared__ double cellbuf[S::CPB * S::M];
    const int local = threadIdx.x / S::P, node = threadIdx.x % S::P;
    const long long cell = static_cast<long long>(blockIdx.x) * S::CPB + local;
    const bool active = local < S::CPB && cell < g.cells;
    bool ok = true;
    if (active) {
        double s[4];
        for (int v = 0; v < 4; ++v) {
            s[v] = u[cell * S::M + v * S::P + node];
            cellbuf[local * S::M + v * S::P + node] = s[v];
            ok = ok && isfinite(s[v]);
        }
        ok = ok && s[0] > 1e-13 && pressure(s, g.gamma) > 1e-13;
    }
    __syncthreads();
    for (int t = node; active && t < 4 * N; t += S::P) {
        const int side = t / N, q = t % N;
        const double *e = side % 2 ? B->right : B->left;
        double s[4];
        for (int v = 0; v < 4; ++v) {
            const double *sheet = cellbuf + local * S::M + v * S::P;
            double acc = 0;
#pragma unroll
            for (int k = 0; k < N; ++k) acc += (side < 2 ? sheet[k + N * q] : sheet[q + N * k]) * e[k];
            s[v] = acc;
            ok = ok && isfinite(acc);
        }
        ok = ok && s[0] > 1e-13 && pressure(s, g.gamma) > 1e-13;
    }
    if (active && !ok) atomicAnd(flag, 0);
}

// ------------------------------------------------------ linearisation caches
// volume_jacobian.cpp:13-20: (u, v, H) per volume node, [cell][3][n^2].
template <int N>
__global__ void k_volume_cache(const double *__restrict__ u, double *__restrict__ vc, Geo g) {
    constexpr int P = N * N, M = 4 * P;
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= static_cast<long long>(g.cells) * P) return;
    const long long cell = gid / P;
    const int node = static_cast<int>(gid % P);
    double s[4];
    for (int v = 0; v < 4; ++v) s[v] = u[cell * M + v * P + node];
    const Prim p = primitive(s, g.gamma);
    vc[cell * 3 * P + node] = p.u;
    vc[cell * 3 * P + P + node] = p.v;
    vc[cell * 3 * P + 2 * P + node] = p.H;
}

// face_jacobian.cpp:16-37: per face node (u, v, H) of both sides, HLL blends,
// penalty and the frozen speeds (HDG tau), structure of arrays.
template <int N>
__global__ void k_face_cache(const double *__restrict__ tr, const double *__restrict__ speeds,
                             double *__restrict__ fc, const Basis *__restrict__ B, Geo g) {
    const long long nodes = 2LL * g.cells * N;
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= nodes) return;
    const int face = static_cast<int>(gid / N), q = static_cast<int>(gid % N);
    const int axis = face >= g.cells, cell = face - axis * g.cells;
    const int next = neighbor(cell, 2 * axis + 1, g.nx, g.ny);
    double l[4], r[4];
    for (int v = 0; v < 4; ++v) {
        l[v] = tr[static_cast<long long>(cell) * 16 * N + (2 * axis + 1) * 4 * N + v * N + q];
        r[v] = tr[static_cast<long long>(next) * 16 * N + (2 * axis) * 4 * N + v * N + q];
    }
    const Prim pl = primitive(l, g.gamma), pr = primitive(r, g.gamma);
    const double a = speeds[2 * gid], b = speeds[2 * gid + 1];
    const double weight = .5 * (axis == 0 ? g.hy : g.hx) * B->w[q];
    fc[F_UL * nodes + gid] = pl.u;
    fc[F_VL * nodes + gid] = pl.v;
    fc[F_HL * nodes + gid] = pl.H;
    fc[F_UR * nodes + gid] = pr.u;
    fc[F_VR * nodes + gid] = pr.v;
    fc[F_HR * nodes + gid] = pr.H;
    fc[F_BL * nodes + gid] = weight * b / (a + b);
    fc[F_BR * nodes + gid] = weight * a / (a + b);
    fc[F_PEN * nodes + gid] = weight * a * b / (a + b);
    fc[F_TAUL * nodes + gid] = a;
    fc[F_TAUR * nodes + gid] = b;
}

// face_jacobian.cpp:56-64: linearised HLL flux per face node.
template <int N>
__global__ void k_jv_flux(const double *__restrict__ tr, const double *__restrict__ fc, double *__restrict__ out,
                          Geo g) {
    const long long nodes = 2LL * g.cells * N;
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= nodes) return;
    const int face = static_cast<int>(gid / N), q = static_cast<int>(gid % N);
    const int axis = face >= g.cells, cell = face - axis * g.cells;
    const int next = neighbor(cell, 2 * axis + 1, g.nx, g.ny);
    double l[4], r[4], fl[4], fr[4];
    for (int v = 0; v < 4; ++v) {
        l[v] = tr[static_cast<long long>(cell) * 16 * N + (2 * axis + 1) * 4 * N + v * N + q];
        r[v] = tr[static_cast<long long>(next) * 16 * N + (2 * axis) * 4 * N + v * N + q];
    }
    const double gm = g.gamma - 1.;
    normal_action(fc[F_UL * nodes + gid], fc[F_VL * nodes + gid], fc[F_HL * nodes + gid], l, axis, gm, fl);
    normal_action(fc[F_UR * nodes + gid], fc[F_VR * nodes + gid], fc[F_HR * nodes + gid], r, axis, gm, fr);
    const double bl = fc[F_BL * nodes + gid], br = fc[F_BR * nodes + gid], pen = fc[F_PEN * nodes + gid];
    for (int v = 0; v < 4; ++v)
        out[static_cast<long long>(face) * 4 * N + v * N + q] = bl * fl[v] + br * fr[v] + pen * (l[v] - r[v]);
}

// volume_jacobian.cpp:59-71 + dg_operator.cpp:10-21: y = M x / tau - weak(dF x) + face lifts.
template <int N, typename TX, typename TY>
__global__ __launch_bounds__(BLOCK) void k_element_jv(const TX *__restrict__ x, const double *__restrict__ vc,
                                                      const double *__restrict__ fq, TY *__restrict__ y,
                                                      double inv_step, const Basis *__restrict__ B, Geo g) {
    using S = Shape<N>;
    __shared__ double sfx[S::CPB * S::M], sfy[S::CPB * S::M];
    __shared__ double sfq[S::CPB * 16 * N];
    __shared__ ElementTables<N> tab;
    stage_tables<N>(B, tab);
    const int local = threadIdx.x / S::P, node = threadIdx.x % S::P;
    const long long cell = static_cast<long long>(blockIdx.x) * S::CPB + local;
    const bool active = local < S::CPB && cell < g.cells;
    const int i = node % N, j = node / N;
    double res[4];
    stage_faces<N>(fq, 1, sfq, g, SideFaces{g});
    if (active) {
        double d[4];
        for (int v = 0; v < 4; ++v) {
            d[v] = static_cast<double>(x[cell * S::M + v * S::P + node]);
            res[v] = B->mass[node] * d[v] * inv_step;
        }
        if (N > 1) {
            const double uu = vc[cell * 3 * S::P + node], vv = vc[cell * 3 * S::P + S::P + node];
            const double H = vc[cell * 3 * S::P + 2 * S::P + node];
            double fx[4], fy[4];
            normal_action(uu, vv, H, d, 0, g.gamma - 1., fx);
            normal_action(uu, vv, H, d, 1, g.gamma - 1., fy);
            for (int v = 0; v < 4; ++v) {
                sfx[local * S::M + v * S::P + node] = fx[v];
                sfy[local * S::M + v * S::P + node] = fy[v];
            }
        }
    }
    __syncthreads();
    if (!active) return;
    for (int v = 0; v < 4; ++v) {
        if (N > 1) res[v] = weak_volume<N>(sfx + local * S::M + v * S::P, sfy + local * S::M + v * S::P, tab, i, j, g,
                                           res[v]);
        const double *f = sfq + local * 16 * N + v * N;
        const double fv[4] = {f[j], f[4 * N + j], f[8 * N + i], f[12 * N + i]};
        y[cell * S::M + v * S::P + node] = static_cast<TY>(lift_staged<N>(fv, tab, i, j, res[v]));
    }
}

// ------------------------------------------------------------- adjoint J^T w
// The face part of J maps side traces (restrictions R_s x) through the HLL
// derivative and lifts with the same shapes (lift = R_s^T), so its transpose
// restricts w to the sides (k_restrict), applies the transposed face flux
// derivative per face node and lifts the result back (lift_faces).
// Per face node with w_L = R_{left side} w (scale +1) and w_R = R_{right side} w
// (scale -1): flux' = bl A_L x_L + br A_R x_R + pen (x_L - x_R), so
// g_L = bl A_L^T (w_L - w_R) + pen (w_L - w_R) and
// g_R = br A_R^T (w_L - w_R) - pen (w_L - w_R).
template <int N>
__global__ void k_adjoint_face(const double *__restrict__ wtr, const double *__restrict__ fc,
                               double *__restrict__ gtr, Geo g) {
    const long long nodes = 2LL * g.cells * N;
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= nodes) return;
    const int face = static_cast<int>(gid / N), q = static_cast<int>(gid % N);
    const int axis = face >= g.cells, cell = face - axis * g.cells;
    const int next = neighbor(cell, 2 * axis + 1, g.nx, g.ny);
    double dw[4], gl[4], gr[4];
    for (int v = 0; v < 4; ++v)
        dw[v] = wtr[static_cast<long long>(cell) * 16 * N + (2 * axis + 1) * 4 * N + v * N + q] -
                wtr[static_cast<long long>(next) * 16 * N + (2 * axis) * 4 * N + v * N + q];
    const double gm = g.gamma - 1.;
    normal_action_t(fc[F_UL * nodes + gid], fc[F_VL * nodes + gid], fc[F_HL * nodes + gid], dw, axis, gm, g.gamma, gl);
    normal_action_t(fc[F_UR * nodes + gid], fc[F_VR * nodes + gid], fc[F_HR * nodes + gid], dw, axis, gm, g.gamma, gr);
    const double bl = fc[F_BL * nodes + gid], br = fc[F_BR * nodes + gid], pen = fc[F_PEN * nodes + gid];
    // Left element (+side) receives +g_L lifted with the +side shape; the right
    // element's -side shape lift carries g_R. Store as [face][v][q] pairs.
    for (int v = 0; v < 4; ++v) {
        gtr[static_cast<long long>(face) * 8 * N + v * N + q] = bl * gl[v] + pen * dw[v];
        gtr[static_cast<long long>(face) * 8 * N + 4 * N + v * N + q] = br * gr[v] - pen * dw[v];
    }
}

// Volume transpose: y_vol = M x/tau - sum_axis D_axis^T-contractions of A(x), so
// J_vol^T w = M w/tau - A_x^T (D-contraction^T w) - A_y^T (...). With
// r(i,j) -= sx w_j sum_q d(q,i) Fx(q,j), the transpose maps w to
// Gx(q,j) = sx w_j sum_i d(q,i) w(i,j) and Gy(i,q) = sy w_i sum_j d(q,j) w(i,j),
// and z(node) = M w/tau - A_x(node)^T Gx(node) - A_y(node)^T Gy(node).
template <int N>
__global__ __launch_bounds__(BLOCK) void k_element_jtv(const double *__restrict__ w, const double *__restrict__ vc,
                                                       const double *__restrict__ gtr, double *__restrict__ z,
                                                       double inv_step, const Basis *__restrict__ B, Geo g) {
    using S = Shape<N>;
    __shared__ double sw[S::CPB * S::M];
    __shared__ ElementTables<N> tab;
    stage_tables<N>(B, tab);
    const int local = threadIdx.x / S::P, node = threadIdx.x % S::P;
    const long long cell = static_cast<long long>(blockIdx.x) * S::CPB + local;
    const bool active = local < S::CPB && cell < g.cells;
    const int i = node % N, j = node / N;
    // Face adjoint lifts: the element is the left element of its +x and +y
    // faces and the right element of its -x and -y faces (AdjointSides).
    __shared__ double sgt[S::CPB * 16 * N];
    stage_faces<N>(gtr, 2, sgt, g, AdjointSides{g});
    double wv[4], res[4], uu = 0, vv = 0, H = 0;
    if (active && N > 1) {
        uu = vc[cell * 3 * S::P + node];
        vv = vc[cell * 3 * S::P + S::P + node];
        H = vc[cell * 3 * S::P + 2 * S::P + node];
    }
    if (active)
        for (int v = 0; v < 4; ++v) {
            wv[v] = w[cell * S::M + v * S::P + node];
            sw[local * S::M + v * S::P + node] = wv[v];
            res[v] = B->mass[node] * wv[v] * inv_step;
        }
    __syncthreads();
    if (!active) return;
    if (N > 1) {
        const double sx = .5 * g.hy, sy = .5 * g.hx;
        double gx[4], gy[4];
        for (int v = 0; v < 4; ++v) {
            const double *W = sw + local * S::M + v * S::P;
            double ax = 0, ay = 0;
#pragma unroll
            for (int k = 0; k < N; ++k) ax += tab.wd[i + N * k] * W[k + N * j];
#pragma unroll
            for (int k = 0; k < N; ++k) ay += tab.wd[j + N * k] * W[i + N * k];
            gx[v] = sx * tab.w[j] * ax;
            gy[v] = sy * tab.w[i] * ay;
        }
        double tx[4], ty[4];
        normal_action_t(uu, vv, H, gx, 0, g.gamma - 1., g.gamma, tx);
        normal_action_t(uu, vv, H, gy, 1, g.gamma - 1., g.gamma, ty);
        for (int v = 0; v < 4; ++v) res[v] -= tx[v] + ty[v];
    }
    for (int v = 0; v < 4; ++v) {
        const double *gl = sgt + local * 16 * N + v * N;
        double acc = res[v];
        acc += tab.right[i] * gl[j];
        acc += tab.left[i] * gl[4 * N + j];
        acc += tab.right[j] * gl[8 * N + i];
        acc += tab.left[j] * gl[12 * N + i];
        z[cell * S::M + v * S::P + node] = acc;
    }
}

// --------------------------------------------------------- over-integration
// Volume terms on the Q-point rule (quad Q > n): interpolation to the points
// (tensor_product.cpp:11-22), fluxes and their linearisation there, and the
// weak form back to the nodes (tensor_product.cpp:54-72); the face terms are
// those of the collocated kernels. One element per block; per-point work is
// spread over the block's threads with the reference's summation orders.
template <int N> struct OI {
    static constexpr int P = N * N, M = 4 * P, ITEMS = (M + BLOCK - 1) / BLOCK;
};

// nodal [4][n^2] -> points [4][Q^2], through tmp [4][Q n]
template <int N>
__device__ void oi_interpolate(const double *nodal, double *tmp, double *pts, const Basis *__restrict__ B, int Q) {
    for (int t = threadIdx.x; t < 4 * Q * N; t += BLOCK) {
        const int v = t / (Q * N), r = t % (Q * N), q = r % Q, j = r / Q;
        const double *in = nodal + v * N * N;
        double s = 0;
        for (int k = 0; k < N; ++k) s += in[k + N * j] * B->vq[q + Q * k];
        tmp[v * Q * N + q + Q * j] = s;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 4 * Q * Q; t += BLOCK) {
        const int v = t / (Q * Q), r = t % (Q * Q), qa = r % Q, qb = r / Q;
        double s = 0;
        for (int k = 0; k < N; ++k) s += B->vq[qb + Q * k] * tmp[v * Q * N + qa + Q * k];
        pts[v * Q * Q + qa + Q * qb] = s;
    }
    __syncthreads();
}

// first pass of the weak form: scr_x[v][i + n qb] = sum_k X[v][k + Q qb] d(k, i),
// scr_y[v][i + n qb] = sum_k Y[v][k + Q qb] (qw_k vq(k, i))
template <int N>
__device__ void oi_weak_pass1(const double *X, const double *Y, double *scr_x, double *scr_y,
                              const Basis *__restrict__ B, int Q) {
    for (int t = threadIdx.x; t < 4 * N * Q; t += BLOCK) {
        const int v = t / (N * Q), r = t % (N * Q), i = r % N, qb = r / N;
        const double *x = X + v * Q * Q, *y = Y + v * Q * Q;
        double sx = 0, sy = 0;
        for (int k = 0; k < Q; ++k) sx += x[k + Q * qb] * B->dq[k + Q * i];
        for (int k = 0; k < Q; ++k) sy += y[k + Q * qb] * (B->qw[k] * B->vq[k + Q * i]);
        scr_x[v * N * Q + i + N * qb] = sx;
        scr_y[v * N * Q + i + N * qb] = sy;
    }
    __syncthreads();
}

// second pass for node (i, j) of variable v: res -= sx * sum_k (qw_k vq(k, j)) scr_x,
// then res -= sy * sum_k d(k, j) scr_y
template <int N>
__device__ __forceinline__ double oi_weak_pass2(const double *scr_x, const double *scr_y, int v, int i, int j,
                                                double res, const Basis *__restrict__ B, const Geo &g, int Q) {
    double sx = 0, sy = 0;
    for (int k = 0; k < Q; ++k) sx += (B->qw[k] * B->vq[k + Q * j]) * scr_x[v * N * Q + i + N * k];
    res -= (.5 * g.hy) * sx;
    for (int k = 0; k < Q; ++k) sy += B->dq[k + Q * j] * scr_y[v * N * Q + i + N * k];
    res -= (.5 * g.hx) * sy;
    return res;
}

#define DGH_OI_SHARED                                                                            \
    __shared__ double nodal[OI<N>::M], tmp[4 * MAXQ * N], pts[4 * MAXQ * MAXQ];                   \
    __shared__ double fxs[4 * MAXQ * MAXQ], fys[4 * MAXQ * MAXQ], scr_x[4 * N * MAXQ], scr_y[4 * N * MAXQ];

// discretization.cpp:111-140 with tensor_product.cpp:54-72
template <int N>
__global__ __launch_bounds__(BLOCK) void k_oi_element_residual(const double *__restrict__ u,
                                                               const double *__restrict__ old,
                                                               const double *__restrict__ fq, double *__restrict__ r,
                                                               double tau, const Basis *__restrict__ B, Geo g) {
    using O = OI<N>;
    DGH_OI_SHARED
    const int cell = blockIdx.x, Q = g.nq, QQ = Q * Q;
    const long long base = static_cast<long long>(cell) * O::M;
    for (int t = threadIdx.x; t < O::M; t += BLOCK) nodal[t] = u[base + t];
    __syncthreads();
    double res[O::ITEMS];
#pragma unroll
    for (int it = 0; it < O::ITEMS; ++it) {
        const int item = threadIdx.x + it * BLOCK;
        if (item < O::M) res[it] = B->mass[item % O::P] * (nodal[item] - old[base + item]) / tau;
    }
    oi_interpolate<N>(nodal, tmp, pts, B, Q);
    for (int q = threadIdx.x; q < QQ; q += BLOCK) {
        double st[4], fx[4], fy[4];
        for (int v = 0; v < 4; ++v) st[v] = pts[v * QQ + q];
        const Prim p = primitive(st, g.gamma);
        flux(st, p, 0, fx);
        flux(st, p, 1, fy);
        for (int v = 0; v < 4; ++v) {
            fxs[v * QQ + q] = fx[v];
            fys[v * QQ + q] = fy[v];
        }
    }
    __syncthreads();
    oi_weak_pass1<N>(fxs, fys, scr_x, scr_y, B, Q);
#pragma unroll
    for (int it = 0; it < O::ITEMS; ++it) {
        const int item = threadIdx.x + it * BLOCK;
        if (item >= O::M) continue;
        const int v = item / O::P, node = item % O::P, i = node % N, j = node / N;
        double acc = oi_weak_pass2<N>(scr_x, scr_y, v, i, j, res[it], B, g, Q);
        acc = lift_faces<N>(fq, B, cell, i, j, v, g, acc);
        r[base + item] = acc;
    }
}

// volume_jacobian.cpp:13-20: (u, v, H) at the Q^2 points, [cell][3][Q^2]
template <int N>
__global__ __launch_bounds__(BLOCK) void k_oi_volume_cache(const double *__restrict__ u, double *__restrict__ vc,
                                                           const Basis *__restrict__ B, Geo g) {
    using O = OI<N>;
    __shared__ double nodal[O::M], tmp[4 * MAXQ * N], pts[4 * MAXQ * MAXQ];
    const int cell = blockIdx.x, Q = g.nq, QQ = Q * Q;
    for (int t = threadIdx.x; t < O::M; t += BLOCK) nodal[t] = u[static_cast<long long>(cell) * O::M + t];
    __syncthreads();
    oi_interpolate<N>(nodal, tmp, pts, B, Q);
    for (int q = threadIdx.x; q < QQ; q += BLOCK) {
        double st[4];
        for (int v = 0; v < 4; ++v) st[v] = pts[v * QQ + q];
        const Prim p = primitive(st, g.gamma);
        const long long vb = static_cast<long long>(cell) * 3 * QQ;
        vc[vb + q] = p.u;
        vc[vb + QQ + q] = p.v;
        vc[vb + 2 * QQ + q] = p.H;
    }
}

// volume_jacobian.cpp:43-71 on the points, plus the face lifts of k_element_jv
template <int N, typename TX, typename TY>
__global__ __launch_bounds__(BLOCK) void k_oi_element_jv(const TX *__restrict__ x, const double *__restrict__ vc,
                                                         const double *__restrict__ fq, TY *__restrict__ y,
                                                         double inv_step, const Basis *__restrict__ B, Geo g) {
    using O = OI<N>;
    DGH_OI_SHARED
    const int cell = blockIdx.x, Q = g.nq, QQ = Q * Q;
    const long long base = static_cast<long long>(cell) * O::M;
    for (int t = threadIdx.x; t < O::M; t += BLOCK) nodal[t] = static_cast<double>(x[base + t]);
    __syncthreads();
    double res[O::ITEMS];
#pragma unroll
    for (int it = 0; it < O::ITEMS; ++it) {
        const int item = threadIdx.x + it * BLOCK;
        if (item < O::M) res[it] = B->mass[item % O::P] * nodal[item] * inv_step;
    }
    oi_interpolate<N>(nodal, tmp, pts, B, Q);
    const double gm = g.gamma - 1.;
    for (int q = threadIdx.x; q < QQ; q += BLOCK) {
        double d[4], fx[4], fy[4];
        for (int v = 0; v < 4; ++v) d[v] = pts[v * QQ + q];
        const long long vb = static_cast<long long>(cell) * 3 * QQ;
        normal_action(vc[vb + q], vc[vb + QQ + q], vc[vb + 2 * QQ + q], d, 0, gm, fx);
        normal_action(vc[vb + q], vc[vb + QQ + q], vc[vb + 2 * QQ + q], d, 1, gm, fy);
        for (int v = 0; v < 4; ++v) {
            fxs[v * QQ + q] = fx[v];
            fys[v * QQ + q] = fy[v];
        }
    }
    __syncthreads();
    oi_weak_pass1<N>(fxs, fys, scr_x, scr_y, B, Q);
#pragma unroll
    for (int it = 0; it < O::ITEMS; ++it) {
        const int item = threadIdx.x + it * BLOCK;
        if (item >= O::M) continue;
        const int v = item / O::P, node = item % O::P, i = node % N, j = node / N;
        double acc = oi_weak_pass2<N>(scr_x, scr_y, v, i, j, res[it], B, g, Q);
        acc = lift_faces<N>(fq, B, cell, i, j, v, g, acc);
        y[base + item] = static_cast<TY>(acc);
    }
}

// Transpose of k_oi_element_jv's volume part: with the weak form
// r(i,j) -= sx sum d(k,i) qw_b vq(b,j) X(k,b) + sy sum qw_k vq(k,i) d(b,j) Y(k,b),
// w maps to GX(k,b) = sx qw_b sum_j vq(b,j) sum_i d(k,i) w(i,j) and
// GY(k,b) = sy qw_k sum_j d(b,j) sum_i vq(k,i) w(i,j); then A^T per point and
// the transposed interpolation back to the nodes. Face adjoint as k_element_jtv.
template <int N>
__global__ __launch_bounds__(BLOCK) void k_oi_element_jtv(const double *__restrict__ w, const double *__restrict__ vc,
                                                          const double *__restrict__ gtr, double *__restrict__ z,
                                                          double inv_step, const Basis *__restrict__ B, Geo g) {
    using O = OI<N>;
    DGH_OI_SHARED
    const int cell = blockIdx.x, Q = g.nq, QQ = Q * Q;
    const long long base = static_cast<long long>(cell) * O::M;
    for (int t = threadIdx.x; t < O::M; t += BLOCK) nodal[t] = w[base + t];
    __syncthreads();
    // T1[v][k + Q j] = sum_i d(k, i) w(i, j) (in scr_x), T2 = sum_i vq(k, i) w(i, j) (in scr_y)
    for (int t = threadIdx.x; t < 4 * Q * N; t += BLOCK) {
        const int v = t / (Q * N), r = t % (Q * N), k = r % Q, j = r / Q;
        const double *in = nodal + v * O::P;
        double a = 0, b = 0;
        for (int i = 0; i < N; ++i) {
            a += B->dq[k + Q * i] * in[i + N * j];
            b += B->vq[k + Q * i] * in[i + N * j];
        }
        scr_x[v * Q * N + k + Q * j] = a;
        scr_y[v * Q * N + k + Q * j] = b;
    }
    __syncthreads();
    const double sx = .5 * g.hy, sy = .5 * g.hx, gm = g.gamma - 1.;
    for (int t = threadIdx.x; t < 4 * QQ; t += BLOCK) {
        const int v = t / QQ, q = t % QQ, k = q % Q, b = q / Q;
        double a = 0, c = 0;
        for (int j = 0; j < N; ++j) {
            a += B->vq[b + Q * j] * scr_x[v * Q * N + k + Q * j];
            c += B->dq[b + Q * j] * scr_y[v * Q * N + k + Q * j];
        }
        fxs[v * QQ + q] = sx * B->qw[b] * a;
        fys[v * QQ + q] = sy * B->qw[k] * c;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < QQ; q += BLOCK) {
        double gx[4], gy[4], tx[4], ty[4];
        for (int v = 0; v < 4; ++v) {
            gx[v] = fxs[v * QQ + q];
            gy[v] = fys[v * QQ + q];
        }
        const long long vb = static_cast<long long>(cell) * 3 * QQ;
        normal_action_t(vc[vb + q], vc[vb + QQ + q], vc[vb + 2 * QQ + q], gx, 0, gm, g.gamma, tx);
        normal_action_t(vc[vb + q], vc[vb + QQ + q], vc[vb + 2 * QQ + q], gy, 1, gm, g.gamma, ty);
        for (int v = 0; v < 4; ++v) pts[v * QQ + q] = tx[v] + ty[v];
    }
    __syncthreads();
    // transposed interpolation: T3[v][k + n b] = sum_a vq(a, k) P[v][a + Q b]
    for (int t = threadIdx.x; t < 4 * N * Q; t += BLOCK) {
        const int v = t / (N * Q), r = t % (N * Q), k = r % N, b = r / N;
        double s = 0;
        for (int a = 0; a < Q; ++a) s += B->vq[a + Q * k] * pts[v * QQ + a + Q * b];
        tmp[v * N * Q + k + N * b] = s;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < O::ITEMS; ++it) {
        const int item = threadIdx.x + it * BLOCK;
        if (item >= O::M) continue;
        const int v = item / O::P, node = item % O::P, i = node % N, j = node / N;
        double vol = 0;
        for (int b = 0; b < Q; ++b) vol += B->vq[b + Q * j] * tmp[v * N * Q + i + N * b];
        double acc = B->mass[node] * nodal[item] * inv_step - vol;
        const long long f0 = neighbor(cell, 0, g.nx, g.ny), f1 = cell;
        const long long f2 = g.cells + neighbor(cell, 2, g.nx, g.ny), f3 = g.cells + cell;
        acc += B->right[i] * gtr[f1 * 8 * N + v * N + j];
        acc += B->left[i] * gtr[f0 * 8 * N + 4 * N + v * N + j];
        acc += B->right[j] * gtr[f3 * 8 * N + v * N + i];
        acc += B->left[j] * gtr[f2 * 8 * N + 4 * N + v * N + i];
        z[base + item] = acc;
    }
}

// discretization.cpp:163-195 with the volume speeds at the Q^2 points
template <int N>
__global__ __launch_bounds__(BLOCK) void k_oi_freeze_rate(const double *__restrict__ u,
                                                          const double *__restrict__ speeds,
                                                          unsigned long long *rate_bits, int *flag,
                                                          const Basis *__restrict__ B, Geo g) {
    using O = OI<N>;
    __shared__ double nodal[O::M], tmp[4 * MAXQ * N], pts[4 * MAXQ * MAXQ];
    __shared__ unsigned long long alpha[2];
    const int cell = blockIdx.x, Q = g.nq, QQ = Q * Q;
    for (int t = threadIdx.x; t < O::M; t += BLOCK) nodal[t] = u[static_cast<long long>(cell) * O::M + t];
    if (threadIdx.x < 2) alpha[threadIdx.x] = 0;
    __syncthreads();
    oi_interpolate<N>(nodal, tmp, pts, B, Q);
    for (int q = threadIdx.x; q < QQ; q += BLOCK) {
        double st[4];
        for (int v = 0; v < 4; ++v) st[v] = pts[v * QQ + q];
        const Prim p = primitive(st, g.gamma);
        if (!(st[0] > 0 && p.p > 0)) atomicOr(flag, 2);
        const double sound = sqrt(g.gamma * p.p * p.inv);
        atomicMax(&alpha[0], static_cast<unsigned long long>(__double_as_longlong(fabs(p.u) + sound)));
        atomicMax(&alpha[1], static_cast<unsigned long long>(__double_as_longlong(fabs(p.v) + sound)));
    }
    for (int t = threadIdx.x; t < 4 * N; t += BLOCK) {
        const int side = t / N, q = t % N;
        const int face = side == 0 ? neighbor(cell, 0, g.nx, g.ny)
                       : side == 1 ? cell
                       : side == 2 ? g.cells + neighbor(cell, 2, g.nx, g.ny)
                                   : g.cells + cell;
        const double a = speeds[2LL * (static_cast<long long>(face) * N + q)];
        const double b = speeds[2LL * (static_cast<long long>(face) * N + q) + 1];
        atomicMax(&alpha[side / 2], static_cast<unsigned long long>(__double_as_longlong(fmax(a, b))));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double ax = __longlong_as_double(static_cast<long long>(alpha[0]));
        const double ay = __longlong_as_double(static_cast<long long>(alpha[1]));
        atomicMax(rate_bits, static_cast<unsigned long long>(__double_as_longlong(ax / g.hx + ay / g.hy)));
    }
}

// discretization.cpp:220-233: the quadrature points of an over-integrated rule
// must be physical too (the nodes and faces are checked by k_admissible)
template <int N>
__global__ __launch_bounds__(BLOCK) void k_oi_admissible(const double *__restrict__ u, int *flag,
                                                         const Basis *__restrict__ B, Geo g) {
    using O = OI<N>;
    __shared__ double nodal[O::M], tmp[4 * MAXQ * N], pts[4 * MAXQ * MAXQ];
    const int cell = blockIdx.x, Q = g.nq, QQ = Q * Q;
    for (int t = threadIdx.x; t < O::M; t += BLOCK) nodal[t] = u[static_cast<long long>(cell) * O::M + t];
    __syncthreads();
    oi_interpolate<N>(nodal, tmp, pts, B, Q);
    for (int q = threadIdx.x; q < QQ; q += BLOCK) {
        double st[4];
        bool ok = true;
        for (int v = 0; v < 4; ++v) {
            st[v] = pts[v * QQ + q];
            ok = ok && isfinite(st[v]);
        }
        ok = ok && st[0] > 1e-13 && pressure(st, g.gamma) > 1e-13;
        if (!ok) atomicAnd(flag, 0);
    }
}

// ------------------------------------------------------------ initial data
// example.cpp:20-39 (rotated simple wave) and :12-16 (entropy wave), t = 0.
__device__ void exact_state(double x, double y, const Geo &g, double out[4]) {
    const double gam = g.gamma, pi = 3.141592653589793;
    if (g.testcase == 1) {
        const double rho = 1 + .2 * sin(pi * x) * sin(pi * y);
        out[0] = rho;
        out[1] = rho;
        out[2] = .5 * rho;
        out[3] = 1 / (gam - 1) + .625 * rho;
        return;
    }
    const bool diagonal = g.testcase == 0;
    if (diagonal) x = g.cosine * x + g.sine * y;
    const double xi = x;  // t = 0: the characteristic foot is the point itself
    const double jm = sin(pi * xi), velocity = .5 * (2 + jm), sound = (gam - 1) * (2 - jm) / 4;
    const double rho = pow(sound * sound / gam, 1 / (gam - 1));
    out[0] = rho;
    out[1] = rho * velocity * (diagonal ? g.cosine : 1.);
    out[2] = diagonal ? rho * velocity * g.sine : 0.;
    out[3] = pow(rho, gam) / (gam - 1) + .5 * rho * velocity * velocity;
}

__device__ __forceinline__ bool physical_1e12(const double z[4], double gamma) {
    return z[0] > 1e-12 && pressure(z, gamma) > 1e-12;
}

// initial_and_error.cpp:19-51 for the element held in `nodal` (shared memory,
// [v][n^2]); `checks` receives the rows [node | error point | face] x 4.
// Executed by the whole block; thread 0 runs the sequential theta scan.
template <int N>
__device__ void limit_cell(double *nodal, double *checks, double *mean, double *theta_out,
                           const Basis *__restrict__ B, const Geo &g) {
    constexpr int P = N * N;
    const int EQ = g.eq, QQ = EQ * EQ;
    const int rows = P + QQ + 4 * N;
    for (int k = threadIdx.x; k < rows; k += blockDim.x) {
        double s[4];
        if (k < P) {
            for (int v = 0; v < 4; ++v) s[v] = nodal[v * P + k];
        } else if (k < P + QQ) {
            const int e = k - P, qa = e % EQ, qb = e / EQ;
            // error_tensor.interpolate: sum_j ev(qb, j) * sum_i ev(qa, i) nodal(i, j)
            for (int v = 0; v < 4; ++v) {
                double acc = 0;
                for (int jj = 0; jj < N; ++jj) {
                    double t = 0;
                    for (int ii = 0; ii < N; ++ii) t += nodal[v * P + ii + N * jj] * B->ev[qa + EQ * ii];
                    acc += B->ev[qb + EQ * jj] * t;
                }
                s[v] = acc;
            }
        } else {
            const int f = k - P - QQ, side = f / N, q = f % N;
            const double *e = side % 2 ? B->right : B->left;
            for (int v = 0; v < 4; ++v) {
                double acc = 0;
                for (int t = 0; t < N; ++t) acc += (side < 2 ? nodal[v * P + t + N * q] : nodal[v * P + q + N * t]) * e[t];
                s[v] = acc;
            }
        }
        for (int v = 0; v < 4; ++v) checks[4 * k + v] = s[v];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m[4] = {0, 0, 0, 0};
        for (int jj = 0; jj < N; ++jj)
            for (int ii = 0; ii < N; ++ii)
                for (int v = 0; v < 4; ++v) m[v] += .25 * B->w[ii] * B->w[jj] * nodal[v * P + ii + N * jj];
        double theta = (m[0] > 0 && pressure(m, g.gamma) > 0) ? 1. : -1.;
        for (int k = 0; k < rows && theta >= 0; ++k) {
            double d[4], z[4];
            for (int v = 0; v < 4; ++v) d[v] = checks[4 * k + v] - m[v];
            for (int v = 0; v < 4; ++v) z[v] = m[v] + theta * d[v];
            if (!physical_1e12(z, g.gamma)) {
                double low = 0, high = theta;
                for (int it = 0; it < 60; ++it) {
                    const double middle = .5 * (low + high);
                    for (int v = 0; v < 4; ++v) z[v] = m[v] + middle * d[v];
                    if (physical_1e12(z, g.gamma)) low = middle; else high = middle;
                }
                theta = .99 * low;
            }
        }
        for (int v = 0; v < 4; ++v) mean[v] = m[v];
        *theta_out = theta;
    }
    __syncthreads();
    const double theta = *theta_out;
    if (theta >= 0)
        for (int k = threadIdx.x; k < 4 * P; k += blockDim.x) {
            const int v = k / P;
            nodal[k] = mean[v] + theta * (nodal[k] - mean[v]);
        }
    __syncthreads();
}

// initial_and_error.cpp:4-54: L2 projection of the exact state from EQ x EQ
// Gauss samples per element, then positivity scaling. One block per element.
template <int N>
__global__ __launch_bounds__(BLOCK) void k_initial(double *__restrict__ u, int *flag, const Basis *__restrict__ B,
                                                   Geo g) {
    constexpr int P = N * N, M = 4 * P;
    __shared__ double samples[4 * MAXEQ * MAXEQ];
    __shared__ double scratch[4 * MAXN * MAXEQ];
    __shared__ double nodal[M];
    __shared__ double checks[4 * (P + MAXEQ * MAXEQ + 4 * N)];
    __shared__ double mean[4], theta;
    const long long cell = blockIdx.x;
    const int EQ = g.eq, QQ = EQ * EQ;
    const double cx = -.5 * g.lx + g.hx * (cell % g.nx + .5), cy = -.5 * g.ly + g.hy * (cell / g.nx + .5);
    for (int k = threadIdx.x; k < QQ; k += blockDim.x) {
        const int qi = k % EQ, qj = k / EQ;
        double s[4];
        exact_state(cx + .5 * g.hx * B->ex[qi], cy + .5 * g.hy * B->ex[qj], g, s);
        for (int v = 0; v < 4; ++v) samples[v * QQ + k] = s[v];
    }
    __syncthreads();
    // scratch(i, qb) = sum_qa proj(i, qa) * S(qa, qb)
    for (int k = threadIdx.x; k < 4 * N * EQ; k += blockDim.x) {
        const int v = k / (N * EQ), r = k % (N * EQ), i = r % N, qb = r / N;
        double acc = 0;
        for (int qa = 0; qa < EQ; ++qa) acc += samples[v * QQ + qa + EQ * qb] * B->proj[i + N * qa];
        scratch[k] = acc;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < M; k += blockDim.x) {
        const int v = k / P, node = k % P, i = node % N, j = node / N;
        double acc = 0;
        for (int qb = 0; qb < EQ; ++qb) acc += B->proj[j + N * qb] * scratch[v * N * EQ + i + N * qb];
        nodal[k] = acc;
    }
    __syncthreads();
    limit_cell<N>(nodal, checks, mean, &theta, B, g);
    if (theta < 0 && threadIdx.x == 0) atomicOr(flag, 4);
    for (int k = threadIdx.x; k < M; k += blockDim.x) u[cell * M + k] = nodal[k];
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double uniform01(unsigned long long seed, unsigned long long index) {
    return static_cast<double>(mix64(seed + (index + 1) * 0x9E3779B97F4A7C15ull) >> 11) * 0x1.0p-53;
}

// Seeded state: rho ~ U[0.5,1.5], u,v ~ U[-0.5,0.5], p ~ U[0.5,1.5] i.i.d. per
// node (splitmix64 counter stream), converted to conserved variables and
// scaled into the admissible set with the reference's positivity scaling.
template <int N>
__global__ __launch_bounds__(BLOCK) void k_seeded(double *__restrict__ u, unsigned long long seed, int *flag,
                                                  const Basis *__restrict__ B, Geo g) {
    constexpr int P = N * N, M = 4 * P;
    __shared__ double nodal[M];
    __shared__ double checks[4 * (P + MAXEQ * MAXEQ + 4 * N)];
    __shared__ double mean[4], theta;
    const long long cell = blockIdx.x;
    for (int k = threadIdx.x; k < P; k += blockDim.x) {
        const unsigned long long base = 4ull * (static_cast<unsigned long long>(cell) * P + k);
        const double rho = .5 + uniform01(seed, base), vx = -.5 + uniform01(seed, base + 1);
        const double vy = -.5 + uniform01(seed, base + 2), p = .5 + uniform01(seed, base + 3);
        nodal[k] = rho;
        nodal[P + k] = rho * vx;
        nodal[2 * P + k] = rho * vy;
        nodal[3 * P + k] = p / (g.gamma - 1) + .5 * rho * (vx * vx + vy * vy);
    }
    __syncthreads();
    limit_cell<N>(nodal, checks, mean, &theta, B, g);
    if (theta < 0 && threadIdx.x == 0) atomicOr(flag, 4);
    for (int k = threadIdx.x; k < M; k += blockDim.x) u[cell * M + k] = nodal[k];
}

// initial_and_error.cpp:55-73: per-block partial sums of the squared L2 error.
template <int N>
__global__ __launch_bounds__(BLOCK) void k_error(const double *__restrict__ u, double time, double *partials,
                                                 const Basis *__restrict__ B, Geo g) {
    constexpr int P = N * N, M = 4 * P;
    __shared__ double red[4][BLOCK];
    double acc[4] = {0, 0, 0, 0};
    const int EQ = g.eq, QQ = EQ * EQ;
    for (long long w = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
         w < static_cast<long long>(g.cells) * QQ; w += static_cast<long long>(gridDim.x) * blockDim.x) {
        const long long cell = w / QQ;
        const int k = static_cast<int>(w % QQ), qa = k % EQ, qb = k / EQ;
        double val[4];
        for (int v = 0; v < 4; ++v) {
            double s = 0;
            for (int jj = 0; jj < N; ++jj) {
                double t = 0;
                for (int ii = 0; ii < N; ++ii) t += u[cell * M + v * P + ii + N * jj] * B->ev[qa + EQ * ii];
                s += B->ev[qb + EQ * jj] * t;
            }
            val[v] = s;
        }
        const double cx = -.5 * g.lx + g.hx * (cell % g.nx + .5), cy = -.5 * g.ly + g.hy * (cell / g.nx + .5);
        double ex[4];
        double x = cx + .5 * g.hx * B->ex[qa], y = cy + .5 * g.hy * B->ex[qb];
        if (g.testcase == 1) {
            const double pi = 3.141592653589793, gam = g.gamma;
            const double rho = 1 + .2 * sin(pi * (x - time)) * sin(pi * (y - .5 * time));
            ex[0] = rho; ex[1] = rho; ex[2] = .5 * rho; ex[3] = 1 / (gam - 1) + .625 * rho;
        } else if (g.testcase == 3) {
            for (int v = 0; v < 4; ++v) ex[v] = nan("");
        } else {
            const double gam = g.gamma, pi = 3.141592653589793;
            const bool diagonal = g.testcase == 0;
            if (diagonal) x = g.cosine * x + g.sine * y;
            const double a = (3 - gam) / 2, b = (gam + 1) / 4;
            double lo = x - time * (a + b), hi = x - time * (a - b), xi = x - time * a;
            for (int it = 0; it < 80; ++it) {
                const double f = xi + time * (a + b * sin(pi * xi)) - x;
                if (fabs(f) < 2e-15) break;
                if (f > 0) hi = xi; else lo = xi;
                const double zz = xi - f / (1 + time * b * pi * cos(pi * xi));
                xi = zz > lo && zz < hi ? zz : .5 * (lo + hi);
            }
            const double jm = sin(pi * xi), velocity = .5 * (2 + jm), sound = (gam - 1) * (2 - jm) / 4;
            const double rho = pow(sound * sound / gam, 1 / (gam - 1));
            ex[0] = rho;
            ex[1] = rho * velocity * (diagonal ? g.cosine : 1.);
            ex[2] = diagonal ? rho * velocity * g.sine : 0.;
            ex[3] = pow(rho, gam) / (gam - 1) + .5 * rho * velocity * velocity;
        }
        const double wgt = .25 * g.hx * g.hy * B->ew[qa] * B->ew[qb];
        for (int v = 0; v < 4; ++v) {
            const double d = val[v] - ex[v];
            acc[v] += wgt * (d * d);
        }
    }
    for (int v = 0; v < 4; ++v) red[v][threadIdx.x] = acc[v];
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int v = 0; v < 4; ++v) red[v][threadIdx.x] += red[v][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 4) partials[blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

unsigned grid_for(long long threads, int block) { return static_cast<unsigned>((threads + block - 1) / block); }

}  // namespace

#define DGH_DISPATCH_N(n, ...)                   \
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
    default: return hipErrorInvalidValue;         \
    }                                             \
    return hipGetLastError();

template <typename TX>
static hipError_t restrict_any(int n, const TX *x, double *traces, const Basis *b, const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, {
        k_restrict<N, TX><<<grid_for(4LL * g.cells * N, 256), 256, 0, s>>>(x, traces, g, b);
    })
}
hipError_t launch_restrict(int n, const double *x, double *traces, const Basis *b, const Geo &g, hipStream_t s) {
    return restrict_any(n, x, traces, b, g, s);
}
hipError_t launch_restrict(int n, const float *x, double *traces, const Basis *b, const Geo &g, hipStream_t s) {
    return restrict_any(n, x, traces, b, g, s);
}
hipError_t launch_freeze_speeds(int n, const double *traces, double *speeds, int *flag, const Geo &g,
                                hipStream_t s) {
    DGH_DISPATCH_N(n, { k_freeze_speeds<N><<<grid_for(2LL * g.cells * N, BLOCK), BLOCK, 0, s>>>(traces, speeds, flag, g); })
}
hipError_t launch_freeze_rate(int n, const double *u, const double *speeds, unsigned long long *rate_bits,
                              int *flag, const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, {
        k_freeze_rate<N><<<grid_for(g.cells, Shape<N>::CPB), BLOCK, 0, s>>>(u, speeds, rate_bits, flag, g);
    })
}
hipError_t launch_residual_flux(int n, const double *traces, const double *speeds, double *fq, const Basis *b,
                                const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, { k_residual_flux<N><<<grid_for(2LL * g.cells * N, BLOCK), BLOCK, 0, s>>>(traces, speeds, fq, b, g); })
}
hipError_t launch_element_residual(int n, const double *u, const double *old, const double *fq, double *r,
                                   double tau, const Basis *b, const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, {
        k_element_residual<N><<<grid_for(g.cells, Shape<N>::CPB), BLOCK, 0, s>>>(u, old, fq, r, tau, b, g);
    })
}
hipError_t launch_admissible(int n, const double *u, int *flag, const Basis *b, const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, { k_admissible<N><<<grid_for(g.cells, Shape<N>::CPB), BLOCK, 0, s>>>(u, flag, b, g); })
}
hipError_t launch_volume_cache(int n, const double *u, double *vcache, const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, { k_volume_cache<N><<<grid_for(static_cast<long long>(g.cells) * N * N, BLOCK), BLOCK, 0, s>>>(u, vcache, g); })
}
hipError_t launch_face_cache(int n, const double *traces, const double *speeds, double *fcache, const Basis *b,
                             const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, { k_face_cache<N><<<grid_for(2LL * g.cells * N, BLOCK), BLOCK, 0, s>>>(traces, speeds, fcache, b, g); })
}
hipError_t launch_jv_flux(int n, const double *traces, const double *fcache, double *fq, const Geo &g,
                          hipStream_t s) {
    DGH_DISPATCH_N(n, { k_jv_flux<N><<<grid_for(2LL * g.cells * N, BLOCK), BLOCK, 0, s>>>(traces, fcache, fq, g); })
}
template <typename TX, typename TY>
static hipError_t element_jv_any(int n, const TX *x, const double *vcache, const double *fq, TY *y, double inv_step,
                                 const Basis *b, const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, {
        k_element_jv<N, TX, TY><<<grid_for(g.cells, Shape<N>::CPB), BLOCK, 0, s>>>(x, vcache, fq, y, inv_step, b, g);
    })
}
hipError_t launch_element_jv(int n, const double *x, const double *vcache, const double *fq, double *y,
                             double inv_step, const Basis *b, const Geo &g, hipStream_t s) {
    return element_jv_any(n, x, vcache, fq, y, inv_step, b, g, s);
}
hipError_t launch_element_jv(int n, const float *x, const double *vcache, const double *fq, float *y,
                             double inv_step, const Basis *b, const Geo &g, hipStream_t s) {
    return element_jv_any(n, x, vcache, fq, y, inv_step, b, g, s);
}
hipError_t launch_element_jv(int n, const float *x, const double *vcache, const double *fq, double *y,
                             double inv_step, const Basis *b, const Geo &g, hipStream_t s) {
    return element_jv_any(n, x, vcache, fq, y, inv_step, b, g, s);
}
hipError_t launch_oi_element_residual(int n, const double *u, const double *old, const double *fq, double *r,
                                      double tau, const Basis *b, const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, { k_oi_element_residual<N><<<g.cells, BLOCK, 0, s>>>(u, old, fq, r, tau, b, g); })
}
hipError_t launch_oi_volume_cache(int n, const double *u, double *vcache, const Basis *b, const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, { k_oi_volume_cache<N><<<g.cells, BLOCK, 0, s>>>(u, vcache, b, g); })
}
template <typename TX, typename TY>
static hipError_t oi_jv_any(int n, const TX *x, const double *vcache, const double *fq, TY *y, double inv_step,
                            const Basis *b, const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, { k_oi_element_jv<N, TX, TY><<<g.cells, BLOCK, 0, s>>>(x, vcache, fq, y, inv_step, b, g); })
}
hipError_t launch_oi_element_jv(int n, const double *x, const double *vcache, const double *fq, double *y,
                                double inv_step, const Basis *b, const Geo &g, hipStream_t s) {
    return oi_jv_any(n, x, vcache, fq, y, inv_step, b, g, s);
}
hipError_t launch_oi_element_jv(int n, const float *x, const double *vcache, const double *fq, float *y,
                                double inv_step, const Basis *b, const Geo &g, hipStream_t s) {
    return oi_jv_any(n, x, vcache, fq, y, inv_step, b, g, s);
}
hipError_t launch_oi_element_jv(int n, const float *x, const double *vcache, const double *fq, double *y,
                                double inv_step, const Basis *b, const Geo &g, hipStream_t s) {
    return oi_jv_any(n, x, vcache, fq, y, inv_step, b, g, s);
}
hipError_t launch_oi_element_jtv(int n, const double *w, const double *vcache, const double *gtraces, double *z,
                                 double inv_step, const Basis *b, const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, { k_oi_element_jtv<N><<<g.cells, BLOCK, 0, s>>>(w, vcache, gtraces, z, inv_step, b, g); })
}
hipError_t launch_oi_freeze_rate(int n, const double *u, const double *speeds, unsigned long long *rate_bits,
                                 int *flag, const Basis *b, const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, { k_oi_freeze_rate<N><<<g.cells, BLOCK, 0, s>>>(u, speeds, rate_bits, flag, b, g); })
}
hipError_t launch_oi_admissible(int n, const double *u, int *flag, const Basis *b, const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, { k_oi_admissible<N><<<g.cells, BLOCK, 0, s>>>(u, flag, b, g); })
}
hipError_t launch_adjoint_face(int n, const double *wtraces, const double *fcache, double *gtraces, const Geo &g,
                               hipStream_t s) {
    DGH_DISPATCH_N(n, { k_adjoint_face<N><<<grid_for(2LL * g.cells * N, BLOCK), BLOCK, 0, s>>>(wtraces, fcache, gtraces, g); })
}
hipError_t launch_element_jtv(int n, const double *w, const double *vcache, const double *gtraces, double *z,
                              double inv_step, const Basis *b, const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, {
        k_element_jtv<N><<<grid_for(g.cells, Shape<N>::CPB), BLOCK, 0, s>>>(w, vcache, gtraces, z, inv_step, b, g);
    })
}
hipError_t launch_initial(int n, double *u, int *flag, const Basis *b, const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, { k_initial<N><<<g.cells, BLOCK, 0, s>>>(u, flag, b, g); })
}
hipError_t launch_seeded(int n, double *u, unsigned long long seed, int *flag, const Basis *b, const Geo &g,
                         hipStream_t s) {
    DGH_DISPATCH_N(n, { k_seeded<N><<<g.cells, BLOCK, 0, s>>>(u, seed, flag, b, g); })
}
hipError_t launch_error_partials(int n, const double *u, double time, double *partials, int nblocks,
                                 const Basis *b, const Geo &g, hipStream_t s) {
    DGH_DISPATCH_N(n, { k_error<N><<<nblocks, BLOCK, 0, s>>>(u, time, partials, b, g); })
}

}  // namespace dgh
