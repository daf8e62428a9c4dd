// extern "C" surface of the CPU oracle, loaded through ctypes by tests/,
// __graft_entry__.smoke() and the CPU-baseline leg of bench.py only.
// TEST INFRASTRUCTURE: never linked into the product library.
#include <cstring>
#include <memory>
#include <string>

#include "dgh_oracle.hpp"

using namespace oracle;

namespace {
thread_local std::string last_error;

struct Handle {
    Problem problem;
    std::unique_ptr<DGOperator> dg;
    std::unique_ptr<BlockJacobi> bj;
    std::unique_ptr<HybridOperator> hdg;
    std::unique_ptr<BlockJacobi> face_bj;
    std::unique_ptr<BlockILU0> ilu;
    explicit Handle(const Config &c) : problem(c) {}
    DGOperator &op() {
        if (!dg) dg.reset(new DGOperator(problem));
        return *dg;
    }
    HybridOperator &hy() {
        if (!hdg) hdg.reset(new HybridOperator(problem));
        return *hdg;
    }
};

template <class F> int guarded(F &&f) {
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        last_error = e.what();
        return 1;
    }
}
Vec copy_in(const double *p, size_t n) { return Vec(p, p + n); }
void copy_out(const Vec &v, double *p) { std::memcpy(p, v.data(), sizeof(double) * v.size()); }
}  // namespace

extern "C" {
const char *orc_error(void) { return last_error.c_str(); }

void *orc_create(int argc, const char *const *argv) {
    try {
        return new Handle(parse(argc, argv));
    } catch (const std::exception &e) {
        last_error = e.what();
        return nullptr;
    }
}
void orc_destroy(void *h) { delete static_cast<Handle *>(h); }

// out: volume size, trace size, cells, nodes, points, nx, ny, face_count
int orc_dims(void *h, long long *out) {
    auto &p = static_cast<Handle *>(h)->problem;
    out[0] = p.layout.size();
    out[1] = p.trace_size();
    out[2] = p.layout.cells;
    out[3] = p.basis.nodes;
    out[4] = p.basis.points;
    out[5] = p.layout.nx;
    out[6] = p.layout.ny;
    out[7] = p.face_count();
    return 0;
}
// out: hx, hy, lx, ly, step_size
int orc_geometry(void *h, double *out) {
    auto &p = static_cast<Handle *>(h)->problem;
    out[0] = p.hx; out[1] = p.hy; out[2] = p.example.lx; out[3] = p.example.ly; out[4] = p.step_size;
    return 0;
}
// basis tables: x, w (n each), left, right (n each), weighted derivatives (n*n column-major)
int orc_basis(void *h, double *out) {
    auto &b = static_cast<Handle *>(h)->problem.basis;
    const int n = b.nodes;
    for (int i = 0; i < n; ++i) {
        out[i] = b.x[i]; out[n + i] = b.w[i]; out[2 * n + i] = b.left[i]; out[3 * n + i] = b.right[i];
    }
    if (b.points == n) std::memcpy(out + 4 * n, b.weighted_derivatives.a.data(), sizeof(double) * n * n);
    return 0;
}
int orc_gauss(int n, double *x, double *w) {
    return guarded([&] {
        Vec xs, ws;
        Basis::gauss(n, xs, ws);
        copy_out(xs, x);
        copy_out(ws, w);
    });
}
int orc_initial(void *h, double *u) {
    return guarded([&] {
        Vec v;
        static_cast<Handle *>(h)->problem.initial(v);
        copy_out(v, u);
    });
}
int orc_seeded(void *h, unsigned long long seed, double *u) {
    return guarded([&] {
        Vec v;
        static_cast<Handle *>(h)->problem.seeded(v, seed);
        copy_out(v, u);
    });
}
int orc_uniform(unsigned long long seed, long long count, double *out) {
    for (long long i = 0; i < count; ++i) out[i] = uniform01(seed, i);
    return 0;
}
int orc_normal(unsigned long long seed, long long count, double *out) {
    for (long long i = 0; i < count; ++i) out[i] = normal01(seed, i);
    return 0;
}
int orc_freeze(void *h, const double *u, double *rate) {
    return guarded([&] {
        auto &p = static_cast<Handle *>(h)->problem;
        *rate = p.freeze(copy_in(u, p.layout.size()));
    });
}
// speeds as [face*n + q][2] = (a, b)
int orc_speeds(void *h, double *out) {
    auto &p = static_cast<Handle *>(h)->problem;
    for (size_t i = 0; i < p.speeds.size(); ++i) { out[2 * i] = p.speeds[i].a; out[2 * i + 1] = p.speeds[i].b; }
    return 0;
}
int orc_set_step(void *h, double tau, const double *old) {
    return guarded([&] {
        auto &p = static_cast<Handle *>(h)->problem;
        p.step_size = tau;
        if (old) p.old = copy_in(old, p.layout.size());
    });
}
int orc_dg_residual(void *h, const double *u, double *r) {
    return guarded([&] {
        auto &p = static_cast<Handle *>(h)->problem;
        Vec out;
        p.dg_residual(copy_in(u, p.layout.size()), out);
        copy_out(out, r);
    });
}
int orc_admissible(void *h, const double *u, int *ok) {
    return guarded([&] {
        auto &p = static_cast<Handle *>(h)->problem;
        *ok = p.admissible(copy_in(u, p.layout.size())) ? 1 : 0;
    });
}
int orc_residual_norm(void *h, const double *r, double *out) {
    return guarded([&] {
        auto &p = static_cast<Handle *>(h)->problem;
        *out = p.residual_norm(copy_in(r, p.layout.size()));
    });
}
int orc_integral(void *h, const double *u, double *out) {
    return guarded([&] {
        auto &p = static_cast<Handle *>(h)->problem;
        State s = p.integral(copy_in(u, p.layout.size()));
        for (int v = 0; v < NVAR; ++v) out[v] = s[v];
    });
}
int orc_error_l2(void *h, const double *u, double t, double *out) {
    return guarded([&] {
        auto &p = static_cast<Handle *>(h)->problem;
        State s = p.error(copy_in(u, p.layout.size()), t);
        for (int v = 0; v < NVAR; ++v) out[v] = s[v];
    });
}
int orc_trace(void *h, const double *u, double *lambda) {
    return guarded([&] {
        auto &p = static_cast<Handle *>(h)->problem;
        Vec out;
        p.trace(copy_in(u, p.layout.size()), out);
        copy_out(out, lambda);
    });
}
int orc_hybrid_residual(void *h, const double *u, const double *lambda, double *r, double *g) {
    return guarded([&] {
        auto &p = static_cast<Handle *>(h)->problem;
        Vec ro, go;
        p.hybrid_residual(copy_in(u, p.layout.size()), copy_in(lambda, p.trace_size()), ro, go);
        copy_out(ro, r);
        copy_out(go, g);
    });
}
int orc_dg_refresh(void *h, const double *u) {
    return guarded([&] {
        auto *H = static_cast<Handle *>(h);
        H->op().refresh(copy_in(u, H->problem.layout.size()));
    });
}
int orc_dg_apply(void *h, const double *x, double *y) {
    return guarded([&] {
        auto *H = static_cast<Handle *>(h);
        Vec out;
        H->op().apply(copy_in(x, H->problem.layout.size()), out);
        copy_out(out, y);
    });
}
int orc_dg_apply_transpose(void *h, const double *w, double *z) {
    return guarded([&] {
        auto *H = static_cast<Handle *>(h);
        Vec out;
        H->op().apply_transpose_blocks(copy_in(w, H->problem.layout.size()), out);
        copy_out(out, z);
    });
}
int orc_dg_block(void *h, int row, int col, double *out) {
    return guarded([&] {
        Mat b;
        static_cast<Handle *>(h)->op().block(row, col, b);
        copy_out(b.a, out);
    });
}
// volume-cache (u, v, H) per point of a refreshed operator: [cells][3][points^2]
int orc_volume_cache(void *h, double *out) {
    auto &vj = static_cast<Handle *>(h)->op().volume;
    const int P = vj.primitives.rows, cols = vj.primitives.cols;
    for (int c = 0; c < cols; ++c)
        for (int q = 0; q < P; ++q) out[static_cast<size_t>(c) * P + q] = vj.primitives(q, c);
    return 0;
}
int orc_bj_refresh(void *h) {
    return guarded([&] {
        auto *H = static_cast<Handle *>(h);
        auto &p = H->problem;
        if (!H->bj) H->bj.reset(new BlockJacobi(p.layout.cells, p.layout.block_size()));
        auto &op = H->op();
        H->bj->refresh([&](int r, int c, Mat &b) { op.block(r, c, b); });
    });
}
int orc_bj_apply(void *h, const double *r, double *z) {
    return guarded([&] {
        auto *H = static_cast<Handle *>(h);
        require(H->bj != nullptr, "Refresh the block Jacobi preconditioner first");
        Vec out;
        H->bj->apply(copy_in(r, H->problem.layout.size()), out);
        copy_out(out, z);
    });
}
// DG block ILU(0) over the cartesian element graph; ordering 0 natural, 1 red-black
int orc_ilu_refresh(void *h, int ordering) {
    return guarded([&] {
        auto *H = static_cast<Handle *>(h);
        auto &p = H->problem;
        std::vector<int> order(p.layout.cells);
        for (int i = 0; i < p.layout.cells; ++i) order[i] = i;
        if (ordering == 1) order = red_black_order(p.layout.nx, p.layout.ny);
        H->ilu.reset(new BlockILU0(cartesian_graph(p.layout.nx, p.layout.ny), p.layout.block_size(), order));
        auto &op = H->op();
        H->ilu->refresh([&](int r, int c, Mat &b) { op.block(r, c, b); });
    });
}
int orc_ilu_apply(void *h, const double *r, double *z) {
    return guarded([&] {
        auto *H = static_cast<Handle *>(h);
        require(H->ilu != nullptr, "Refresh the block ILU(0) preconditioner first");
        Vec out;
        H->ilu->apply(copy_in(r, H->problem.layout.size()), out);
        copy_out(out, z);
    });
}
int orc_ilu_gmres(void *h, const double *rhs, double *x, int restart, int limit, double tol, int *its, double *rel) {
    return guarded([&] {
        auto *H = static_cast<Handle *>(h);
        auto &p = H->problem;
        require(H->ilu != nullptr, "Refresh the block ILU(0) preconditioner first");
        GMRES g(p.layout.size(), restart, p.inverse_mass);
        Vec out;
        auto &op = H->op();
        auto &ilu = *H->ilu;
        auto r = g.solve([&](const Vec &a, Vec &b) { op.apply(a, b); }, [&](const Vec &a, Vec &b) { ilu.apply(a, b); },
                         copy_in(rhs, p.layout.size()), out, limit, tol);
        copy_out(out, x);
        *its = r.iterations;
        *rel = r.relative_residual;
    });
}
// Standalone block ILU(0) of a block matrix given in CSR over blocks (each row
// holding its diagonal; blocks column-major per slot), in the order `order`;
// solves with rhs into x.
int orc_block_ilu_solve(int rows, const int *row_ptr, const int *col_idx, int width, const double *blocks,
                        const int *order, const double *rhs, double *x) {
    return guarded([&] {
        std::vector<std::vector<int>> adjacency(rows);
        for (int r = 0; r < rows; ++r) adjacency[r].assign(col_idx + row_ptr[r], col_idx + row_ptr[r + 1]);
        auto slot = [&](int r, int c) {
            for (int p = row_ptr[r]; p < row_ptr[r + 1]; ++p)
                if (col_idx[p] == c) return p;
            throw std::runtime_error("Requested structural zero block");
        };
        BlockILU0 ilu(adjacency, width, std::vector<int>(order, order + rows));
        const size_t bb = static_cast<size_t>(width) * width;
        ilu.refresh([&](int r, int c, Mat &b) { std::memcpy(b.a.data(), blocks + slot(r, c) * bb, bb * sizeof(double)); });
        Vec out;
        ilu.apply(Vec(rhs, rhs + static_cast<size_t>(rows) * width), out);
        std::memcpy(x, out.data(), out.size() * sizeof(double));
    });
}
// LU factors of one refreshed block (column-major, m*m) and its row sources
int orc_bj_factor(void *h, int block, double *lu, int *rowsrc) {
    return guarded([&] {
        auto *H = static_cast<Handle *>(h);
        require(H->bj != nullptr, "Refresh the block Jacobi preconditioner first");
        const DenseLU &f = H->bj->lu(block);
        copy_out(f.lu.a, lu);
        for (int i = 0; i < f.n; ++i) rowsrc[i] = f.rowsrc[i];
    });
}
// Dense LU solve of one column-major n x n matrix (forced-permutation checks).
int orc_lu_solve(int n, const double *a, const double *b, double *x) {
    return guarded([&] {
        Mat m(n, n);
        std::memcpy(m.a.data(), a, sizeof(double) * n * n);
        DenseLU lu;
        require(lu.factor(m), "Singular block Jacobi factor");
        std::memcpy(x, b, sizeof(double) * n);
        lu.solve_inplace(x);
    });
}
int orc_dg_gmres(void *h, const double *rhs, double *x, int restart, int limit, double tol, int *its, double *rel,
                 double *residual) {
    return guarded([&] {
        auto *H = static_cast<Handle *>(h);
        auto &p = H->problem;
        require(H->bj != nullptr, "Refresh the block Jacobi preconditioner first");
        GMRES g(p.layout.size(), restart, p.inverse_mass);
        Vec out;
        auto &op = H->op();
        auto &bj = *H->bj;
        auto r = g.solve([&](const Vec &a, Vec &b) { op.apply(a, b); }, [&](const Vec &a, Vec &b) { bj.apply(a, b); },
                         copy_in(rhs, p.layout.size()), out, limit, tol);
        copy_out(out, x);
        if (residual) copy_out(g.residual(), residual);
        *its = r.iterations;
        *rel = r.relative_residual;
    });
}
int orc_hdg_refresh_projected(void *h, const double *u, const double *r) {
    return guarded([&] {
        auto *H = static_cast<Handle *>(h);
        const size_t N = H->problem.layout.size();
        H->hy().refresh_projected(copy_in(u, N), copy_in(r, N));
    });
}
int orc_hdg_rhs(void *h, double *out) {
    copy_out(static_cast<Handle *>(h)->hy().rhs(), out);
    return 0;
}
int orc_hdg_apply(void *h, const double *x, double *y) {
    return guarded([&] {
        auto *H = static_cast<Handle *>(h);
        Vec out;
        H->hy().apply(copy_in(x, H->problem.trace_size()), out);
        copy_out(out, y);
    });
}
int orc_hdg_block(void *h, int row, int col, double *out) {
    return guarded([&] {
        Mat b;
        static_cast<Handle *>(h)->hy().block(row, col, b);
        copy_out(b.a, out);
    });
}
int orc_hdg_recover(void *h, const double *dl, double *du) {
    return guarded([&] {
        auto *H = static_cast<Handle *>(h);
        Vec out;
        H->hy().recover(copy_in(dl, H->problem.trace_size()), out);
        copy_out(out, du);
    });
}
int orc_hdg_recovered_residual(void *h, const double *e, double *out) {
    return guarded([&] {
        auto *H = static_cast<Handle *>(h);
        Vec o;
        H->hy().recovered_residual_from_schur(copy_in(e, H->problem.trace_size()), o);
        copy_out(o, out);
    });
}
// X [cells][m x t], S [cells][t x t], y [cells][m] of the last condensation
int orc_hdg_factors(void *h, double *x, double *s, double *y) {
    auto &hy = static_cast<Handle *>(h)->hy();
    if (x) copy_out(hy.x_, x);
    if (s) copy_out(hy.schur_, s);
    if (y) copy_out(hy.y_, y);
    return 0;
}
int orc_hdg_bj_refresh(void *h) {
    return guarded([&] {
        auto *H = static_cast<Handle *>(h);
        auto &p = H->problem;
        if (!H->face_bj) H->face_bj.reset(new BlockJacobi(p.face_count(), p.face_width()));
        auto &hy = H->hy();
        H->face_bj->refresh([&](int r, int c, Mat &b) { hy.block(r, c, b); });
    });
}
int orc_hdg_bj_apply(void *h, const double *r, double *z) {
    return guarded([&] {
        auto *H = static_cast<Handle *>(h);
        require(H->face_bj != nullptr, "Refresh the face block Jacobi preconditioner first");
        Vec out;
        H->face_bj->apply(copy_in(r, H->problem.trace_size()), out);
        copy_out(out, z);
    });
}
// Physics kernels for the check of physics.cpp: flux, analytic and cached
// Jacobian (row-major 4x4) of one state.
int orc_physics(const double *state, int axis, double *flux, double *analytic, double *cached) {
    return guarded([&] {
        Gas gas;
        State u = {state[0], state[1], state[2], state[3]};
        State f = gas.flux(u, axis);
        Jac a = gas.jacobian(u, axis);
        auto p = gas.primitive(u);
        Jac c = gas.cached_jacobian(p.velocity[0], p.velocity[1], p.enthalpy, axis);
        for (int k = 0; k < 4; ++k) flux[k] = f[k];
        for (int k = 0; k < 16; ++k) { analytic[k] = a[k]; cached[k] = c[k]; }
    });
}
// TensorProduct::subtract_weak on one element with Q quadrature points.
int orc_subtract_weak(int degree, int quadrature, double hx, double hy, const double *fx, const double *fy,
                      double *residual) {
    return guarded([&] {
        Basis b(degree, quadrature);
        TensorProduct t(b, hx, hy);
        t.subtract_weak(fx, fy, residual);
    });
}
// Basis tables of a (degree, quadrature) pair: values [Q x n], weighted derivatives [Q x n], qw [Q].
int orc_basis_tables(int degree, int quadrature, double *values, double *wd, double *qw) {
    return guarded([&] {
        Basis b(degree, quadrature);
        copy_out(b.values.a, values);
        copy_out(b.weighted_derivatives.a, wd);
        copy_out(b.qw, qw);
    });
}

// Dense DG Jacobian (volume x volume, column-major) from the block provider.
int orc_dense_jacobian(void *h, double *out) {
    return guarded([&] {
        auto *H = static_cast<Handle *>(h);
        auto &p = H->problem;
        const long long N = p.layout.size();
        require(N <= 20000, "Dense Jacobian limited to small meshes");
        const int m = p.layout.block_size();
        std::memset(out, 0, sizeof(double) * N * N);
        auto graph = cartesian_graph(p.layout.nx, p.layout.ny);
        Mat b;
        for (int row = 0; row < p.layout.cells; ++row)
            for (int col : graph[row]) {
                H->op().block(row, col, b);
                for (int j = 0; j < m; ++j)
                    for (int i = 0; i < m; ++i) out[(static_cast<long long>(col) * m + j) * N + row * m + i] = b(i, j);
            }
    });
}
// Full TimeIntegrator::advance. run[16]: final_time, drift, seconds, max_lin, max_dg_lin,
// min_step, max_step, measured_newton_s, steps, newton, krylov, measured_updates.
// updates[cap][13]: step, stage, iteration, krylov, time, stage_time, dt, cfl, rms,
// damping, linear_rel, dg_linear_rel, iteration_s.
int orc_advance(void *h, const char *method, double *u, double *run, double *updates, int cap, int *count) {
    return guarded([&] {
        auto *H = static_cast<Handle *>(h);
        auto &p = H->problem;
        Vec solution = copy_in(u, p.layout.size());
        TimeIntegrator ti(p, std::string(method) == "hdg" ? Method::HDG : Method::DG);
        RunStatistics s = ti.advance(solution);
        copy_out(solution, u);
        double r[12] = {s.final_time, s.conservation_drift, s.seconds, s.max_linear_residual, s.max_dg_linear_residual,
                        s.minimum_step, s.maximum_step, s.measured_newton_seconds, double(s.steps),
                        double(s.newton_iterations), double(s.krylov_iterations), double(s.measured_updates)};
        std::memcpy(run, r, sizeof(r));
        int k = 0;
        for (const auto &x : s.updates) {
            if (k >= cap) break;
            double row[13] = {double(x.step), double(x.stage), double(x.iteration), double(x.krylov_iterations), x.time,
                              x.stage_time, x.dt, x.cfl, x.residual_physical_rms, x.damping,
                              x.linear_relative_residual, x.dg_linear_relative_residual, x.iteration_s};
            std::memcpy(updates + 13 * k, row, sizeof(row));
            ++k;
        }
        *count = static_cast<int>(s.updates.size());
    });
}
}
