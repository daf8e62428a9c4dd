// CPU FP64 oracle for the 2D tensor-product DG / condensed-HDG Newton-Krylov path.
//
// TEST INFRASTRUCTURE ONLY. This is a from-scratch restatement of the reference
// algorithm (YueWu2002/DG_HDG_Euler_Cpp, dg_hdg_2d/ and common/) in plain C++
// without Eigen. It is the test oracle of the HIP path: only tests/, the smoke
// entry point and the CPU-baseline leg of bench.py may load it. The product
// library (dg_hdg_euler_cpp_amd/) never links or calls it.
//
// Every function cites the reference file:line it restates. Parity status of
// each part (which golden vector / invariant pins it) is listed in the header
// comment of oracle/dgh_oracle.cpp.
#pragma once
#include <array>
#include <cmath>
#include <algorithm>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

namespace oracle {
constexpr int NDIM = 2, NVAR = 4, NFACES = 4;
using Vec = std::vector<double>;

inline void require(bool condition, const char *message) {
    if (!condition) throw std::runtime_error(message);
}

// ---------------------------------------------------------------- threading
// The CPU-baseline build of the oracle is compiled with -fopenmp; the tests'
// build is not, and then these loops run in index order. Both builds give
// bit-identical results: threaded loops write disjoint entries (cells, faces,
// blocks, vector entries); reductions keep fixed 256-term blocks, each summed
// in order by one thread, with the block sums compensated in block order
// (compensated_sum in dgh_oracle.cpp); face lifts are gathered per cell in the
// order the serial face loop applies them (Problem::lift_faces). An exception
// thrown for one index is rethrown after the loop (the lowest such index's
// message when the loop runs in order).
template <class MakeScratch, class Body>
void parallel_for(long long count, MakeScratch make_scratch, Body body) {
    bool failed = false;
    std::string message;
#pragma omp parallel
    {
        auto scratch = make_scratch();
#pragma omp for schedule(static)
        for (long long i = 0; i < count; ++i) {
            try {
                body(i, scratch);
            } catch (const std::exception &e) {
#pragma omp critical(oracle_parallel_error)
                if (!failed) {
                    failed = true;
                    message = e.what();
                }
            }
        }
    }
    if (failed) throw std::runtime_error(message);
}
template <class Body> void parallel_for(long long count, Body body) {
    parallel_for(count, [] { return 0; }, [&](long long i, int &) { body(i); });
}

// Column-major dense matrix (types.hpp:13).
struct Mat {
    int rows = 0, cols = 0;
    Vec a;
    Mat() = default;
    Mat(int r, int c) : rows(r), cols(c), a(static_cast<size_t>(r) * c, 0.0) {}
    double &operator()(int i, int j) { return a[static_cast<size_t>(j) * rows + i]; }
    double operator()(int i, int j) const { return a[static_cast<size_t>(j) * rows + i]; }
    double *col(int j) { return a.data() + static_cast<size_t>(j) * rows; }
    const double *col(int j) const { return a.data() + static_cast<size_t>(j) * rows; }
    void zero() { std::fill(a.begin(), a.end(), 0.0); }
};

// Config defaults and flags: config.hpp:7-18, config.cpp:6-122.
struct Config {
    int nx = 8, ny = 8, degree = 3, quadrature = 0;
    int restart = 30, max_linear = 1500, max_newton = 30, steps = 0;
    int newton_steps = 0, warmup_updates = 0;
    double cfl = 2., final_time = .2, safety = 1.1;
    double nonlinear_tolerance = 1e-12, linear_tolerance = 1e-12;
    bool direct = false, colamd = false, timing_only = false, check_update = false, save_solution = true;
    // Extension beside the reference's options: --precond block-ilu0 with
    // --ordering redblack runs the DG block ILU(0) of block_ilu.cpp in
    // red-black order (the reference's 2D parser rejects block-ilu0).
    bool ilu_red_black = false;
    std::string integrator = "ssp-dirk3", method = "both", testcase = "diagonal", prefix = "output/run_2d";
};
Config parse(int argc, const char *const *argv);  // config.cpp:23-122
void validate(const Config &c);                    // config.cpp:6-22

// Gauss rule + nodal basis: nodal_basis.hpp:13-65.
struct Basis {
    int nodes, points;
    Vec x, w, qx, qw, left, right, error_x, error_w;
    Mat values, weighted_derivatives, error_values, projection, derivative;
    static void gauss(int n, Vec &x, Vec &w);
    Vec evaluate(double z) const;
    Basis(int degree, int quadrature = 0);
};

// Gas physics: physics.cpp:4-70.
struct Primitive { double inverse_density, pressure, enthalpy, velocity[2]; };
using State = std::array<double, NVAR>;
using Jac = std::array<double, 16>;  // row-major A[v*4+w] = dF_v/dU_w
struct Gas {
    double gamma = 3.0;
    double pressure(const State &u) const;
    Primitive primitive(const State &u) const;
    State flux(const State &u, const Primitive &p, int axis) const;
    State flux(const State &u, int axis) const { return flux(u, primitive(u), axis); }
    Jac jacobian(const State &u, int axis) const;
    Jac cached_jacobian(double u, double v, double H, int axis) const;
};
struct Speed { double a, b; };
Speed einfeldt(const State &l, const State &r, const Gas &gas, double safety, int axis);

// Test cases: example.cpp:4-39.
struct Example {
    Gas gas;
    std::string testcase;
    double cosine, sine, lx, ly;
    explicit Example(const std::string &name, double degrees = 45.);
    State exact(double x, double y, double t) const;
    double shock_time() const;
};

struct Layout {
    int nx, ny, nodes, cells;
    int plane() const { return nodes * nodes; }
    int block_size() const { return NVAR * plane(); }
    int size() const { return cells * block_size(); }
    int local(int v, int node) const { return v * plane() + node; }
    size_t index(int cell, int v, int node) const { return static_cast<size_t>(cell) * block_size() + local(v, node); }
};

// Sum factorisation: tensor_product.cpp:4-73.
struct TensorProduct {
    const Basis *b;
    double scale_x, scale_y;
    TensorProduct(const Basis &basis, double hx, double hy) : b(&basis), scale_x(.5 * hy), scale_y(.5 * hx) {}
    void interpolate(const double *nodal, double *points) const;  // [plane x NVAR] -> [Q^2 x NVAR]
    void subtract_weak(const double *fx, const double *fy, double *residual) const;
};

// Discretisation: discretization.cpp:4-299 (line numbers of the .cpp file).
class Problem {
  public:
    Config config;
    Basis basis;
    Layout layout;
    Example example;
    double hx, hy, step_size = 0;
    Vec old, mass, inverse_mass, trace_mass, inverse_trace_mass;
    std::vector<Speed> speeds;
    TensorProduct tensor;
    explicit Problem(const Config &c);
    int neighbor(int cell, int side) const;
    std::array<int, NFACES> faces(int cell) const;
    int face_width() const { return basis.nodes * NVAR; }
    int face_count() const { return NDIM * layout.cells; }
    int trace_size() const { return face_count() * face_width(); }
    int local_trace_size() const { return NFACES * face_width(); }
    double face_weight(int face, int q) const;
    int face_node(int side, int tangent, int normal) const;
    void face_values(const double *nodal, int side, double *values) const;
    void lift(int side, const double *values, double *nodal, double scale = 1) const;
    // r += the lifts of face values [face][face_width] into both cells of each face
    void lift_faces(const Vec &face_values, Vec &r) const;
    void initial(Vec &u) const;
    void seeded(Vec &u, uint64_t seed) const;  // build addition: seeded admissible state
    double freeze(const Vec &u);
    void trace(const Vec &u, Vec &lambda) const;
    void volume_residual(const Vec &u, Vec &r) const;
    void dg_residual(const Vec &u, Vec &r) const;
    void hybrid_residual(const Vec &u, const Vec &lambda, Vec &r, Vec &g) const;
    bool admissible(const Vec &u) const;
    double residual_norm(const Vec &r) const;
    State integral(const Vec &u) const;
    State error(const Vec &u, double time) const;
    void limit_cell(double *nodal) const;  // initial_and_error.cpp:19-51 positivity scaling

  private:
    bool physical(const double *states, int rows) const;
};

class VolumeJacobian {  // volume_jacobian.cpp:3-126
  public:
    explicit VolumeJacobian(const Problem &p);
    void refresh(const Vec &state);
    void apply(const Vec &direction, Vec &result) const;
    void block(int cell, Mat &result) const;
    Mat primitives;  // [points, 3*cells]
    double inverse_step = 0;

  private:
    const Problem &p_;
    bool ready_ = false;
    void point_action(int cell, const double *direction, double *fx, double *fy) const;
};

class FaceJacobian {  // face_jacobian.cpp:3-110
  public:
    explicit FaceJacobian(const Problem &p);
    void refresh(const Vec &state);
    void hll_action(int face, const double *l, const double *r, double *flux) const;
    Jac hll(int face, int q, bool left) const;
    Jac hdg(int face, int q, bool left) const;
    double tau(int face, int q, bool left) const { return left ? tau_left(q, face) : tau_right(q, face); }
    void cell_c_many(int cell, const Mat &directions, int columns, Mat &result) const;
    Vec diagonal;
    Mat primitives_left, primitives_right, tau_left, tau_right, weights, blend_left, blend_right, penalty;

  private:
    const Problem &p_;
    Jac physical(int face, int q, bool left) const;
};

class DGOperator {  // dg_operator.cpp:3-48
  public:
    explicit DGOperator(const Problem &p) : volume(p), faces(p), p_(p) {}
    void refresh(const Vec &state) { volume.refresh(state); faces.refresh(state); }
    void apply(const Vec &direction, Vec &result) const;
    void block(int row, int col, Mat &result) const;
    // Build addition: J^T w by transposing each assembled 5-point block row.
    void apply_transpose_blocks(const Vec &w, Vec &z) const;
    VolumeJacobian volume;
    FaceJacobian faces;

  private:
    const Problem &p_;
};

// Partial-pivot LU (dense_lu.hpp:10-46; Eigen PartialPivLU semantics: max |a|,
// first index on ties, LAPACK row interchanges across the full row).
struct DenseLU {
    int n = 0;
    Mat lu;
    std::vector<int> rowsrc;  // (P b)[i] = b[rowsrc[i]]
    void compute(const Mat &a);   // "Singular or nonfinite LU factors"
    bool factor(const Mat &a);    // returns false when singular / non-finite
    void solve_inplace(double *x) const;  // x <- A^-1 x
    void solve_multiple(const Mat &b, int cols, Mat &x) const;
};

class BlockJacobi {  // block_jacobi.hpp:12-39
  public:
    BlockJacobi(int blocks, int width) : blocks_(blocks), width_(width), lus_(blocks) {}
    template <class Provider> void refresh(Provider &&provider) {
        parallel_for(blocks_, [&] { return Mat(width_, width_); }, [&](long long i, Mat &block) {
            provider(static_cast<int>(i), static_cast<int>(i), block);
            require(lus_[i].factor(block), "Singular block Jacobi factor");
        });
    }
    void apply(const Vec &rhs, Vec &result) const;
    const DenseLU &lu(int i) const { return lus_[i]; }

  private:
    int blocks_, width_;
    std::vector<DenseLU> lus_;
};

// block_ilu.cpp:3-84 with block_graph.cpp:7-44: block ILU(0) on the permuted
// block graph (no fill). order[new] = old row; the graph rows hold the
// diagonal. Refresh asks the provider for every pattern block in the original
// numbering, factorizes in the permuted order and keeps L (unit lower, strictly
// lower slots), U (upper slots) and the LU of each updated diagonal block.
class BlockILU0 {
  public:
    BlockILU0(const std::vector<std::vector<int>> &adjacency, int width, std::vector<int> order);
    template <class Provider> void refresh(Provider &&provider) {
        for (int row = 0; row < rows_; ++row)
            for (int p = row_ptr_[row]; p < row_ptr_[row + 1]; ++p) {
                blocks_[p] = Mat(width_, width_);
                provider(order_[row], order_[col_idx_[p]], blocks_[p]);
            }
        factorize();
    }
    void apply(const Vec &rhs, Vec &result) const;
    int rows() const { return rows_; }
    const std::vector<int> &row_ptr() const { return row_ptr_; }
    const std::vector<int> &col_idx() const { return col_idx_; }
    const std::vector<int> &order() const { return order_; }
    const Mat &block(int slot) const { return blocks_[slot]; }
    const DenseLU &diagonal(int row) const { return diagonal_[row]; }

  private:
    int rows_, width_;
    std::vector<int> order_, row_ptr_, col_idx_, diagonal_slots_, update_offsets_, update_targets_, update_upper_;
    std::vector<Mat> blocks_;
    std::vector<DenseLU> diagonal_;
    mutable Vec work_;
    int find(int row, int column) const;
    void factorize();
};
// X <- X A^-1 with the LU of A, one row of X at a time (dense_lu.hpp:38-45).
void right_solve(const DenseLU &lu, Mat &x);
// Red-black order of an nx x ny grid: cells with x + y even, then odd, each in
// natural order (a two-colour ordering; not one of the reference's orderings).
std::vector<int> red_black_order(int nx, int ny);

struct LinearResult { int iterations; double relative_residual; };
using Action = std::function<void(const Vec &, Vec &)>;  // gmres.hpp:5-14

class GMRES {  // gmres.cpp:5-86
  public:
    GMRES(int n, int restart, const Vec &weights);
    LinearResult solve(const Action &A, const Action &M, const Vec &rhs, Vec &x, int limit, double tolerance);
    const Vec &residual() const { return residual_; }

  private:
    int n_, restart_;
    Mat basis_, hessenberg_;
    Vec cosine_, sine_, rotated_rhs_, coefficients_, work_, residual_, column_, weights_, projection_;
    double dot(const Vec &a, const Vec &b) const;
    double norm(const Vec &a) const { return std::sqrt(std::max(0.0, dot(a, a))); }
};

class HybridOperator {  // hybrid.cpp:3-121
  public:
    explicit HybridOperator(const Problem &p);
    void refresh_projected(const Vec &state, const Vec &dg_residual);
    void apply(const Vec &direction, Vec &result) const;
    void block(int row, int col, Mat &result) const;
    void recover(const Vec &trace_direction, Vec &volume_direction) const;
    void recovered_residual_from_schur(const Vec &error, Vec &result) const;
    const Vec &rhs() const { return rhs_; }
    VolumeJacobian volume;
    FaceJacobian faces_jacobian;
    Vec x_, schur_, y_;  // X [cells][m x t], S [cells][t x t], y [cells][m]

  private:
    const Problem &p_;
    Vec volume_residual_, skeleton_residual_, rhs_;
    void assemble_cell(int cell, Mat &a, Mat &local_rhs) const;
    void condense(const Vec &state);
};

// Dense global reference: direct.hpp:16-35 with BlockGraph::cartesian /
// Problem::trace_graph connectivity (block_graph.cpp:20-35, discretization.cpp:290-298).
class DirectSolver {
  public:
    DirectSolver(int blocks, int width) : blocks_(blocks), width_(width), matrix_(blocks * width, blocks * width) {}
    template <class Provider> void refresh(const std::vector<std::vector<int>> &graph, Provider &&provider) {
        matrix_.zero();
        Mat block(width_, width_);
        for (int row = 0; row < blocks_; ++row)
            for (int col : graph[row]) {
                provider(row, col, block);
                for (int j = 0; j < width_; ++j)
                    for (int i = 0; i < width_; ++i) matrix_(row * width_ + i, col * width_ + j) = block(i, j);
            }
        lu_.compute(matrix_);
    }
    LinearResult solve(const Vec &rhs, Vec &x, const Vec &weights) const;
    const Mat &matrix() const { return matrix_; }

  private:
    int blocks_, width_;
    Mat matrix_;
    DenseLU lu_;
};

// time_integrator.hpp:9-20
struct UpdateStatistics {
    int step, stage, iteration, krylov_iterations;
    double time, stage_time, dt, cfl, residual_physical_rms, damping;
    double linear_relative_residual, dg_linear_relative_residual, iteration_s;
};
struct RunStatistics {
    double final_time = 0, conservation_drift = 0, seconds = 0, max_linear_residual = 0;
    double max_dg_linear_residual = 0, minimum_step = 0, maximum_step = 0;
    double measured_newton_seconds = 0;
    int steps = 0, newton_iterations = 0, krylov_iterations = 0, measured_updates = 0;
    std::vector<UpdateStatistics> updates;
};

enum class Method { DG, HDG };
class TimeIntegrator {  // time_integrator.cpp:20-213
  public:
    TimeIntegrator(Problem &p, Method method);
    RunStatistics advance(Vec &solution);
    std::vector<std::vector<int>> graph;

  private:
    Problem &p_;
    Method method_;
    double dual_norm(const Vec &x) const;
    double state_norm(const Vec &x) const;
};

std::vector<std::vector<int>> cartesian_graph(int nx, int ny);
std::vector<std::vector<int>> trace_graph(const Problem &p);

// Counter-based generator for seeded inputs (build addition, no reference
// counterpart): splitmix64 finaliser on seed + (index+1) * golden ratio.
uint64_t splitmix64(uint64_t z);
double uniform01(uint64_t seed, uint64_t index);            // (z >> 11) * 2^-53
double normal01(uint64_t seed, uint64_t index);             // Box-Muller on two uniforms
}  // namespace oracle
