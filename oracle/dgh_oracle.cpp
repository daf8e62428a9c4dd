// CPU FP64 oracle: from-scratch restatement of the reference 2D DG / HDG path.
// TEST INFRASTRUCTURE ONLY (see dgh_oracle.hpp). Reference paths below are
// relative to the reference repository root.
//
// What pins each part (tests/test_oracle_*.py):
//   Gauss rule / basis  : closed forms (sum w = 2, exactness x^p p <= 2n-1, D on
//                         polynomials) -- nodal_basis.hpp:13-65
//   physics             : cached == analytic Jacobian, central FD of the flux
//                         (check_2d.cpp:26-36)
//   sum factorisation   : brute-force quadruple sum (check_2d.cpp:37-51)
//   residual/freeze/IC  : golden RMS0 and dt of the reference's archived runs
//                         (4x4 k=2; 256^2 k=8 `diagonal`)
//   Jv / blocks         : central FD of the residual, block vs matrix-free
//                         (check_2d.cpp:52-71)
//   LU / block Jacobi   : forced-permutation exact solve (check_2d.cpp:12-25)
//   GMRES / Newton      : golden per-update RMS + Krylov counts (4x4 k=2 BE and
//                         DIRK3, 32^2 k=2 DIRK3 L2 errors)
//   HDG                 : projected residual identity, recovered residual
//                         identity, Schur blocks vs matrix-free, DG == HDG
//                         Newton RMS sequences of the archived runs
//   Bit-level agreement with Eigen's blocked PartialPivLU rounding is parity
//   unpinned (the reference cannot be built here: Eigen 3.4 is absent).
#include "dgh_oracle.hpp"

#include <chrono>
#include <memory>
#include <cstring>
#include <limits>

namespace oracle {

// Global reductions over millions of terms: 256-term blocks summed naively,
// block sums accumulated with Neumaier compensation. Keeps the oracle's norms
// at ~1e-15 relative (the reference's vectorised Eigen reductions are as
// accurate; a plain running sum over 2e7 terms is not).
struct Accumulator {
    double block = 0, sum = 0, comp = 0;
    int count = 0;
    void flush() {
        const double t = sum + block;
        comp += std::abs(sum) >= std::abs(block) ? (sum - t) + block : (block - t) + sum;
        sum = t;
        block = 0;
        count = 0;
    }
    void add(double x) {
        block += x;
        if (++count == 256) flush();
    }
    double result() {
        flush();
        return sum + comp;
    }
    // a whole block's sum, summed elsewhere in the same order add() would
    void add_block(double block_sum) {
        block = block_sum;
        flush();
    }
};

// The Accumulator's sum of term(0..count-1): each 256-term block summed in
// order (one thread per block), the block sums compensated in block order.
template <class Term> double compensated_sum(size_t count, Term term) {
    const long long blocks = static_cast<long long>((count + 255) / 256);
    Vec part(static_cast<size_t>(blocks));
    parallel_for(blocks, [&](long long b) {
        const size_t begin = static_cast<size_t>(b) * 256, end = std::min(count, begin + 256);
        Accumulator s;  // one block: its result is the plain in-order sum
        for (size_t i = begin; i < end; ++i) s.add(term(i));
        part[static_cast<size_t>(b)] = s.result();
    });
    Accumulator acc;
    for (double x : part) acc.add_block(x);
    return acc.result();
}

// ----------------------------------------------------------------- config
// dg_hdg_2d/src/config.cpp:6-22
void validate(const Config &c) {
    require(c.nx >= 3 && c.ny >= 3 && c.degree >= 0 && c.degree <= 8 && c.quadrature >= 0, "Invalid mesh or degree");
    require(c.testcase == "diagonal" || c.testcase == "wave" || c.testcase == "axis" || c.testcase == "seeded",
            "Invalid testcase");
    require(c.quadrature == 0 || c.quadrature >= c.degree + 1, "Quadrature must be >= degree+1");
    require(c.restart > 0 && c.max_linear > 0 && c.max_newton > 0 && c.steps >= 0, "Invalid iteration limit");
    require(c.newton_steps >= 0 && c.warmup_updates >= 0, "Invalid timing update count");
    require(c.integrator == "be" || c.integrator == "ssp-dirk3", "Invalid integrator");
    require(std::isfinite(c.cfl) && c.cfl > 0 && std::isfinite(c.final_time) && c.final_time >= 0,
            "CFL must be positive and final time nonnegative");
    require(std::isfinite(c.safety) && c.safety >= 1 && std::isfinite(c.nonlinear_tolerance) &&
                c.nonlinear_tolerance > 0 && std::isfinite(c.linear_tolerance) && c.linear_tolerance > 0 &&
                c.linear_tolerance < 1,
            "Invalid solver tolerance or wave-speed safety");
    require(c.method == "dg" || c.method == "hdg" || c.method == "both", "Method must be dg, hdg, or both");
}

// dg_hdg_2d/src/config.cpp:23-122 (flags the oracle does not need are parsed and ignored)
Config parse(int argc, const char *const *argv) {
    Config c;
    bool time_seen = false, steps_seen = false, ilu = false, red_black = false;
    for (int i = 1; i < argc; ++i) {
        std::string option = argv[i];
        auto value = [&]() {
            require(i + 1 < argc, "Missing option value");
            return std::string(argv[++i]);
        };
        auto integer = [&]() {
            auto text = value();
            size_t end;
            int n = std::stoi(text, &end);
            require(end == text.size(), "Invalid integer");
            return n;
        };
        auto real = [&]() {
            auto text = value();
            size_t end;
            double n = std::stod(text, &end);
            require(end == text.size(), "Invalid number");
            return n;
        };
        if (option == "--nx") c.nx = integer();
        else if (option == "--ny") c.ny = integer();
        else if (option == "--testcase") c.testcase = value();
        else if (option == "--no-save-solution") c.save_solution = false;
        else if (option == "--newton-steps") c.newton_steps = integer();
        else if (option == "--warmup-updates") c.warmup_updates = integer();
        else if (option == "--integrator") c.integrator = value();
        else if (option == "--timing-only") c.timing_only = true;
        else if (option == "--check-update") c.check_update = true;
        else if (option == "--k") c.degree = integer();
        else if (option == "--quad") c.quadrature = integer();
        else if (option == "--cfl") c.cfl = real();
        else if (option == "--T") { c.final_time = real(); time_seen = true; }
        else if (option == "--steps") { c.steps = integer(); steps_seen = true; }
        else if (option == "--safety") c.safety = real();
        else if (option == "--method") c.method = value();
        else if (option == "--prefix") c.prefix = value();
        else if (option == "--ntol") c.nonlinear_tolerance = real();
        else if (option == "--ltol") c.linear_tolerance = real();
        else if (option == "--restart") c.restart = integer();
        else if (option == "--max-linear") c.max_linear = integer();
        else if (option == "--max-newton") c.max_newton = integer();
        else if (option == "--linear") {
            auto name = value();
            require(name == "direct" || name == "gmres", "Invalid linear solver");
            c.direct = name == "direct";
        } else if (option == "--ordering") {
            auto order = value();
            require(order == "natural" || order == "colamd" || order == "redblack", "Invalid ordering");
            c.colamd = order == "colamd";
            red_black = order == "redblack";
        } else if (option == "--precond") {
            auto name = value();
            require(name == "block-jacobi" || name == "block-ilu0", "2D currently supports block-jacobi only");
            ilu = name == "block-ilu0";
        } else
            throw std::runtime_error("Unknown option: " + option);
    }
    require(!ilu || red_black, "2D currently supports block-jacobi only");
    require(!red_black || ilu, "Invalid ordering");
    require(!ilu || c.method == "dg", "Block ILU(0) is built for --method dg");
    c.ilu_red_black = ilu;
    require(!(time_seen && steps_seen), "Choose either --T or --steps");
    require(!steps_seen || c.steps > 0, "steps must be positive");
    validate(c);
    return c;
}

// ------------------------------------------------------------------ basis
// Gauss-Legendre rule by Newton iteration on P_n (the reference uses fastgl,
// vendor/fastgl/fastgl.h:46; nodal_basis.hpp:13-25 fixes ordering/symmetry).
void Basis::gauss(int n, Vec &x, Vec &w) {
    require(n > 0, "Gauss rule requires a positive point count");
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    const double pi = std::acos(-1.0);
    for (int k = 1; k <= (n + 1) / 2; ++k) {
        double z = std::cos(pi * (k - 0.25) / (n + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int l = 2; l <= n; ++l) {
                double p2 = ((2.0 * l - 1.0) * z * p1 - (l - 1.0) * p0) / l;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p0 = 1.0; p1 = z; }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            double dz = p1 / dp;
            z -= dz;
            if (std::abs(dz) <= 1e-16 * std::max(1.0, std::abs(z)) && it > 2) break;
        }
        {
            double p0 = 1.0, p1 = z;
            for (int l = 2; l <= n; ++l) {
                double p2 = ((2.0 * l - 1.0) * z * p1 - (l - 1.0) * p0) / l;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
        }
        x[k - 1] = -z;
        x[n - k] = z;
        w[k - 1] = w[n - k] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
    if (n % 2) x[n / 2] = 0;
}

// nodal_basis.hpp:26-32
Vec Basis::evaluate(double z) const {
    Vec v(nodes, 1.0);
    for (int i = 0; i < nodes; ++i)
        for (int j = 0; j < nodes; ++j)
            if (i != j) v[i] *= (z - x[j]) / (x[i] - x[j]);
    return v;
}

// nodal_basis.hpp:34-65
Basis::Basis(int degree, int quadrature) : nodes(degree + 1), points(quadrature ? quadrature : degree + 1) {
    require(nodes > 0 && points >= nodes, "Require quadrature >= degree+1");
    gauss(nodes, x, w);
    gauss(points, qx, qw);
    Vec bary(nodes, 1.0);
    for (int i = 0; i < nodes; ++i)
        for (int j = 0; j < nodes; ++j)
            if (i != j) bary[i] /= x[i] - x[j];
    derivative = Mat(nodes, nodes);
    for (int i = 0; i < nodes; ++i) {
        for (int j = 0; j < nodes; ++j)
            if (i != j) derivative(i, j) = bary[j] / (bary[i] * (x[i] - x[j]));
        double s = 0;
        for (int j = 0; j < nodes; ++j) s += derivative(i, j);
        derivative(i, i) = -s;
    }
    values = Mat(points, nodes);
    for (int q = 0; q < points; ++q) {
        Vec e = evaluate(qx[q]);
        for (int j = 0; j < nodes; ++j) values(q, j) = points == nodes ? (q == j ? 1.0 : 0.0) : e[j];
    }
    weighted_derivatives = Mat(points, nodes);
    for (int q = 0; q < points; ++q)
        for (int j = 0; j < nodes; ++j) {
            double s = 0;
            for (int k = 0; k < nodes; ++k) s += (qw[q] * values(q, k)) * derivative(k, j);
            weighted_derivatives(q, j) = s;
        }
    left = evaluate(-1);
    right = evaluate(1);
    gauss(std::max(20, 2 * nodes + 3), error_x, error_w);
    const int eq = static_cast<int>(error_x.size());
    error_values = Mat(eq, nodes);
    for (int q = 0; q < eq; ++q) {
        Vec e = evaluate(error_x[q]);
        for (int j = 0; j < nodes; ++j) error_values(q, j) = e[j];
    }
    projection = Mat(nodes, eq);
    for (int i = 0; i < nodes; ++i)
        for (int q = 0; q < eq; ++q) projection(i, q) = ((1.0 / w[i]) * error_values(q, i)) * error_w[q];
}

// ---------------------------------------------------------------- physics
// dg_hdg_2d/src/physics.cpp:4-6
double Gas::pressure(const State &u) const {
    return (gamma - 1) * (u[3] - .5 * (u[1] * u[1] + u[2] * u[2]) / u[0]);
}
// physics.cpp:7-14
Primitive Gas::primitive(const State &u) const {
    Primitive p;
    p.inverse_density = 1 / u[0];
    p.velocity[0] = u[1] * p.inverse_density;
    p.velocity[1] = u[2] * p.inverse_density;
    p.pressure = (gamma - 1) * (u[3] - .5 * (u[1] * p.velocity[0] + u[2] * p.velocity[1]));
    p.enthalpy = (u[3] + p.pressure) * p.inverse_density;
    return p;
}
// physics.cpp:15-22
State Gas::flux(const State &u, const Primitive &p, int axis) const {
    State f;
    f[0] = u[1 + axis];
    f[1] = u[1] * p.velocity[axis];
    f[2] = u[2] * p.velocity[axis];
    f[1 + axis] += p.pressure;
    f[3] = (u[3] + p.pressure) * p.velocity[axis];
    return f;
}
// physics.cpp:23-41 (analytic derivative of the flux)
Jac Gas::jacobian(const State &u, int axis) const {
    const Primitive p = primitive(u);
    double dp[4] = {.5 * (gamma - 1) * (p.velocity[0] * p.velocity[0] + p.velocity[1] * p.velocity[1]),
                    -(gamma - 1) * p.velocity[0], -(gamma - 1) * p.velocity[1], gamma - 1};
    double dv[4] = {-p.velocity[axis] * p.inverse_density, 0, 0, 0};
    dv[1 + axis] = p.inverse_density;
    Jac a{};
    a[0 * 4 + 1 + axis] = 1;
    for (int d = 0; d < 2; ++d) {
        int m = 1 + d;
        for (int w = 0; w < 4; ++w) a[m * 4 + w] = u[m] * dv[w];
        a[m * 4 + m] += p.velocity[axis];
    }
    for (int w = 0; w < 4; ++w) a[(1 + axis) * 4 + w] += dp[w];
    for (int w = 0; w < 4; ++w) a[3 * 4 + w] = (u[3] + p.pressure) * dv[w] + p.velocity[axis] * dp[w];
    a[3 * 4 + 3] += p.velocity[axis];
    return a;
}
// physics.cpp:42-57 (Jacobian from the cached (u, v, H))
Jac Gas::cached_jacobian(double u, double v, double H, int axis) const {
    const double gm = gamma - 1., speed2 = u * u + v * v;
    Jac a{};
    if (axis == 0) {
        a[0 * 4 + 1] = 1.;
        a[1 * 4 + 0] = .5 * (gamma - 3) * u * u + .5 * gm * v * v; a[1 * 4 + 1] = (3 - gamma) * u;
        a[1 * 4 + 2] = -gm * v; a[1 * 4 + 3] = gm;
        a[2 * 4 + 0] = -u * v; a[2 * 4 + 1] = v; a[2 * 4 + 2] = u; a[2 * 4 + 3] = 0.;
        a[3 * 4 + 0] = u * (.5 * gm * speed2 - H); a[3 * 4 + 1] = H - gm * u * u;
        a[3 * 4 + 2] = -gm * u * v; a[3 * 4 + 3] = gamma * u;
    } else {
        a[0 * 4 + 2] = 1.;
        a[1 * 4 + 0] = -u * v; a[1 * 4 + 1] = v; a[1 * 4 + 2] = u; a[1 * 4 + 3] = 0.;
        a[2 * 4 + 0] = .5 * gm * u * u + .5 * (gamma - 3) * v * v; a[2 * 4 + 1] = -gm * u;
        a[2 * 4 + 2] = (3 - gamma) * v; a[2 * 4 + 3] = gm;
        a[3 * 4 + 0] = v * (.5 * gm * speed2 - H); a[3 * 4 + 1] = -gm * u * v;
        a[3 * 4 + 2] = H - gm * v * v; a[3 * 4 + 3] = gamma * v;
    }
    return a;
}
// physics.cpp:58-70 (Einfeldt / Roe HLL speeds with safety factor)
Speed einfeldt(const State &l, const State &r, const Gas &gas, double safety, int axis) {
    auto pl = gas.primitive(l), pr = gas.primitive(r);
    require(l[0] > 0 && r[0] > 0 && pl.pressure > 0 && pr.pressure > 0, "Invalid frozen face state");
    double sl = std::sqrt(l[0]), sr = std::sqrt(r[0]);
    double vel[2] = {(sl * pl.velocity[0] + sr * pr.velocity[0]) / (sl + sr),
                     (sl * pl.velocity[1] + sr * pr.velocity[1]) / (sl + sr)};
    double H = (sl * pl.enthalpy + sr * pr.enthalpy) / (sl + sr);
    double sound = std::sqrt((gas.gamma - 1) * (H - .5 * (vel[0] * vel[0] + vel[1] * vel[1])));
    double cl = std::sqrt(gas.gamma * pl.pressure * pl.inverse_density);
    double cr = std::sqrt(gas.gamma * pr.pressure * pr.inverse_density);
    return {safety * std::max({0., -pl.velocity[axis] + cl, -vel[axis] + sound}),
            safety * std::max({0., pr.velocity[axis] + cr, vel[axis] + sound})};
}

// ---------------------------------------------------------------- example
// dg_hdg_2d/include/euler/example.hpp:9-13, src/example.cpp:4-39
Example::Example(const std::string &name, double degrees)
    : testcase(name), cosine(std::cos(degrees * std::acos(-1.) / 180.)),
      sine(std::sin(degrees * std::acos(-1.) / 180.)), lx(testcase == "diagonal" ? 2. / cosine : 2.),
      ly(testcase == "diagonal" ? 2. / sine : 2.) {}
double Example::shock_time() const {
    if (testcase == "wave" || testcase == "seeded") return std::numeric_limits<double>::infinity();
    return 4 / ((gas.gamma + 1) * std::acos(-1.));
}
State Example::exact(double x, double y, double t) const {
    double g = gas.gamma, pi = std::acos(-1.);
    State u;
    if (testcase == "seeded") {
        u.fill(std::numeric_limits<double>::quiet_NaN());
        return u;
    }
    if (testcase == "wave") {
        double rho = 1 + .2 * std::sin(pi * (x - t)) * std::sin(pi * (y - .5 * t));
        u = {rho, rho, .5 * rho, 1 / (g - 1) + .625 * rho};
        return u;
    }
    require(t < shock_time(), "Simple wave reaches shock time");
    bool diagonal = testcase == "diagonal";
    if (diagonal) x = cosine * x + sine * y;
    double a = (3 - g) / 2, b = (g + 1) / 4;
    double lo = x - t * (a + b), hi = x - t * (a - b), xi = x - t * a;
    for (int it = 0; it < 80; ++it) {
        double f = xi + t * (a + b * std::sin(pi * xi)) - x;
        if (std::abs(f) < 2e-15) break;
        if (f > 0) hi = xi; else lo = xi;
        double z = xi - f / (1 + t * b * pi * std::cos(pi * xi));
        xi = z > lo && z < hi ? z : .5 * (lo + hi);
    }
    double jm = std::sin(pi * xi), velocity = .5 * (2 + jm), sound = (g - 1) * (2 - jm) / 4;
    double rho = std::pow(sound * sound / g, 1 / (g - 1));
    u[0] = rho;
    u[1] = rho * velocity * (diagonal ? cosine : 1.);
    u[2] = diagonal ? rho * velocity * sine : 0.;
    u[3] = std::pow(rho, g) / (g - 1) + .5 * rho * velocity * velocity;
    return u;
}

// ----------------------------------------------------------- tensor product
// tensor_product.cpp:11-22 (points = values * nodal * values^T per variable)
void TensorProduct::interpolate(const double *nodal, double *points) const {
    const int n = b->nodes, Q = b->points;
    if (Q == n) {
        std::memcpy(points, nodal, sizeof(double) * n * n * NVAR);
        return;
    }
    Vec tmp(static_cast<size_t>(Q) * n);
    for (int v = 0; v < NVAR; ++v) {
        const double *in = nodal + v * n * n;
        double *out = points + v * Q * Q;
        for (int j = 0; j < n; ++j)
            for (int q = 0; q < Q; ++q) {
                double s = 0;
                for (int k = 0; k < n; ++k) s += in[k + n * j] * b->values(q, k);
                tmp[q + Q * j] = s;
            }
        for (int qb = 0; qb < Q; ++qb)
            for (int qa = 0; qa < Q; ++qa) {
                double s = 0;
                for (int k = 0; k < n; ++k) s += b->values(qb, k) * tmp[qa + Q * k];
                out[qa + Q * qb] = s;
            }
    }
}
// tensor_product.cpp:23-73: r(i,j) -= (hy/2) w_j sum_q d(q,i) Fx(q,j) + (hx/2) w_i sum_q Fy(i,q) d(q,j)
void TensorProduct::subtract_weak(const double *fx, const double *fy, double *residual) const {
    const int n = b->nodes, Q = b->points;
    if (n == 1) return;
    const Mat &d = b->weighted_derivatives;
    if (Q == n) {
        for (int v = 0; v < NVAR; ++v) {
            const double *X = fx + v * n * n, *Y = fy + v * n * n;
            double *out = residual + v * n * n;
            for (int j = 0; j < n; ++j) {
                for (int i = 0; i < n; ++i) {
                    double wx = 0, wy = 0;
                    for (int q = 0; q < n; ++q) wx += d(q, i) * X[q + n * j];
                    for (int q = 0; q < n; ++q) wy += Y[i + n * q] * d(q, j);
                    out[i + n * j] -= (scale_x * b->w[j]) * wx;
                    out[i + n * j] -= scale_y * b->w[i] * wy;
                }
            }
        }
        return;
    }
    Vec scratch(static_cast<size_t>(n) * Q);
    for (int v = 0; v < NVAR; ++v) {
        const double *X = fx + v * Q * Q, *Y = fy + v * Q * Q;
        double *out = residual + v * n * n;
        for (int qb = 0; qb < Q; ++qb)
            for (int i = 0; i < n; ++i) {
                double s = 0;
                for (int k = 0; k < Q; ++k) s += X[k + Q * qb] * d(k, i);
                scratch[i + n * qb] = s;
            }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                double s = 0;
                for (int k = 0; k < Q; ++k) s += (b->qw[k] * b->values(k, j)) * scratch[i + n * k];
                out[i + n * j] -= scale_x * s;
            }
        for (int qb = 0; qb < Q; ++qb)
            for (int i = 0; i < n; ++i) {
                double s = 0;
                for (int k = 0; k < Q; ++k) s += Y[k + Q * qb] * (b->qw[k] * b->values(k, i));
                scratch[i + n * qb] = s;
            }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                double s = 0;
                for (int k = 0; k < Q; ++k) s += d(k, j) * scratch[i + n * k];
                out[i + n * j] -= scale_y * s;
            }
    }
}

// ---------------------------------------------------------------- problem
// discretization.cpp:4-25
Problem::Problem(const Config &c)
    : config(c), basis(c.degree, c.quadrature), layout{c.nx, c.ny, c.degree + 1, c.nx * c.ny},
      example(c.testcase), hx(example.lx / c.nx), hy(example.ly / c.ny), tensor(basis, hx, hy) {
    validate(c);
    const int n = basis.nodes;
    old.assign(layout.size(), 0.0);
    mass.assign(layout.size(), 0.0);
    for (int cell = 0; cell < layout.cells; ++cell)
        for (int v = 0; v < NVAR; ++v)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    mass[layout.index(cell, v, i + n * j)] = .25 * hx * hy * basis.w[i] * basis.w[j];
    trace_mass.assign(trace_size(), 0.0);
    for (int f = 0; f < face_count(); ++f)
        for (int v = 0; v < NVAR; ++v)
            for (int q = 0; q < n; ++q) trace_mass[f * face_width() + v * n + q] = face_weight(f, q);
    inverse_mass.resize(mass.size());
    for (size_t i = 0; i < mass.size(); ++i) inverse_mass[i] = 1.0 / mass[i];
    inverse_trace_mass.resize(trace_mass.size());
    for (size_t i = 0; i < trace_mass.size(); ++i) inverse_trace_mass[i] = 1.0 / trace_mass[i];
    speeds.assign(static_cast<size_t>(face_count()) * n, Speed{0, 0});
}
// discretization.cpp:26-31
int Problem::neighbor(int cell, int side) const {
    int x = cell % layout.nx, y = cell / layout.nx;
    if (side < 2) x = (x + (side == 0 ? layout.nx - 1 : 1)) % layout.nx;
    else y = (y + (side == 2 ? layout.ny - 1 : 1)) % layout.ny;
    return x + layout.nx * y;
}
// discretization.cpp:32-34
std::array<int, NFACES> Problem::faces(int cell) const {
    return {neighbor(cell, 0), cell, layout.cells + neighbor(cell, 2), layout.cells + cell};
}
// discretization.cpp:35-37
double Problem::face_weight(int face, int q) const { return .5 * (face < layout.cells ? hy : hx) * basis.w[q]; }
// discretization.cpp:38-40
int Problem::face_node(int side, int q, int i) const { return side < 2 ? i + basis.nodes * q : q + basis.nodes * i; }
// discretization.cpp:50-71 (restriction to side `side`; left/right shape = l(-1)/l(+1))
void Problem::face_values(const double *nodal, int side, double *values) const {
    const int n = basis.nodes;
    const Vec &e = side % 2 ? basis.right : basis.left;
    for (int v = 0; v < NVAR; ++v) {
        const double *sheet = nodal + v * n * n;
        for (int q = 0; q < n; ++q) {
            double s = 0;
            if (side < 2) for (int i = 0; i < n; ++i) s += sheet[i + n * q] * e[i];
            else for (int j = 0; j < n; ++j) s += sheet[q + n * j] * e[j];
            values[v * n + q] = s;
        }
    }
}
// discretization.cpp:61-102 (lift = outer product with the same shapes)
void Problem::lift(int side, const double *values, double *nodal, double scale) const {
    const int n = basis.nodes;
    const Vec &e = side % 2 ? basis.right : basis.left;
    for (int v = 0; v < NVAR; ++v) {
        double *sheet = nodal + v * n * n;
        const double *f = values + v * n;
        for (int q = 0; q < n; ++q)
            for (int i = 0; i < n; ++i) {
                if (side < 2) sheet[i + n * q] += (scale * e[i]) * f[q];
                else sheet[q + n * i] += (scale * f[q]) * e[i];
            }
    }
}

// The face loops of discretization.cpp:123-140 and dg_operator.cpp:10-21 visit
// face = axis * cells + cell in order and lift each face value into its cell
// (side 2 axis + 1, +1) and then into the +axis neighbour (side 2 axis, -1).
// Per cell that is: the two x faces in face order, then the two y faces in
// face order; gathering them per cell in that order gives the same sums.
void Problem::lift_faces(const Vec &face_values, Vec &r) const {
    const int m = layout.block_size(), fw = face_width(), C = layout.cells;
    parallel_for(C, [&](long long c) {
        const int cell = static_cast<int>(c);
        double *nodal = r.data() + static_cast<size_t>(cell) * m;
        for (int axis = 0; axis < NDIM; ++axis) {
            const int own = axis * C + cell, other = axis * C + neighbor(cell, 2 * axis);
            const double *fo = face_values.data() + static_cast<size_t>(own) * fw;
            const double *fn = face_values.data() + static_cast<size_t>(other) * fw;
            if (other < own) {
                lift(2 * axis, fn, nodal, -1);
                lift(2 * axis + 1, fo, nodal);
            } else {
                lift(2 * axis + 1, fo, nodal);
                lift(2 * axis, fn, nodal, -1);
            }
        }
    });
}
// discretization.cpp:111-122
void Problem::volume_residual(const Vec &u, Vec &r) const {
    r.resize(u.size());
    parallel_for(static_cast<long long>(u.size()), [&](long long i) { r[i] = mass[i] * (u[i] - old[i]) / step_size; });
    if (basis.nodes == 1) return;
    const int P = basis.points * basis.points, m = layout.block_size();
    struct Scratch { Vec points, fx, fy; };
    parallel_for(layout.cells, [&] { return Scratch{Vec(static_cast<size_t>(P) * NVAR), Vec(static_cast<size_t>(P) * NVAR), Vec(static_cast<size_t>(P) * NVAR)}; },
                 [&](long long c, Scratch &w) {
        const int cell = static_cast<int>(c);
        tensor.interpolate(u.data() + static_cast<size_t>(cell) * m, w.points.data());
        for (int q = 0; q < P; ++q) {
            State s = {w.points[q], w.points[P + q], w.points[2 * P + q], w.points[3 * P + q]};
            auto prim = example.gas.primitive(s);
            State a = example.gas.flux(s, prim, 0), b = example.gas.flux(s, prim, 1);
            for (int v = 0; v < NVAR; ++v) { w.fx[v * P + q] = a[v]; w.fy[v * P + q] = b[v]; }
        }
        tensor.subtract_weak(w.fx.data(), w.fy.data(), r.data() + static_cast<size_t>(cell) * m);
    });
}
// discretization.cpp:123-140
void Problem::dg_residual(const Vec &u, Vec &r) const {
    volume_residual(u, r);
    const int n = basis.nodes, m = layout.block_size(), fw = face_width();
    Vec fluxes(static_cast<size_t>(face_count()) * fw);
    parallel_for(face_count(), [&] { return std::make_pair(Vec(n * NVAR), Vec(n * NVAR)); },
                 [&](long long f, std::pair<Vec, Vec> &lr) {
        const int face = static_cast<int>(f), axis = face / layout.cells, cell = face % layout.cells;
        const int side = 2 * axis + 1, next = neighbor(cell, side);
        Vec &left = lr.first, &right = lr.second;
        double *flux = fluxes.data() + static_cast<size_t>(face) * fw;
        face_values(u.data() + static_cast<size_t>(cell) * m, side, left.data());
        face_values(u.data() + static_cast<size_t>(next) * m, side - 1, right.data());
        for (int q = 0; q < n; ++q) {
            State l, rr;
            for (int v = 0; v < NVAR; ++v) { l[v] = left[v * n + q]; rr[v] = right[v * n + q]; }
            auto s = speeds[static_cast<size_t>(face) * n + q];
            State fl = example.gas.flux(l, axis), fr = example.gas.flux(rr, axis);
            double scale = face_weight(face, q) / (s.a + s.b);
            for (int v = 0; v < NVAR; ++v)
                flux[v * n + q] = scale * (s.b * fl[v] + s.a * fr[v] + s.a * s.b * (l[v] - rr[v]));
        }
    });
    lift_faces(fluxes, r);
}
// discretization.cpp:141-162
void Problem::hybrid_residual(const Vec &u, const Vec &lambda, Vec &r, Vec &g) const {
    volume_residual(u, r);
    g.assign(trace_size(), 0.0);
    const int n = basis.nodes, m = layout.block_size(), fw = face_width();
    Vec left(n * NVAR), right(n * NVAR), flux(n * NVAR), other(n * NVAR);
    for (int axis = 0; axis < NDIM; ++axis)
        for (int cell = 0; cell < layout.cells; ++cell) {
            int side = 2 * axis + 1, next = neighbor(cell, side), face = axis * layout.cells + cell;
            face_values(u.data() + static_cast<size_t>(cell) * m, side, left.data());
            face_values(u.data() + static_cast<size_t>(next) * m, side - 1, right.data());
            for (int q = 0; q < n; ++q) {
                State l, rr, z;
                for (int v = 0; v < NVAR; ++v) {
                    l[v] = left[v * n + q];
                    rr[v] = right[v * n + q];
                    z[v] = lambda[static_cast<size_t>(face) * fw + v * n + q];
                }
                auto s = speeds[static_cast<size_t>(face) * n + q];
                double weight = face_weight(face, q);
                State fl = example.gas.flux(l, axis), fr = example.gas.flux(rr, axis);
                for (int v = 0; v < NVAR; ++v) {
                    flux[v * n + q] = weight * (fl[v] + s.a * (l[v] - z[v]));
                    other[v * n + q] = weight * (fr[v] + s.b * (z[v] - rr[v]));
                }
            }
            for (int k = 0; k < fw; ++k) g[static_cast<size_t>(face) * fw + k] = flux[k] - other[k];
            lift(side, flux.data(), r.data() + static_cast<size_t>(cell) * m);
            lift(side - 1, other.data(), r.data() + static_cast<size_t>(next) * m, -1);
        }
}
// discretization.cpp:163-195
double Problem::freeze(const Vec &u) {
    old = u;
    const int n = basis.nodes, m = layout.block_size();
    parallel_for(face_count(), [&] { return std::make_pair(Vec(n * NVAR), Vec(n * NVAR)); },
                 [&](long long f, std::pair<Vec, Vec> &lr) {
        const int face = static_cast<int>(f), axis = face / layout.cells, cell = face % layout.cells;
        const int side = 2 * axis + 1, next = neighbor(cell, side);
        Vec &left = lr.first, &right = lr.second;
        face_values(u.data() + static_cast<size_t>(cell) * m, side, left.data());
        face_values(u.data() + static_cast<size_t>(next) * m, side - 1, right.data());
        for (int q = 0; q < n; ++q) {
            State l, r;
            for (int v = 0; v < NVAR; ++v) { l[v] = left[v * n + q]; r[v] = right[v * n + q]; }
            speeds[static_cast<size_t>(face) * n + q] = einfeldt(l, r, example.gas, config.safety, axis);
        }
    });
    const int P = basis.points * basis.points;
    Vec cell_rate(static_cast<size_t>(layout.cells));  // the max over cells below is exact in any order
    parallel_for(layout.cells, [&] { return Vec(static_cast<size_t>(P) * NVAR); }, [&](long long c, Vec &points) {
        const int cell = static_cast<int>(c);
        double alpha[2] = {0, 0};
        auto fs = faces(cell);
        for (int side = 0; side < NFACES; ++side)
            for (int q = 0; q < n; ++q) {
                auto s = speeds[static_cast<size_t>(fs[side]) * n + q];
                alpha[side / 2] = std::max({alpha[side / 2], s.a, s.b});
            }
        tensor.interpolate(u.data() + static_cast<size_t>(cell) * m, points.data());
        for (int q = 0; q < P; ++q) {
            State s = {points[q], points[P + q], points[2 * P + q], points[3 * P + q]};
            auto prim = example.gas.primitive(s);
            require(s[0] > 0 && prim.pressure > 0, "Invalid CFL volume state");
            double sound = std::sqrt(example.gas.gamma * prim.pressure * prim.inverse_density);
            for (int axis = 0; axis < NDIM; ++axis)
                alpha[axis] = std::max(alpha[axis], std::abs(prim.velocity[axis]) + sound);
        }
        cell_rate[static_cast<size_t>(cell)] = alpha[0] / hx + alpha[1] / hy;
    });
    double rate = 0;
    for (double x : cell_rate) rate = std::max(rate, x);
    return rate;
}
// discretization.cpp:196-210 (trace projection Phi(U))
void Problem::trace(const Vec &u, Vec &lambda) const {
    lambda.assign(trace_size(), 0.0);
    const int n = basis.nodes, m = layout.block_size(), fw = face_width();
    Vec left(n * NVAR), right(n * NVAR);
    for (int axis = 0; axis < NDIM; ++axis)
        for (int cell = 0; cell < layout.cells; ++cell) {
            int side = 2 * axis + 1, next = neighbor(cell, side), face = axis * layout.cells + cell;
            face_values(u.data() + static_cast<size_t>(cell) * m, side, left.data());
            face_values(u.data() + static_cast<size_t>(next) * m, side - 1, right.data());
            for (int q = 0; q < n; ++q) {
                State l, r;
                for (int v = 0; v < NVAR; ++v) { l[v] = left[v * n + q]; r[v] = right[v * n + q]; }
                auto s = speeds[static_cast<size_t>(face) * n + q];
                State fl = example.gas.flux(l, axis), fr = example.gas.flux(r, axis);
                for (int v = 0; v < NVAR; ++v)
                    lambda[static_cast<size_t>(face) * fw + v * n + q] =
                        (s.a * l[v] + s.b * r[v] + fl[v] - fr[v]) / (s.a + s.b);
            }
        }
}
// discretization.cpp:211-219 (rows of a [rows x NVAR] column-major block)
bool Problem::physical(const double *states, int rows) const {
    for (int k = 0; k < rows * NVAR; ++k)
        if (!std::isfinite(states[k])) return false;
    for (int q = 0; q < rows; ++q) {
        State s = {states[q], states[rows + q], states[2 * rows + q], states[3 * rows + q]};
        if (s[0] <= 1e-13 || !(example.gas.pressure(s) > 1e-13)) return false;
    }
    return true;
}
// discretization.cpp:220-233
bool Problem::admissible(const Vec &u) const {
    if (static_cast<int>(u.size()) != layout.size()) return false;
    const int n = basis.nodes, m = layout.block_size(), P = basis.points * basis.points;
    std::vector<char> ok(static_cast<size_t>(layout.cells), 1);
    parallel_for(layout.cells, [&] { return std::make_pair(Vec(static_cast<size_t>(P) * NVAR), Vec(n * NVAR)); },
                 [&](long long c, std::pair<Vec, Vec> &w) {
        const double *nodal = u.data() + static_cast<size_t>(c) * m;
        bool good = physical(nodal, n * n);
        if (good && basis.points != basis.nodes) {
            tensor.interpolate(nodal, w.first.data());
            good = physical(w.first.data(), P);
        }
        for (int side = 0; good && side < NFACES; ++side) {
            face_values(nodal, side, w.second.data());
            good = physical(w.second.data(), n);
        }
        ok[static_cast<size_t>(c)] = good;
    });
    return std::all_of(ok.begin(), ok.end(), [](char x) { return x != 0; });
}
// discretization.cpp:234-236
double Problem::residual_norm(const Vec &r) const {
    const double s = compensated_sum(r.size(), [&](size_t i) { return r[i] * r[i] * inverse_mass[i]; });
    return std::sqrt(s / (example.lx * example.ly));
}
// discretization.cpp:237-242
State Problem::integral(const Vec &u) const {
    Accumulator acc[NVAR];
    const int P = layout.plane(), m = layout.block_size();
    for (int cell = 0; cell < layout.cells; ++cell)
        for (int v = 0; v < NVAR; ++v) {
            double s = 0;
            for (int k = 0; k < P; ++k) {
                size_t i = static_cast<size_t>(cell) * m + v * P + k;
                s += u[i] * mass[i];
            }
            acc[v].add(s);
        }
    State result;
    for (int v = 0; v < NVAR; ++v) result[v] = acc[v].result();
    return result;
}

// initial_and_error.cpp:19-51: cell-mean-preserving positivity scaling over
// nodes, quadrature points, error points and face values.
void Problem::limit_cell(double *nodal) const {
    const int n = basis.nodes, P = basis.points * basis.points;
    const int Q = static_cast<int>(basis.error_x.size());
    Basis error_basis(config.degree, Q);
    TensorProduct error_tensor(error_basis, hx, hy);
    State mean{0, 0, 0, 0};
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i)
            for (int v = 0; v < NVAR; ++v) mean[v] += .25 * basis.w[i] * basis.w[j] * nodal[v * n * n + i + n * j];
    require(mean[0] > 0 && example.gas.pressure(mean) > 0, "Nonphysical cell average");
    const int rows = n * n + P + Q * Q + NFACES * n;
    std::vector<State> checks;
    checks.reserve(rows);
    for (int k = 0; k < n * n; ++k)
        checks.push_back({nodal[k], nodal[n * n + k], nodal[2 * n * n + k], nodal[3 * n * n + k]});
    Vec points(static_cast<size_t>(P) * NVAR);
    tensor.interpolate(nodal, points.data());
    for (int k = 0; k < P; ++k) checks.push_back({points[k], points[P + k], points[2 * P + k], points[3 * P + k]});
    Vec values(static_cast<size_t>(Q) * Q * NVAR);
    error_tensor.interpolate(nodal, values.data());
    const int QQ = Q * Q;
    for (int k = 0; k < QQ; ++k) checks.push_back({values[k], values[QQ + k], values[2 * QQ + k], values[3 * QQ + k]});
    Vec face(n * NVAR);
    for (int side = 0; side < NFACES; ++side) {
        face_values(nodal, side, face.data());
        for (int q = 0; q < n; ++q) checks.push_back({face[q], face[n + q], face[2 * n + q], face[3 * n + q]});
    }
    double theta = 1;
    for (const State &c : checks) {
        State difference;
        for (int v = 0; v < NVAR; ++v) difference[v] = c[v] - mean[v];
        auto physical = [&](double factor) {
            State z;
            for (int v = 0; v < NVAR; ++v) z[v] = mean[v] + factor * difference[v];
            return z[0] > 1e-12 && example.gas.pressure(z) > 1e-12;
        };
        if (!physical(theta)) {
            double low = 0, high = theta;
            for (int it = 0; it < 60; ++it) {
                double middle = .5 * (low + high);
                if (physical(middle)) low = middle; else high = middle;
            }
            theta = .99 * low;
        }
    }
    for (int v = 0; v < NVAR; ++v)
        for (int k = 0; k < n * n; ++k) nodal[v * n * n + k] = mean[v] + theta * (nodal[v * n * n + k] - mean[v]);
}

// initial_and_error.cpp:4-54 (L2 projection from max(20, 2n+3)^2 Gauss samples)
void Problem::initial(Vec &u) const {
    const int n = basis.nodes, Q = static_cast<int>(basis.error_x.size()), m = layout.block_size();
    u.assign(layout.size(), 0.0);
    std::vector<State> samples(static_cast<size_t>(Q) * Q);
    Vec scratch(static_cast<size_t>(n) * Q);
    for (int cell = 0; cell < layout.cells; ++cell) {
        for (int j = 0; j < Q; ++j)
            for (int i = 0; i < Q; ++i)
                samples[i + Q * j] = example.exact(
                    -.5 * example.lx + hx * (cell % layout.nx + .5) + .5 * hx * basis.error_x[i],
                    -.5 * example.ly + hy * (cell / layout.nx + .5) + .5 * hy * basis.error_x[j], 0);
        double *nodal = u.data() + static_cast<size_t>(cell) * m;
        for (int v = 0; v < NVAR; ++v) {
            for (int j = 0; j < Q; ++j)
                for (int i = 0; i < n; ++i) {
                    double s = 0;
                    for (int k = 0; k < Q; ++k) s += samples[k + Q * j][v] * basis.projection(i, k);
                    scratch[i + n * j] = s;
                }
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    double s = 0;
                    for (int k = 0; k < Q; ++k) s += basis.projection(j, k) * scratch[i + n * k];
                    nodal[v * n * n + i + n * j] = s;
                }
        }
        limit_cell(nodal);
    }
    require(admissible(u), "Invalid projected initial state");
}

// Seeded state (no reference counterpart): rho ~ U[0.5,1.5], u,v ~ U[-0.5,0.5],
// p ~ U[0.5,1.5] i.i.d. per node, then the reference's own positivity scaling.
void Problem::seeded(Vec &u, uint64_t seed) const {
    const int n2 = layout.plane(), m = layout.block_size();
    const double g = example.gas.gamma;
    u.assign(layout.size(), 0.0);
    for (int cell = 0; cell < layout.cells; ++cell) {
        double *nodal = u.data() + static_cast<size_t>(cell) * m;
        for (int k = 0; k < n2; ++k) {
            uint64_t base = 4ull * (static_cast<uint64_t>(cell) * n2 + k);
            double rho = .5 + uniform01(seed, base), vx = -.5 + uniform01(seed, base + 1);
            double vy = -.5 + uniform01(seed, base + 2), p = .5 + uniform01(seed, base + 3);
            nodal[k] = rho;
            nodal[n2 + k] = rho * vx;
            nodal[2 * n2 + k] = rho * vy;
            nodal[3 * n2 + k] = p / (g - 1) + .5 * rho * (vx * vx + vy * vy);
        }
        limit_cell(nodal);
    }
    require(admissible(u), "Invalid projected initial state");
}

// initial_and_error.cpp:55-73
State Problem::error(const Vec &u, double time) const {
    const int Q = static_cast<int>(basis.error_x.size()), m = layout.block_size();
    Basis error_basis(config.degree, Q);
    TensorProduct error_tensor(error_basis, hx, hy);
    Vec values(static_cast<size_t>(Q) * Q * NVAR);
    State result{0, 0, 0, 0};
    const int QQ = Q * Q;
    for (int cell = 0; cell < layout.cells; ++cell) {
        error_tensor.interpolate(u.data() + static_cast<size_t>(cell) * m, values.data());
        for (int j = 0; j < Q; ++j)
            for (int i = 0; i < Q; ++i) {
                State e = example.exact(-.5 * example.lx + hx * (cell % layout.nx + .5) + .5 * hx * basis.error_x[i],
                                        -.5 * example.ly + hy * (cell / layout.nx + .5) + .5 * hy * basis.error_x[j],
                                        time);
                for (int v = 0; v < NVAR; ++v) {
                    double d = values[v * QQ + i + Q * j] - e[v];
                    result[v] += .25 * hx * hy * basis.error_w[i] * basis.error_w[j] * (d * d);
                }
            }
    }
    for (auto &r : result) r = std::sqrt(r);
    return result;
}

// ------------------------------------------------------ volume linearisation
VolumeJacobian::VolumeJacobian(const Problem &p)
    : primitives(p.basis.points * p.basis.points, 3 * p.layout.cells), p_(p) {}
// volume_jacobian.cpp:13-20, 30-42
void VolumeJacobian::refresh(const Vec &state) {
    require(p_.step_size > 0, "Set the time step before refreshing the Jacobian");
    inverse_step = 1 / p_.step_size;
    const int P = p_.basis.points * p_.basis.points, m = p_.layout.block_size();
    if (p_.basis.nodes > 1)
        parallel_for(p_.layout.cells, [&] { return Vec(static_cast<size_t>(P) * NVAR); }, [&](long long c, Vec &points) {
            const int cell = static_cast<int>(c);
            p_.tensor.interpolate(state.data() + static_cast<size_t>(cell) * m, points.data());
            for (int q = 0; q < P; ++q) {
                auto prim = p_.example.gas.primitive({points[q], points[P + q], points[2 * P + q], points[3 * P + q]});
                primitives(q, 3 * cell) = prim.velocity[0];
                primitives(q, 3 * cell + 1) = prim.velocity[1];
                primitives(q, 3 * cell + 2) = prim.enthalpy;
            }
        });
    ready_ = true;
}
// volume_jacobian.cpp:43-58 (dF_x, dF_y from the cached (u, v, H))
void VolumeJacobian::point_action(int cell, const double *d, double *fx, double *fy) const {
    const int P = p_.basis.points * p_.basis.points;
    const double gm = p_.example.gas.gamma - 1.;
    for (int q = 0; q < P; ++q) {
        const double u = primitives(q, 3 * cell), v = primitives(q, 3 * cell + 1), H = primitives(q, 3 * cell + 2);
        const double dr = d[q], dx = d[P + q], dy = d[2 * P + q], de = d[3 * P + q];
        const double dp = gm * (de - u * dx - v * dy + .5 * (u * u + v * v) * dr);
        fx[q] = dx;
        fy[q] = dy;
        fx[P + q] = 2 * u * dx - u * u * dr + dp;
        fx[2 * P + q] = u * dy + v * dx - u * v * dr;
        fy[P + q] = fx[2 * P + q];
        fy[2 * P + q] = 2 * v * dy - v * v * dr + dp;
        fx[3 * P + q] = H * (dx - u * dr) + u * (de + dp);
        fy[3 * P + q] = H * (dy - v * dr) + v * (de + dp);
    }
}
// volume_jacobian.cpp:59-71
void VolumeJacobian::apply(const Vec &direction, Vec &result) const {
    require(ready_, "Refresh the volume Jacobian before applying it");
    result.resize(direction.size());
    parallel_for(static_cast<long long>(direction.size()),
                 [&](long long i) { result[i] = p_.mass[i] * direction[i] * inverse_step; });
    if (p_.basis.nodes == 1) return;
    const int P = p_.basis.points * p_.basis.points, m = p_.layout.block_size();
    const size_t width = static_cast<size_t>(P) * NVAR;
    parallel_for(p_.layout.cells, [&] { return Vec(3 * width); }, [&](long long c, Vec &w) {
        const int cell = static_cast<int>(c);
        double *points = w.data(), *fx = points + width, *fy = fx + width;
        p_.tensor.interpolate(direction.data() + static_cast<size_t>(cell) * m, points);
        point_action(cell, points, fx, fy);
        p_.tensor.subtract_weak(fx, fy, result.data() + static_cast<size_t>(cell) * m);
    });
}
// volume_jacobian.cpp:72-126 (explicit element volume block + M/tau)
void VolumeJacobian::block(int cell, Mat &result) const {
    require(ready_, "Refresh the volume Jacobian before assembling a block");
    const auto &b = p_.basis;
    const int n = b.nodes, Q = b.points, plane = p_.layout.plane(), m = p_.layout.block_size();
    result = Mat(m, m);
    if (n > 1) {
        const int P = Q * Q;
        std::vector<Jac> ax(P), ay(P);
        for (int q = 0; q < P; ++q) {
            ax[q] = p_.example.gas.cached_jacobian(primitives(q, 3 * cell), primitives(q, 3 * cell + 1),
                                                   primitives(q, 3 * cell + 2), 0);
            ay[q] = p_.example.gas.cached_jacobian(primitives(q, 3 * cell), primitives(q, 3 * cell + 1),
                                                   primitives(q, 3 * cell + 2), 1);
        }
        const Mat &d = b.weighted_derivatives;
        if (Q == n) {
            for (int w = 0; w < NVAR; ++w)
                for (int jy = 0; jy < n; ++jy)
                    for (int jx = 0; jx < n; ++jx) {
                        int node = jx + n * jy;
                        double *column = result.col(p_.layout.local(w, node));
                        for (int v = 0; v < NVAR; ++v) {
                            double sx = .5 * p_.hy * b.w[jy] * ax[node][v * 4 + w];
                            for (int i = 0; i < n; ++i) column[v * plane + n * jy + i] -= sx * d(jx, i);
                            double scale = -.5 * p_.hx * b.w[jx] * ay[node][v * 4 + w];
                            for (int iy = 0; iy < n; ++iy) column[v * plane + jx + n * iy] += scale * d(jy, iy);
                        }
                    }
        } else {
            Vec line_x(static_cast<size_t>(n) * n), line_y(static_cast<size_t>(n) * n), wxv(Q), wyv(Q);
            for (int qb = 0; qb < Q; ++qb)
                for (int w = 0; w < NVAR; ++w)
                    for (int v = 0; v < NVAR; ++v) {
                        for (int k = 0; k < n; ++k) {
                            for (int qa = 0; qa < Q; ++qa) {
                                wxv[qa] = ax[qb * Q + qa][v * 4 + w] * b.values(qa, k);
                                wyv[qa] = ay[qb * Q + qa][v * 4 + w] * b.values(qa, k);
                            }
                            for (int i = 0; i < n; ++i) {
                                double sx = 0, sy = 0;
                                for (int qa = 0; qa < Q; ++qa) {
                                    sx += d(qa, i) * wxv[qa];
                                    sy += (b.values(qa, i) * b.qw[qa]) * wyv[qa];
                                }
                                line_x[i + n * k] = -sx;
                                line_y[i + n * k] = -sy;
                            }
                        }
                        for (int l = 0; l < n; ++l)
                            for (int k = 0; k < n; ++k) {
                                double *column = result.col(p_.layout.local(w, k + n * l));
                                for (int j = 0; j < n; ++j) {
                                    double cx = .5 * p_.hy * b.qw[qb] * b.values(qb, j) * b.values(qb, l);
                                    double cy = .5 * p_.hx * d(qb, j) * b.values(qb, l);
                                    for (int i = 0; i < n; ++i)
                                        column[v * plane + n * j + i] += cx * line_x[i + n * k] + cy * line_y[i + n * k];
                                }
                            }
                    }
        }
    }
    for (int k = 0; k < m; ++k) result(k, k) += p_.mass[static_cast<size_t>(cell) * m + k] * inverse_step;
}

// -------------------------------------------------------- face linearisation
FaceJacobian::FaceJacobian(const Problem &p)
    : diagonal(p.trace_size()), primitives_left(p.basis.nodes, 3 * p.face_count()),
      primitives_right(p.basis.nodes, 3 * p.face_count()), tau_left(p.basis.nodes, p.face_count()),
      tau_right(p.basis.nodes, p.face_count()), weights(p.basis.nodes, p.face_count()),
      blend_left(p.basis.nodes, p.face_count()), blend_right(p.basis.nodes, p.face_count()),
      penalty(p.basis.nodes, p.face_count()), p_(p) {
    for (int face = 0; face < p.face_count(); ++face)
        for (int q = 0; q < p.basis.nodes; ++q) weights(q, face) = p.face_weight(face, q);
}
// face_jacobian.cpp:16-37
void FaceJacobian::refresh(const Vec &state) {
    const int n = p_.basis.nodes, m = p_.layout.block_size(), fw = p_.face_width();
    parallel_for(p_.face_count(), [&] { return std::make_pair(Vec(n * NVAR), Vec(n * NVAR)); },
                 [&](long long f, std::pair<Vec, Vec> &lr) {
            Vec &left = lr.first, &right = lr.second;
            const int face = static_cast<int>(f), axis = face / p_.layout.cells, cell = face % p_.layout.cells;
            const int side = 2 * axis + 1;
            p_.face_values(state.data() + static_cast<size_t>(cell) * m, side, left.data());
            p_.face_values(state.data() + static_cast<size_t>(p_.neighbor(cell, side)) * m, side - 1, right.data());
            for (int q = 0; q < n; ++q) {
                auto l = p_.example.gas.primitive({left[q], left[n + q], left[2 * n + q], left[3 * n + q]});
                auto r = p_.example.gas.primitive({right[q], right[n + q], right[2 * n + q], right[3 * n + q]});
                primitives_left(q, 3 * face) = l.velocity[0];
                primitives_left(q, 3 * face + 1) = l.velocity[1];
                primitives_left(q, 3 * face + 2) = l.enthalpy;
                primitives_right(q, 3 * face) = r.velocity[0];
                primitives_right(q, 3 * face + 1) = r.velocity[1];
                primitives_right(q, 3 * face + 2) = r.enthalpy;
                const auto s = p_.speeds[static_cast<size_t>(face) * n + q];
                tau_left(q, face) = s.a;
                tau_right(q, face) = s.b;
                blend_left(q, face) = weights(q, face) * s.b / (s.a + s.b);
                blend_right(q, face) = weights(q, face) * s.a / (s.a + s.b);
                penalty(q, face) = weights(q, face) * s.a * s.b / (s.a + s.b);
                for (int v = 0; v < NVAR; ++v) diagonal[static_cast<size_t>(face) * fw + v * n + q] = -weights(q, face) * (s.a + s.b);
            }
        });
}
// face_jacobian.cpp:38-64 (normal_action on both sides, HLL blend + penalty)
void FaceJacobian::hll_action(int face, const double *l, const double *r, double *flux) const {
    const int n = p_.basis.nodes, axis = face / p_.layout.cells;
    const double gm = p_.example.gas.gamma - 1.;
    auto normal = [&](const Mat &prim, const double *d, int q, double *out) {
        const double u = prim(q, 3 * face), v = prim(q, 3 * face + 1), H = prim(q, 3 * face + 2);
        const double dr = d[q], dx = d[n + q], dy = d[2 * n + q], de = d[3 * n + q];
        const double dp = gm * (de - u * dx - v * dy + .5 * (u * u + v * v) * dr);
        if (axis == 0) {
            out[0] = dx;
            out[1] = 2 * u * dx - u * u * dr + dp;
            out[2] = u * dy + v * dx - u * v * dr;
            out[3] = H * (dx - u * dr) + u * (de + dp);
        } else {
            out[0] = dy;
            out[1] = u * dy + v * dx - u * v * dr;
            out[2] = 2 * v * dy - v * v * dr + dp;
            out[3] = H * (dy - v * dr) + v * (de + dp);
        }
    };
    for (int q = 0; q < n; ++q) {
        double fl[4], fr[4];
        normal(primitives_left, l, q, fl);
        normal(primitives_right, r, q, fr);
        for (int v = 0; v < NVAR; ++v)
            flux[v * n + q] = blend_left(q, face) * fl[v] + blend_right(q, face) * fr[v] +
                              penalty(q, face) * (l[v * n + q] - r[v * n + q]);
    }
}
// face_jacobian.cpp:65-68
Jac FaceJacobian::physical(int face, int q, bool left) const {
    const Mat &a = left ? primitives_left : primitives_right;
    return p_.example.gas.cached_jacobian(a(q, 3 * face), a(q, 3 * face + 1), a(q, 3 * face + 2), face / p_.layout.cells);
}
// face_jacobian.cpp:69-73
Jac FaceJacobian::hll(int face, int q, bool left) const {
    Jac r = physical(face, q, left);
    const double blend = left ? blend_left(q, face) : blend_right(q, face);
    for (auto &x : r) x *= blend;
    for (int k = 0; k < 4; ++k) r[k * 4 + k] += left ? penalty(q, face) : -penalty(q, face);
    return r;
}
// face_jacobian.cpp:77-82
Jac FaceJacobian::hdg(int face, int q, bool left) const {
    Jac r = physical(face, q, left);
    for (auto &x : r) x *= left ? 1. : -1.;
    for (int k = 0; k < 4; ++k) r[k * 4 + k] += tau(face, q, left);
    for (auto &x : r) x *= weights(q, face);
    return r;
}
// face_jacobian.cpp:88-110 (C * directions via face restriction, per side)
void FaceJacobian::cell_c_many(int cell, const Mat &directions, int columns, Mat &result) const {
    const int n = p_.basis.nodes, fw = p_.face_width(), plane = p_.layout.plane();
    const auto fs = p_.faces(cell);
    Vec samples(static_cast<size_t>(fw) * columns);
    for (int side = 0; side < NFACES; ++side) {
        for (int j = 0; j < columns; ++j) p_.face_values(directions.col(j), side, samples.data() + static_cast<size_t>(j) * fw);
        (void)plane;
        for (int q = 0; q < n; ++q) {
            const Jac derivative = hdg(fs[side], q, side % 2 != 0);
            for (int v = 0; v < NVAR; ++v)
                for (int j = 0; j < columns; ++j) {
                    const double *s = samples.data() + static_cast<size_t>(j) * fw;
                    double out = s[q] * derivative[v * 4 + 0];
                    for (int w = 1; w < NVAR; ++w) out += s[w * n + q] * derivative[v * 4 + w];
                    result(side * fw + v * n + q, j) = out;
                }
        }
    }
}

// ------------------------------------------------------------ DG operator
// dg_operator.cpp:10-21
void DGOperator::apply(const Vec &direction, Vec &result) const {
    volume.apply(direction, result);
    const int n = p_.basis.nodes, m = p_.layout.block_size(), fw = p_.face_width();
    Vec fluxes(static_cast<size_t>(p_.face_count()) * fw);
    parallel_for(p_.face_count(), [&] { return std::make_pair(Vec(n * NVAR), Vec(n * NVAR)); },
                 [&](long long f, std::pair<Vec, Vec> &lr) {
        const int face = static_cast<int>(f), axis = face / p_.layout.cells, cell = face % p_.layout.cells;
        const int side = 2 * axis + 1, next = p_.neighbor(cell, side);
        p_.face_values(direction.data() + static_cast<size_t>(cell) * m, side, lr.first.data());
        p_.face_values(direction.data() + static_cast<size_t>(next) * m, side - 1, lr.second.data());
        faces.hll_action(face, lr.first.data(), lr.second.data(), fluxes.data() + static_cast<size_t>(face) * fw);
    });
    p_.lift_faces(fluxes, result);
}
// dg_operator.cpp:22-48
void DGOperator::block(int row, int col, Mat &result) const {
    const int n = p_.basis.nodes, m = p_.layout.block_size();
    if (row == col) volume.block(row, result);
    else result = Mat(m, m);
    auto fs = p_.faces(row);
    for (int side = 0; side < NFACES; ++side) {
        if (col != row && col != p_.neighbor(row, side)) continue;
        bool left = col == (fs[side] % p_.layout.cells);
        int trial_side = 2 * (side / 2) + (left ? 1 : 0);
        const Vec &test = side % 2 ? p_.basis.right : p_.basis.left;
        const Vec &trial = left ? p_.basis.right : p_.basis.left;
        double sign = side % 2 ? 1 : -1;
        for (int q = 0; q < n; ++q) {
            Jac derivative = faces.hll(fs[side], q, left);
            for (int w = 0; w < NVAR; ++w)
                for (int j = 0; j < n; ++j) {
                    double *column = result.col(p_.layout.local(w, p_.face_node(trial_side, q, j)));
                    for (int v = 0; v < NVAR; ++v) {
                        double scale = sign * trial[j] * derivative[v * 4 + w];
                        for (int i = 0; i < n; ++i) column[p_.layout.local(v, p_.face_node(side, q, i))] += scale * test[i];
                    }
                }
        }
    }
}
// J^T w = sum over the 5-point block rows of B(row, col)^T w_row (adjoint oracle).
void DGOperator::apply_transpose_blocks(const Vec &w, Vec &z) const {
    const int m = p_.layout.block_size();
    z.assign(w.size(), 0.0);
    Mat block;
    for (int row = 0; row < p_.layout.cells; ++row) {
        std::array<int, 5> cols = {row, p_.neighbor(row, 0), p_.neighbor(row, 1), p_.neighbor(row, 2), p_.neighbor(row, 3)};
        for (int c : cols) {
            this->block(row, c, block);
            for (int j = 0; j < m; ++j) {
                double s = 0;
                for (int i = 0; i < m; ++i) s += block(i, j) * w[static_cast<size_t>(row) * m + i];
                z[static_cast<size_t>(c) * m + j] += s;
            }
        }
    }
}

// ------------------------------------------------------------------ dense LU
bool DenseLU::factor(const Mat &a) {
    n = a.rows;
    lu = a;
    rowsrc.resize(n);
    for (int i = 0; i < n; ++i) rowsrc[i] = i;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = std::abs(lu(k, k));
        for (int i = k + 1; i < n; ++i)
            if (std::abs(lu(i, k)) > best) { best = std::abs(lu(i, k)); p = i; }
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(lu(k, j), lu(p, j));
            std::swap(rowsrc[k], rowsrc[p]);
        }
        const double pivot = lu(k, k);
        if (pivot != 0)
            for (int i = k + 1; i < n; ++i) lu(i, k) /= pivot;
        for (int j = k + 1; j < n; ++j) {
            const double ukj = lu(k, j);
            double *cj = lu.col(j);
            const double *ck = lu.col(k);
            for (int i = k + 1; i < n; ++i) cj[i] -= ck[i] * ukj;
        }
    }
    for (double x : lu.a)
        if (!std::isfinite(x)) return false;
    for (int k = 0; k < n; ++k)
        if (!(std::abs(lu(k, k)) > std::numeric_limits<double>::min())) return false;
    return true;
}
// dense_lu.hpp:10-16
void DenseLU::compute(const Mat &a) {
    for (double x : a.a) require(std::isfinite(x), "Nonfinite block LU input");
    require(factor(a), "Singular or nonfinite LU factors");
}
// block_jacobi.hpp:27-35: permute, unit-lower solve, upper solve
void DenseLU::solve_inplace(double *x) const {
    Vec y(n);
    for (int i = 0; i < n; ++i) y[i] = x[rowsrc[i]];
    for (int j = 0; j < n; ++j) {
        const double yj = y[j];
        const double *c = lu.col(j);
        for (int i = j + 1; i < n; ++i) y[i] -= c[i] * yj;
    }
    for (int j = n - 1; j >= 0; --j) {
        const double *c = lu.col(j);
        y[j] /= c[j];
        const double yj = y[j];
        for (int i = 0; i < j; ++i) y[i] -= c[i] * yj;
    }
    for (int i = 0; i < n; ++i) x[i] = y[i];
}
void DenseLU::solve_multiple(const Mat &b, int cols, Mat &x) const {
    x = Mat(b.rows, b.cols);
    for (int j = 0; j < cols; ++j) {
        std::memcpy(x.col(j), b.col(j), sizeof(double) * b.rows);
        solve_inplace(x.col(j));
    }
}
// block_jacobi.hpp:24-39
void BlockJacobi::apply(const Vec &rhs, Vec &result) const {
    result.resize(rhs.size());
    parallel_for(blocks_, [&](long long i) {
        std::memcpy(result.data() + static_cast<size_t>(i) * width_, rhs.data() + static_cast<size_t>(i) * width_,
                    sizeof(double) * width_);
        lus_[i].solve_inplace(result.data() + static_cast<size_t>(i) * width_);
    });
}

// ------------------------------------------------------------- block ILU(0)
// dense_lu.hpp:38-45: for each row x of X solve y U = x (U^T lower), z L = y
// (L^T unit upper), then X(row, rowsrc[i]) = z[i] (the P^T of PA = LU).
void right_solve(const DenseLU &lu, Mat &x) {
    const int n = lu.n;
    Vec y(n);
    for (int r = 0; r < x.rows; ++r) {
        for (int k = 0; k < n; ++k) {
            double s = x(r, k);
            const double *c = lu.lu.col(k);
            for (int i = 0; i < k; ++i) s -= c[i] * y[i];
            y[k] = s / c[k];
        }
        for (int k = n - 1; k >= 0; --k) {
            double s = y[k];
            const double *c = lu.lu.col(k);
            for (int i = k + 1; i < n; ++i) s -= c[i] * y[i];
            y[k] = s;
        }
        for (int i = 0; i < n; ++i) x(r, lu.rowsrc[i]) = y[i];
    }
}
std::vector<int> red_black_order(int nx, int ny) {
    std::vector<int> order;
    for (int colour = 0; colour < 2; ++colour)
        for (int cell = 0; cell < nx * ny; ++cell)
            if ((cell % nx + cell / nx) % 2 == colour) order.push_back(cell);
    return order;
}
// block_ilu.cpp:3-14 (permuted_graph) and :15-47 (symbolic update schedule)
BlockILU0::BlockILU0(const std::vector<std::vector<int>> &adjacency, int width, std::vector<int> order)
    : rows_(static_cast<int>(adjacency.size())), width_(width), order_(std::move(order)) {
    require(static_cast<int>(order_.size()) == rows_, "Block ordering size mismatch");
    std::vector<int> inverse(rows_, -1);
    for (int i = 0; i < rows_; ++i) inverse[order_[i]] = i;
    for (int i = 0; i < rows_; ++i) require(inverse[i] >= 0, "Block ordering is not a permutation");
    row_ptr_.assign(rows_ + 1, 0);
    for (int i = 0; i < rows_; ++i) {
        std::vector<int> row{i};
        for (int c : adjacency[order_[i]]) row.push_back(inverse[c]);
        std::sort(row.begin(), row.end());
        row.erase(std::unique(row.begin(), row.end()), row.end());
        col_idx_.insert(col_idx_.end(), row.begin(), row.end());
        row_ptr_[i + 1] = static_cast<int>(col_idx_.size());
    }
    blocks_.resize(col_idx_.size());
    diagonal_.resize(rows_);
    work_.resize(static_cast<size_t>(rows_) * width_);
    diagonal_slots_.resize(rows_);
    update_offsets_.resize(col_idx_.size() + 1);
    for (int row = 0; row < rows_; ++row) {
        diagonal_slots_[row] = find(row, row);
        for (int slot = row_ptr_[row]; slot < row_ptr_[row + 1]; ++slot) {
            update_offsets_[slot] = static_cast<int>(update_targets_.size());
            const int eliminated = col_idx_[slot];
            if (eliminated >= row) continue;
            for (int upper = row_ptr_[eliminated]; upper < row_ptr_[eliminated + 1]; ++upper) {
                const int column = col_idx_[upper];
                if (column <= eliminated) continue;
                const int target = find(row, column);
                if (target >= 0) {
                    update_targets_.push_back(target);
                    update_upper_.push_back(upper);
                }
            }
        }
    }
    update_offsets_.back() = static_cast<int>(update_targets_.size());
}
int BlockILU0::find(int row, int column) const {  // block_graph.cpp:37-42
    auto begin = col_idx_.begin() + row_ptr_[row], end = col_idx_.begin() + row_ptr_[row + 1];
    auto found = std::lower_bound(begin, end, column);
    return found != end && *found == column ? static_cast<int>(found - col_idx_.begin()) : -1;
}
// block_ilu.cpp:48-61; subtract_product is dense_lu.hpp:50-55
void BlockILU0::factorize() {
    for (int i = 0; i < rows_; ++i) {
        for (int p = row_ptr_[i]; p < diagonal_slots_[i]; ++p) {
            right_solve(diagonal_[col_idx_[p]], blocks_[p]);
            const Mat &left = blocks_[p];
            for (int u = update_offsets_[p]; u < update_offsets_[p + 1]; ++u) {
                Mat &dest = blocks_[update_targets_[u]];
                const Mat &right = blocks_[update_upper_[u]];
                for (int j = 0; j < right.cols; ++j)
                    for (int k = 0; k < left.cols; ++k) {
                        const double rkj = right(k, j);
                        double *d = dest.col(j);
                        const double *l = left.col(k);
                        for (int r = 0; r < dest.rows; ++r) d[r] -= rkj * l[r];
                    }
            }
        }
        diagonal_[i].compute(blocks_[diagonal_slots_[i]]);
    }
}
// block_ilu.cpp:62-84
void BlockILU0::apply(const Vec &rhs, Vec &result) const {
    const size_t w = width_;
    for (int i = 0; i < rows_; ++i)
        std::memcpy(work_.data() + i * w, rhs.data() + order_[i] * w, sizeof(double) * w);
    auto subtract = [&](int slot, int row) {
        const Mat &b = blocks_[slot];
        const double *x = work_.data() + col_idx_[slot] * w;
        double *y = work_.data() + row * w;
        for (size_t j = 0; j < w; ++j) {
            const double xj = x[j];
            const double *c = b.col(static_cast<int>(j));
            for (size_t r = 0; r < w; ++r) y[r] -= c[r] * xj;
        }
    };
    for (int i = 0; i < rows_; ++i)
        for (int p = row_ptr_[i]; p < diagonal_slots_[i]; ++p) subtract(p, i);
    for (int i = rows_ - 1; i >= 0; --i) {
        for (int p = diagonal_slots_[i] + 1; p < row_ptr_[i + 1]; ++p) subtract(p, i);
        diagonal_[i].solve_inplace(work_.data() + i * w);
    }
    result.resize(rhs.size());
    for (int i = 0; i < rows_; ++i)
        std::memcpy(result.data() + order_[i] * w, work_.data() + i * w, sizeof(double) * w);
}

// ------------------------------------------------------------------- GMRES
// gmres.cpp:5-12
GMRES::GMRES(int n, int restart, const Vec &weights)
    : n_(n), restart_(std::min(n, restart)), basis_(n, std::min(n, restart) + 1),
      hessenberg_(std::min(n, restart) + 1, std::min(n, restart)), cosine_(restart_), sine_(restart_),
      rotated_rhs_(restart_ + 1), coefficients_(restart_), work_(n), residual_(n), column_(n), weights_(weights),
      projection_(restart_) {
    bool ok = n > 0 && restart > 0 && static_cast<int>(weights_.size()) == n;
    for (double w : weights_) ok = ok && std::isfinite(w) && w > 0;
    require(ok, "Invalid GMRES metric or dimensions");
}
double GMRES::dot(const Vec &a, const Vec &b) const {
    return compensated_sum(static_cast<size_t>(n_), [&](size_t i) { return a[i] * weights_[i] * b[i]; });
}
// gmres.cpp:17-86: restarted right-preconditioned GMRES, weighted CGS2, Givens
LinearResult GMRES::solve(const Action &A, const Action &M, const Vec &rhs, Vec &x, int limit, double tolerance) {
    x.assign(n_, 0.0);
    residual_ = rhs;
    const double rhs_norm = norm(rhs);
    if (rhs_norm == 0) return {0, 0};
    int iterations = 0;
    Vec weighted(n_);
    while (iterations < limit) {
        const double beta = norm(residual_);
        if (beta <= tolerance * rhs_norm) return {iterations, beta / rhs_norm};
        parallel_for(n_, [&](long long i) { basis_(static_cast<int>(i), 0) = residual_[i] / beta; });
        hessenberg_.zero();
        std::fill(rotated_rhs_.begin(), rotated_rhs_.end(), 0.0);
        rotated_rhs_[0] = beta;
        int count = 0;
        for (int j = 0; j < restart_ && iterations < limit; ++j) {
            std::memcpy(column_.data(), basis_.col(j), sizeof(double) * n_);
            M(column_, work_);
            A(work_, column_);
            for (int pass = 0; pass < 2; ++pass) {
                parallel_for(n_, [&](long long i) { weighted[i] = weights_[i] * column_[i]; });
                for (int k = 0; k <= j; ++k) {
                    const double *v = basis_.col(k);
                    const double s = compensated_sum(static_cast<size_t>(n_), [&](size_t i) { return v[i] * weighted[i]; });
                    projection_[k] = s;
                    hessenberg_(k, j) += s;
                }
                parallel_for(n_, [&](long long i) {
                    double s = 0;
                    for (int k = 0; k <= j; ++k) s += basis_(static_cast<int>(i), k) * projection_[k];
                    column_[i] -= s;
                });
            }
            const double next = norm(column_);
            hessenberg_(j + 1, j) = next;
            if (next > 0) parallel_for(n_, [&](long long i) { basis_(static_cast<int>(i), j + 1) = column_[i] / next; });
            for (int i = 0; i < j; ++i) {
                double a = hessenberg_(i, j), b = hessenberg_(i + 1, j);
                hessenberg_(i, j) = cosine_[i] * a + sine_[i] * b;
                hessenberg_(i + 1, j) = -sine_[i] * a + cosine_[i] * b;
            }
            const double d = std::hypot(hessenberg_(j, j), hessenberg_(j + 1, j));
            require(std::isfinite(d) && d > 0, "GMRES Arnoldi breakdown");
            cosine_[j] = hessenberg_(j, j) / d;
            sine_[j] = hessenberg_(j + 1, j) / d;
            hessenberg_(j, j) = d;
            hessenberg_(j + 1, j) = 0;
            rotated_rhs_[j + 1] = -sine_[j] * rotated_rhs_[j];
            rotated_rhs_[j] *= cosine_[j];
            ++iterations;
            count = j + 1;
            if (std::abs(rotated_rhs_[j + 1]) <= tolerance * rhs_norm || next <= std::numeric_limits<double>::min())
                break;
        }
        for (int i = count - 1; i >= 0; --i) {
            double value = rotated_rhs_[i];
            for (int j = i + 1; j < count; ++j) value -= hessenberg_(i, j) * coefficients_[j];
            coefficients_[i] = value / hessenberg_(i, i);
        }
        parallel_for(n_, [&](long long i) {
            double s = 0;
            for (int k = 0; k < count; ++k) s += basis_(static_cast<int>(i), k) * coefficients_[k];
            work_[i] = s;
        });
        M(work_, column_);
        parallel_for(n_, [&](long long i) { x[i] += column_[i]; });
        A(x, work_);
        parallel_for(n_, [&](long long i) { residual_[i] = rhs[i] - work_[i]; });
        const double relative = norm(residual_) / rhs_norm;
        require(std::isfinite(relative), "Nonfinite GMRES residual");
        if (relative <= tolerance) return {iterations, relative};
    }
    throw std::runtime_error("GMRES iteration limit reached");
}

// ------------------------------------------------------------- HDG operator
// hybrid.cpp:3-12
HybridOperator::HybridOperator(const Problem &p) : volume(p), faces_jacobian(p), p_(p) {
    const size_t m = p.layout.block_size(), t = p.local_trace_size(), C = p.layout.cells;
    x_.assign(m * t * C, 0.0);
    schur_.assign(t * t * C, 0.0);
    y_.assign(m * C, 0.0);
    rhs_.assign(p.trace_size(), 0.0);
    skeleton_residual_.assign(p.trace_size(), 0.0);
}
// hybrid.cpp:28-30
void HybridOperator::refresh_projected(const Vec &state, const Vec &r) {
    volume_residual_ = r;
    std::fill(skeleton_residual_.begin(), skeleton_residual_.end(), 0.0);
    condense(state);
}
// hybrid.cpp:31-55
void HybridOperator::assemble_cell(int cell, Mat &a, Mat &local_rhs) const {
    const int n = p_.basis.nodes, fw = p_.face_width(), t = p_.local_trace_size(), m = p_.layout.block_size();
    volume.block(cell, a);
    local_rhs = Mat(m, t + 1);
    const auto fs = p_.faces(cell);
    for (int side = 0; side < NFACES; ++side) {
        const Vec &shape = side % 2 ? p_.basis.right : p_.basis.left;
        for (int q = 0; q < n; ++q) {
            const Jac derivative = faces_jacobian.hdg(fs[side], q, side % 2 != 0);
            const double stabilization = p_.face_weight(fs[side], q) * faces_jacobian.tau(fs[side], q, side % 2 != 0);
            for (int v = 0; v < NVAR; ++v)
                for (int i = 0; i < n; ++i)
                    local_rhs(p_.layout.local(v, p_.face_node(side, q, i)), side * fw + v * n + q) = -stabilization * shape[i];
            for (int w = 0; w < NVAR; ++w)
                for (int j = 0; j < n; ++j) {
                    double *column = a.col(p_.layout.local(w, p_.face_node(side, q, j)));
                    for (int v = 0; v < NVAR; ++v) {
                        const double scale = shape[j] * derivative[v * 4 + w];
                        for (int i = 0; i < n; ++i) column[p_.layout.local(v, p_.face_node(side, q, i))] += scale * shape[i];
                    }
                }
        }
    }
    for (int k = 0; k < m; ++k) local_rhs(k, t) = volume_residual_[static_cast<size_t>(cell) * m + k];
}
// hybrid.cpp:56-72 (element-streamed exact static condensation)
void HybridOperator::condense(const Vec &state) {
    volume.refresh(state);
    faces_jacobian.refresh(state);
    for (size_t i = 0; i < rhs_.size(); ++i) rhs_[i] = -skeleton_residual_[i];
    const int m = p_.layout.block_size(), t = p_.local_trace_size(), fw = p_.face_width();
    Mat a, local_rhs, solved, face_action(t, t + 1);
    DenseLU lu;
    for (int cell = 0; cell < p_.layout.cells; ++cell) {
        assemble_cell(cell, a, local_rhs);
        lu.compute(a);
        lu.solve_multiple(local_rhs, t + 1, solved);
        std::memcpy(x_.data() + static_cast<size_t>(cell) * m * t, solved.a.data(), sizeof(double) * m * t);
        std::memcpy(y_.data() + static_cast<size_t>(cell) * m, solved.col(t), sizeof(double) * m);
        faces_jacobian.cell_c_many(cell, solved, t + 1, face_action);
        double *s = schur_.data() + static_cast<size_t>(cell) * t * t;
        for (int k = 0; k < t * t; ++k) s[k] = -face_action.a[k];
        const auto fs = p_.faces(cell);
        for (int side = 0; side < NFACES; ++side)
            for (int k = 0; k < fw; ++k) rhs_[static_cast<size_t>(fs[side]) * fw + k] += face_action(side * fw + k, t);
    }
}
// hybrid.cpp:73-82 (gather / local Schur GEMV / scatter)
void HybridOperator::apply(const Vec &direction, Vec &result) const {
    const int t = p_.local_trace_size(), fw = p_.face_width();
    result.resize(direction.size());
    for (size_t i = 0; i < direction.size(); ++i) result[i] = faces_jacobian.diagonal[i] * direction[i];
    Vec local(t);
    for (int cell = 0; cell < p_.layout.cells; ++cell) {
        const auto fs = p_.faces(cell);
        for (int s = 0; s < NFACES; ++s)
            for (int k = 0; k < fw; ++k) local[s * fw + k] = direction[static_cast<size_t>(fs[s]) * fw + k];
        const double *S = schur_.data() + static_cast<size_t>(cell) * t * t;
        for (int s = 0; s < NFACES; ++s)
            for (int k = 0; k < fw; ++k) {
                double acc = 0;
                const int row = s * fw + k;
                for (int c = 0; c < t; ++c) acc += S[static_cast<size_t>(c) * t + row] * local[c];
                result[static_cast<size_t>(fs[s]) * fw + k] += acc;
            }
    }
}
// hybrid.cpp:83-96
void HybridOperator::block(int row, int col, Mat &result) const {
    const int fw = p_.face_width(), t = p_.local_trace_size();
    result = Mat(fw, fw);
    if (row == col)
        for (int k = 0; k < fw; ++k) result(k, k) = faces_jacobian.diagonal[static_cast<size_t>(row) * fw + k];
    const int axis = row / p_.layout.cells, first = row % p_.layout.cells;
    for (int adjacent = 0; adjacent < 2; ++adjacent) {
        const int cell = adjacent == 0 ? first : p_.neighbor(first, 2 * axis + 1);
        const int row_side = 2 * axis + (adjacent == 0 ? 1 : 0);
        const auto fs = p_.faces(cell);
        const double *S = schur_.data() + static_cast<size_t>(cell) * t * t;
        for (int side = 0; side < NFACES; ++side)
            if (fs[side] == col)
                for (int j = 0; j < fw; ++j)
                    for (int i = 0; i < fw; ++i) result(i, j) += S[static_cast<size_t>(side * fw + j) * t + row_side * fw + i];
    }
}
// hybrid.cpp:97-105
void HybridOperator::recover(const Vec &direction, Vec &result) const {
    const int m = p_.layout.block_size(), t = p_.local_trace_size(), fw = p_.face_width();
    result.resize(static_cast<size_t>(p_.layout.size()));
    Vec local(t);
    for (int cell = 0; cell < p_.layout.cells; ++cell) {
        const auto fs = p_.faces(cell);
        for (int s = 0; s < NFACES; ++s)
            for (int k = 0; k < fw; ++k) local[s * fw + k] = direction[static_cast<size_t>(fs[s]) * fw + k];
        const double *X = x_.data() + static_cast<size_t>(cell) * m * t;
        for (int i = 0; i < m; ++i) {
            double acc = 0;
            for (int c = 0; c < t; ++c) acc += X[static_cast<size_t>(c) * m + i] * local[c];
            result[static_cast<size_t>(cell) * m + i] = -acc - y_[static_cast<size_t>(cell) * m + i];
        }
    }
}
// hybrid.cpp:106-121
void HybridOperator::recovered_residual_from_schur(const Vec &error, Vec &result) const {
    const int n = p_.basis.nodes, fw = p_.face_width(), m = p_.layout.block_size();
    result.assign(static_cast<size_t>(p_.layout.size()), 0.0);
    Vec flux(fw);
    for (int cell = 0; cell < p_.layout.cells; ++cell) {
        const auto fs = p_.faces(cell);
        for (int side = 0; side < NFACES; ++side) {
            const int face = fs[side];
            for (int v = 0; v < NVAR; ++v)
                for (int q = 0; q < n; ++q) {
                    const double d = faces_jacobian.diagonal[static_cast<size_t>(face) * fw + v * n + q];
                    const double tau = faces_jacobian.tau(face, q, side % 2 != 0);
                    flux[v * n + q] = p_.face_weight(face, q) * tau * error[static_cast<size_t>(face) * fw + v * n + q] / d;
                }
            p_.lift(side, flux.data(), result.data() + static_cast<size_t>(cell) * m);
        }
    }
}

// ------------------------------------------------------------ direct solve
// direct.hpp:24-35
LinearResult DirectSolver::solve(const Vec &rhs, Vec &x, const Vec &weights) const {
    const int N = blocks_ * width_;
    x = rhs;
    lu_.solve_inplace(x.data());
    Accumulator num_acc, den_acc;
    for (int i = 0; i < N; ++i) {
        double s = 0;
        for (int j = 0; j < N; ++j) s += matrix_(i, j) * x[j];
        const double res = s - rhs[i];
        num_acc.add(res * res * weights[i]);
        den_acc.add(rhs[i] * rhs[i] * weights[i]);
    }
    const double num = num_acc.result(), den = den_acc.result();
    const double relative = den > 0 ? std::sqrt(num / den) : std::sqrt(num);
    require(std::isfinite(relative), "Nonfinite direct residual");
    return {0, relative};
}

// block_graph.cpp:20-35
std::vector<std::vector<int>> cartesian_graph(int nx, int ny) {
    std::vector<std::vector<int>> g(nx * ny);
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            int cell = x + nx * y;
            std::vector<int> row = {cell, (x + nx - 1) % nx + nx * y, (x + 1) % nx + nx * y,
                                    x + nx * ((y + ny - 1) % ny), x + nx * ((y + 1) % ny)};
            std::sort(row.begin(), row.end());
            row.erase(std::unique(row.begin(), row.end()), row.end());
            g[cell] = row;
        }
    return g;
}
// discretization.cpp:290-298
std::vector<std::vector<int>> trace_graph(const Problem &p) {
    std::vector<std::vector<int>> g(p.face_count());
    for (int cell = 0; cell < p.layout.cells; ++cell) {
        auto fs = p.faces(cell);
        for (int row : fs)
            for (int col : fs) g[row].push_back(col);
    }
    for (auto &row : g) {
        std::sort(row.begin(), row.end());
        row.erase(std::unique(row.begin(), row.end()), row.end());
    }
    return g;
}

// ------------------------------------------------------- time integration
TimeIntegrator::TimeIntegrator(Problem &p, Method method) : p_(p), method_(method) {
    if (p.config.direct) {
        const int unknowns = method == Method::DG ? p.layout.size() : p.trace_size();
        require(unknowns <= 5000, "Direct reference limited to 5000 unknowns");
    }
    graph = method == Method::DG ? cartesian_graph(p.layout.nx, p.layout.ny) : trace_graph(p);
}
// time_integrator.cpp:42-47
double TimeIntegrator::dual_norm(const Vec &x) const {
    return std::sqrt(compensated_sum(x.size(), [&](size_t i) { return x[i] * x[i] * p_.inverse_mass[i]; }));
}
double TimeIntegrator::state_norm(const Vec &x) const {
    return std::sqrt(compensated_sum(x.size(), [&](size_t i) { return x[i] * x[i] * p_.mass[i]; }));
}

namespace {
using Clock = std::chrono::steady_clock;
double elapsed(Clock::time_point start) { return std::chrono::duration<double>(Clock::now() - start).count(); }
State integral_difference(const State &a, const State &b) {
    State d;
    for (int v = 0; v < NVAR; ++v) d[v] = std::abs(a[v] - b[v]);
    return d;
}
}  // namespace

// time_integrator.cpp:48-213
RunStatistics TimeIntegrator::advance(Vec &solution) {
    require(static_cast<int>(solution.size()) == p_.layout.size() && p_.admissible(solution), "Invalid initial solution");
    const auto &c = p_.config;
    require(c.steps > 0 || c.final_time < p_.example.shock_time(), "Final time reaches simple-wave shock");
    constexpr double alpha = .24169426078821, beta = alpha / 4., eta = .12915286960590;
    constexpr double a[4][4] = {{alpha, 0, 0, 0}, {-alpha, alpha, 0, 0}, {0, 1 - alpha, alpha, 0},
                                {beta, eta, .5 - beta - eta - alpha, alpha}};
    constexpr double b[4] = {0, 1. / 6, 1. / 6, 2. / 3}, stage_c[4] = {alpha, 0, 1, .5};
    const int stages = c.integrator == "be" ? 1 : 4;
    const size_t N = solution.size();
    RunStatistics stats;
    p_.freeze(solution);
    const State initial_integral = p_.integral(solution);
    DGOperator *dg = nullptr;
    HybridOperator *hybrid = nullptr;
    std::unique_ptr<DGOperator> dg_owner, diag_owner;
    std::unique_ptr<HybridOperator> hybrid_owner;
    if (method_ == Method::DG) { dg_owner.reset(new DGOperator(p_)); dg = dg_owner.get(); }
    else { hybrid_owner.reset(new HybridOperator(p_)); hybrid = hybrid_owner.get(); }
    if (method_ == Method::HDG && c.check_update) diag_owner.reset(new DGOperator(p_));
    const int width = method_ == Method::DG ? p_.layout.block_size() : p_.face_width();
    const int blocks = method_ == Method::DG ? p_.layout.cells : p_.face_count();
    const Vec &linear_weights = method_ == Method::DG ? p_.inverse_mass : p_.inverse_trace_mass;
    const int rhs_size = method_ == Method::DG ? p_.layout.size() : p_.trace_size();
    std::unique_ptr<BlockJacobi> preconditioner;
    std::unique_ptr<BlockILU0> ilu;
    std::unique_ptr<GMRES> gmres;
    std::unique_ptr<DirectSolver> direct;
    if (c.direct) direct.reset(new DirectSolver(blocks, width));
    else {
        if (c.ilu_red_black && method_ == Method::DG)
            ilu.reset(new BlockILU0(graph, width, red_black_order(p_.layout.nx, p_.layout.ny)));
        else preconditioner.reset(new BlockJacobi(blocks, width));
        gmres.reset(new GMRES(rhs_size, c.restart, linear_weights));
    }
    Vec state(N), trial(N), residual(N), trial_residual(N), direction(N), rhs(rhs_size);
    Vec lambda(p_.trace_size()), trace_direction(p_.trace_size()), equation_error(N), independent_error(N);
    Vec step_state(N), stage_base(N);
    std::vector<Vec> slopes(4, Vec(N));
    auto begin = Clock::now();
    while (c.steps > 0 ? stats.steps < c.steps : stats.final_time < c.final_time) {
        const double rate = p_.freeze(solution);
        double step = c.cfl / rate;
        if (c.steps == 0) {
            const double remaining = c.final_time - stats.final_time;
            step = std::min(step, remaining);
            if (remaining - step <= 16 * std::numeric_limits<double>::epsilon() * std::max(remaining, c.final_time))
                step = remaining;
        }
        require(std::isfinite(step) && step > 0 && stats.final_time + step > stats.final_time, "Invalid CFL step");
        require(stats.final_time + step < p_.example.shock_time(), "Step reaches simple-wave shock");
        step_state = solution;
        state = solution;
        for (int stage = 0; stage < stages; ++stage) {
            const double diagonal = stages == 1 ? 1. : a[stage][stage];
            stage_base = step_state;
            for (int j = 0; j < stage; ++j)
                for (size_t i = 0; i < N; ++i) stage_base[i] += step * a[stage][j] * slopes[j][i];
            if (stage > 0) p_.freeze(state);
            p_.step_size = diagonal * step;
            p_.old = stage_base;
            const double tau = p_.step_size, old_norm = state_norm(stage_base);
            const int limit = c.newton_steps > 0 ? c.newton_steps : c.max_newton;
            double initial_rms = -1.;
            bool converged = false, residual_ready = false;
            for (int it = 0; it <= limit; ++it) {
                const auto update_start = Clock::now();
                if (!residual_ready) p_.dg_residual(state, residual);
                const double norm = dual_norm(residual), rms = norm / std::sqrt(p_.example.lx * p_.example.ly);
                if (initial_rms < 0) initial_rms = rms;
                if (c.newton_steps == 0 && rms <= c.nonlinear_tolerance * std::max(1., initial_rms)) {
                    converged = true;
                    break;
                }
                if (it == limit) break;
                const double allowance =
                    128 * std::numeric_limits<double>::epsilon() * ((state_norm(state) + old_norm) / tau + norm);
                const double target = std::max(10 * c.linear_tolerance * norm, allowance);
                LinearResult linear{0, 0};
                if (method_ == Method::DG) {
                    dg->refresh(state);
                    for (size_t i = 0; i < N; ++i) rhs[i] = -residual[i];
                    auto provider = [&](int row, int col, Mat &block) { dg->block(row, col, block); };
                    if (direct) {
                        direct->refresh(graph, provider);
                        linear = direct->solve(rhs, direction, linear_weights);
                        dg->apply(direction, equation_error);
                        for (size_t i = 0; i < N; ++i) equation_error[i] += residual[i];
                    } else if (ilu) {
                        ilu->refresh(provider);
                        linear = gmres->solve([&](const Vec &x, Vec &y) { dg->apply(x, y); },
                                              [&](const Vec &x, Vec &y) { ilu->apply(x, y); }, rhs, direction,
                                              c.max_linear, c.linear_tolerance);
                        for (size_t i = 0; i < N; ++i) equation_error[i] = -gmres->residual()[i];
                    } else {
                        preconditioner->refresh(provider);
                        linear = gmres->solve([&](const Vec &x, Vec &y) { dg->apply(x, y); },
                                              [&](const Vec &x, Vec &y) { preconditioner->apply(x, y); }, rhs,
                                              direction, c.max_linear, c.linear_tolerance);
                        for (size_t i = 0; i < N; ++i) equation_error[i] = -gmres->residual()[i];
                    }
                } else {
                    hybrid->refresh_projected(state, residual);
                    rhs = hybrid->rhs();
                    auto provider = [&](int row, int col, Mat &block) { hybrid->block(row, col, block); };
                    if (direct) direct->refresh(graph, provider);
                    else preconditioner->refresh(provider);
                    double tolerance = c.linear_tolerance;
                    bool valid = false;
                    for (int attempt = 0; attempt < 4; ++attempt) {
                        LinearResult current;
                        if (direct) current = direct->solve(rhs, trace_direction, linear_weights);
                        else
                            current = gmres->solve([&](const Vec &x, Vec &y) { hybrid->apply(x, y); },
                                                   [&](const Vec &x, Vec &y) { preconditioner->apply(x, y); }, rhs,
                                                   trace_direction, c.max_linear, tolerance);
                        linear.iterations += current.iterations;
                        linear.relative_residual = current.relative_residual;
                        hybrid->recover(trace_direction, direction);
                        if (direct) {
                            hybrid->apply(trace_direction, lambda);
                            for (size_t i = 0; i < lambda.size(); ++i) lambda[i] -= rhs[i];
                        } else
                            for (size_t i = 0; i < lambda.size(); ++i) lambda[i] = -gmres->residual()[i];
                        hybrid->recovered_residual_from_schur(lambda, equation_error);
                        valid = dual_norm(equation_error) <= target;
                        if (valid) break;
                        require(!direct, "Direct Schur solve fails DG residual criterion");
                        tolerance *= .01;
                    }
                    require(valid, "Recovered DG residual criterion failed after Schur refinement");
                }
                const double error_norm = dual_norm(equation_error);
                require(error_norm <= target, "DG Newton linear residual too large");
                if (diag_owner) {
                    diag_owner->refresh(state);
                    diag_owner->apply(direction, independent_error);
                    for (size_t i = 0; i < N; ++i) independent_error[i] += residual[i];
                    require(dual_norm(independent_error) <= std::max(4 * target, 1e-12),
                            "Independent recovered DG equation check failed");
                    for (size_t i = 0; i < N; ++i) independent_error[i] -= equation_error[i];
                    require(dual_norm(independent_error) <= std::max(4 * allowance, 1e-12), "Recovered residual identity failed");
                }
                double damping = 1.;
                bool accepted = false;
                for (int search = 0; search < 30; ++search) {
                    parallel_for(static_cast<long long>(N), [&](long long i) { trial[i] = state[i] + damping * direction[i]; });
                    if (p_.admissible(trial)) {
                        p_.dg_residual(trial, trial_residual);
                        if (dual_norm(trial_residual) <= (1 - 1e-4 * damping) * norm + allowance) {
                            accepted = true;
                            break;
                        }
                    }
                    damping *= .5;
                }
                require(accepted, "Newton line search failed");
                state.swap(trial);
                residual.swap(trial_residual);
                residual_ready = true;
                const double duration = elapsed(update_start);
                stats.max_linear_residual = std::max(stats.max_linear_residual, linear.relative_residual);
                const double dgrel = norm > 0 ? error_norm / norm : 0.;
                stats.max_dg_linear_residual = std::max(stats.max_dg_linear_residual, dgrel);
                stats.krylov_iterations += linear.iterations;
                if (stats.newton_iterations >= c.warmup_updates) {
                    stats.measured_newton_seconds += duration;
                    ++stats.measured_updates;
                }
                ++stats.newton_iterations;
                stats.updates.push_back({stats.steps, stage, it, linear.iterations, stats.final_time + step,
                                         stats.final_time + (stages == 1 ? 1. : stage_c[stage]) * step, step,
                                         step * rate, rms, damping, linear.relative_residual, dgrel, duration});
            }
            require(converged || c.newton_steps > 0, "Newton iteration limit reached");
            for (size_t i = 0; i < N; ++i) slopes[stage][i] = (state[i] - stage_base[i]) / tau;
        }
        if (stages == 1) solution = state;
        else {
            solution = step_state;
            for (int stage = 0; stage < stages; ++stage)
                for (size_t i = 0; i < N; ++i) solution[i] += step * b[stage] * slopes[stage][i];
            require(p_.admissible(solution), "DIRK3 step produced an inadmissible state");
        }
        stats.final_time = c.steps == 0 && step >= c.final_time - stats.final_time ? c.final_time : stats.final_time + step;
        stats.minimum_step = stats.steps == 0 ? step : std::min(stats.minimum_step, step);
        stats.maximum_step = std::max(stats.maximum_step, step);
        ++stats.steps;
        const State d = integral_difference(p_.integral(solution), initial_integral);
        stats.conservation_drift = std::max(stats.conservation_drift, *std::max_element(d.begin(), d.end()));
    }
    stats.seconds = elapsed(begin);
    return stats;
}

// ------------------------------------------------------- seeded generator
uint64_t splitmix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double uniform01(uint64_t seed, uint64_t index) {
    return static_cast<double>(splitmix64(seed + (index + 1) * 0x9E3779B97F4A7C15ull) >> 11) * 0x1.0p-53;
}
double normal01(uint64_t seed, uint64_t index) {
    const double u1 = uniform01(seed, 2 * index), u2 = uniform01(seed, 2 * index + 1);
    return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(2.0 * std::acos(-1.0) * u2);
}
}  // namespace oracle
