// CPU restatement of the reference's 1D periodic Euler DG solver
// (dg_hdg_1d/src) for the parity tests of the device 1D path. Test
// infrastructure: loaded by tests/ through ctypes, never by the product.
// Shares the nodal basis, dense LU, block Jacobi and weighted GMRES with the
// 2D restatement (common/include/euler/{nodal_basis.hpp, linalg/*},
// common/src/linalg/gmres.cpp), which the 2D checks pin.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "dgh_oracle.hpp"

namespace oracle1d {
using oracle::Basis;
using oracle::BlockJacobi;
using oracle::GMRES;
using oracle::LinearResult;
using oracle::Mat;
using oracle::require;
using oracle::Vec;

constexpr int NV = 3;  // types.hpp:8 (density, momentum, energy)
using State = std::array<double, NV>;
using Jacobian = std::array<double, NV * NV>;  // column-major, J(v, w) at [v + 3 w]

// physics.cpp:4-24
struct Gas {
    double gamma = 3.0;  // simple_wave.hpp:6
    double pressure(const State &u) const { return (gamma - 1) * (u[2] - .5 * u[1] * u[1] / u[0]); }
    State flux(const State &u) const {
        const double v = u[1] / u[0], p = pressure(u);
        return {u[1], u[1] * v + p, (u[2] + p) * v};
    }
    Jacobian jacobian(const State &u) const {
        const double v = u[1] / u[0], H = (u[2] + pressure(u)) / u[0], g = gamma;
        Jacobian a{};
        a[0] = 0, a[3] = 1, a[6] = 0;
        a[1] = .5 * (g - 3) * v * v, a[4] = (3 - g) * v, a[7] = g - 1;
        a[2] = v * (.5 * (g - 1) * v * v - H), a[5] = H - (g - 1) * v * v, a[8] = g * v;
        return a;
    }
    double spectral_radius(const State &u) const {
        return std::abs(u[1] / u[0]) + std::sqrt(gamma * pressure(u) / u[0]);
    }
};
struct Speed { double a, b; };

// physics.cpp:25-39 (Einfeldt speeds with Roe averages)
Speed einfeldt(const State &l, const State &r, const Gas &gas, double safety) {
    require(l[0] > 0 && r[0] > 0 && gas.pressure(l) > 0 && gas.pressure(r) > 0, "Invalid frozen face state");
    const double sl = std::sqrt(l[0]), sr = std::sqrt(r[0]);
    const double ul = l[1] / l[0], ur = r[1] / r[0];
    const double cl = std::sqrt(gas.gamma * gas.pressure(l) / l[0]);
    const double cr = std::sqrt(gas.gamma * gas.pressure(r) / r[0]);
    const double v = (sl * ul + sr * ur) / (sl + sr);
    const double H = (sl * (l[2] + gas.pressure(l)) / l[0] + sr * (r[2] + gas.pressure(r)) / r[0]) / (sl + sr);
    const double c = std::sqrt((gas.gamma - 1) * (H - .5 * v * v));
    return {safety * std::max({0.0, -ul + cl, -v + c}), safety * std::max({0.0, ur + cr, v + c})};
}

// simple_wave.cpp:3-25: the gamma = 3 simple wave through its characteristic
// foot xi (safeguarded Newton on xi + t (a + b J(xi)) = x), J(xi) the
// Riemann invariant u - 2c/(gamma-1) at t = 0 and u + 2c/(gamma-1) = 2.
// The reference's wave has J = sin(pi xi) and p = rho^gamma (amplitude 1,
// entropy 1). The `density` case (no reference counterpart) has
// J = -0.4 sin(pi xi) and p = rho^gamma / gamma, so that for gamma = 3
// rho = c = 1 + 0.2 sin(pi x) and u = 2 - rho at t = 0: density, velocity and
// pressure all positive; it steepens at t = 1 / (b |amplitude| pi).
struct SimpleWave {
    Gas gas;
    double amplitude = 1.0, entropy = 1.0;
    double shock_time() const { return 4 / ((gas.gamma + 1) * std::acos(-1.0) * std::abs(amplitude)); }  // simple_wave.hpp:8
    State exact(double x, double t) const {
        require(t < shock_time(), "Simple wave reaches shock time");
        const double g = gas.gamma, pi = std::acos(-1.0), a = (3 - g) / 2, b = (g + 1) / 4;
        const double bs = b * std::abs(amplitude);
        double lo = x - t * (a + bs), hi = x - t * (a - bs), xi = x - t * a;
        for (int i = 0; i < 80; ++i) {
            const double f = xi + t * (a + b * (amplitude * std::sin(pi * xi))) - x;
            if (std::abs(f) < 2e-15) break;
            if (f > 0) hi = xi;
            else lo = xi;
            const double z = xi - f / (1 + t * (b * amplitude) * pi * std::cos(pi * xi));
            xi = z > lo && z < hi ? z : .5 * (lo + hi);
        }
        const double jm = amplitude * std::sin(pi * xi), v = .5 * (2 + jm), c = (g - 1) * (2 - jm) / 4;
        const double rho = std::pow(c * c / (g * entropy), 1 / (g - 1)), p = entropy * std::pow(rho, g);
        return {rho, rho * v, p / (g - 1) + .5 * rho * v * v};
    }
};

// config.hpp:7-18 (1D defaults)
struct Config {
    int cells = 32, degree = 3, quadrature = 0;
    int restart = 30, max_linear = 1500, max_newton = 30, steps = 0;
    int newton_steps = 0, warmup_updates = 0;
    double cfl = 2.0, final_time = 0.2, safety = 1.1;
    double nonlinear_tolerance = 1e-12, linear_tolerance = 1e-12;
    std::string integrator = "ssp-dirk3";
    std::string testcase = "simple";  // build addition: `density` (rho = 1 + 0.2 sin(pi x) at t = 0)
};

// config.cpp:23-108 (the subset of flags the DG Newton-Krylov path reads)
Config parse(int argc, const char *const *argv) {
    Config c;
    bool time_seen = false, steps_seen = false;
    for (int i = 0; i < argc; ++i) {
        const std::string o = argv[i];
        auto value = [&]() {
            require(i + 1 < argc, "Missing option value");
            return std::string(argv[++i]);
        };
        auto integer = [&]() { return std::stoi(value()); };
        auto real = [&]() { return std::stod(value()); };
        if (o == "--cells") c.cells = integer();
        else if (o == "--k") c.degree = integer();
        else if (o == "--quad") c.quadrature = integer();
        else if (o == "--cfl") c.cfl = real();
        else if (o == "--T") { c.final_time = real(); time_seen = true; }
        else if (o == "--steps") { c.steps = integer(); steps_seen = true; }
        else if (o == "--integrator") c.integrator = value();
        else if (o == "--testcase") c.testcase = value();
        else if (o == "--ntol") c.nonlinear_tolerance = real();
        else if (o == "--ltol") c.linear_tolerance = real();
        else if (o == "--restart") c.restart = integer();
        else if (o == "--max-linear") c.max_linear = integer();
        else if (o == "--max-newton") c.max_newton = integer();
        else if (o == "--newton-steps") c.newton_steps = integer();
        else if (o == "--warmup-updates") c.warmup_updates = integer();
        else if (o == "--safety") c.safety = real();
        else if (o == "--method" || o == "--precond" || o == "--ordering" || o == "--prefix" || o == "--linear")
            (void)value();
        else if (o == "--timing-only" || o == "--check-update") {
        } else throw std::runtime_error("Unknown option: " + o);
    }
    require(!(time_seen && steps_seen), "Choose either --T or --steps");
    require(c.cells >= 3 && c.degree >= 0 && c.degree <= 8 && c.quadrature >= 0, "Invalid mesh or degree");
    require(c.quadrature == 0 || c.quadrature >= c.degree + 1, "Quadrature must be >= degree+1");
    require(c.integrator == "be" || c.integrator == "ssp-dirk3", "Invalid integrator");
    require(c.testcase == "simple" || c.testcase == "density", "Test case must be simple or density");
    return c;
}

// discretization.cpp:4-169
class Problem {
  public:
    Config c;
    SimpleWave example;
    Basis basis;
    int cells, nodes, m;
    double h, step_size = 0;
    Vec old, mass, inverse_mass;
    std::vector<Speed> speeds;

    explicit Problem(const Config &cfg)
        : c(cfg), basis(cfg.degree, cfg.quadrature), cells(cfg.cells), nodes(cfg.degree + 1),
          m(NV * (cfg.degree + 1)), h(2.0 / cfg.cells), old(size()), mass(size()), inverse_mass(size()),
          speeds(cfg.cells) {
        if (cfg.testcase == "density") {
            example.amplitude = -0.4;
            example.entropy = 1 / example.gas.gamma;
        }
        for (int e = 0; e < cells; ++e)
            for (int v = 0; v < NV; ++v)
                for (int i = 0; i < nodes; ++i) mass[index(e, v, i)] = .5 * h * basis.w[i];
        for (size_t k = 0; k < mass.size(); ++k) inverse_mass[k] = 1.0 / mass[k];
    }
    int size() const { return cells * m; }
    int index(int cell, int v, int i) const { return cell * m + v * nodes + i; }  // types.hpp:38-41
    State sample(const Vec &u, int cell, const Vec &shape) const {                 // :15-21
        State r{0, 0, 0};
        for (int v = 0; v < NV; ++v)
            for (int i = 0; i < nodes; ++i) r[v] += shape[i] * u[index(cell, v, i)];
        return r;
    }
    State point(const Vec &u, int cell, int q) const {  // :22-30
        State r{0, 0, 0};
        if (basis.points == nodes) {
            for (int v = 0; v < NV; ++v) r[v] = u[index(cell, v, q)];
            return r;
        }
        for (int v = 0; v < NV; ++v)
            for (int i = 0; i < nodes; ++i) r[v] += basis.values(q, i) * u[index(cell, v, i)];
        return r;
    }
    void initial(Vec &u) const {  // :31-73
        u.assign(size(), 0.0);
        const int eq = static_cast<int>(basis.error_x.size());
        const Gas &gas = example.gas;
        for (int e = 0; e < cells; ++e) {
            Mat samples(eq, NV), nodal(nodes, NV);
            for (int q = 0; q < eq; ++q) {
                const State s = example.exact(-1 + h * (e + .5) + .5 * h * basis.error_x[q], 0);
                for (int v = 0; v < NV; ++v) samples(q, v) = s[v];
            }
            for (int v = 0; v < NV; ++v)
                for (int i = 0; i < nodes; ++i) {
                    double s = 0;
                    for (int q = 0; q < eq; ++q) s += basis.projection(i, q) * samples(q, v);
                    nodal(i, v) = s;
                }
            State mean{0, 0, 0};
            for (int v = 0; v < NV; ++v) {
                double s = 0;
                for (int i = 0; i < nodes; ++i) s += nodal(i, v) * basis.w[i];
                mean[v] = .5 * s;
            }
            require(mean[0] > 0 && gas.pressure(mean) > 0, "Nonphysical cell average");
            std::vector<State> checks;
            for (int i = 0; i < nodes; ++i) checks.push_back({nodal(i, 0), nodal(i, 1), nodal(i, 2)});
            auto mapped = [&](const Mat &map, int rows) {
                for (int q = 0; q < rows; ++q) {
                    State z{0, 0, 0};
                    for (int v = 0; v < NV; ++v)
                        for (int i = 0; i < nodes; ++i) z[v] += map(q, i) * nodal(i, v);
                    checks.push_back(z);
                }
            };
            mapped(basis.values, basis.points);
            mapped(basis.error_values, eq);
            State l{0, 0, 0}, r{0, 0, 0};
            for (int v = 0; v < NV; ++v)
                for (int i = 0; i < nodes; ++i) {
                    l[v] += basis.left[i] * nodal(i, v);
                    r[v] += basis.right[i] * nodal(i, v);
                }
            checks.push_back(l);
            checks.push_back(r);
            double theta = 1;
            for (const State &point : checks) {
                State d;
                for (int v = 0; v < NV; ++v) d[v] = point[v] - mean[v];
                auto physical = [&](double f) {
                    State z;
                    for (int v = 0; v < NV; ++v) z[v] = mean[v] + f * d[v];
                    return z[0] > 1e-12 && gas.pressure(z) > 1e-12;
                };
                if (!physical(theta)) {
                    double low = 0, high = theta;
                    for (int k = 0; k < 60; ++k) {
                        const double mid = .5 * (low + high);
                        if (physical(mid)) low = mid;
                        else high = mid;
                    }
                    theta = .99 * low;
                }
            }
            for (int v = 0; v < NV; ++v)
                for (int i = 0; i < nodes; ++i) u[index(e, v, i)] = mean[v] + theta * (nodal(i, v) - mean[v]);
        }
        require(admissible(u), "Invalid projected initial state");
    }
    double freeze(const Vec &u) {  // :75-92
        old = u;
        double rate = 0;
        for (int e = 0; e < cells; ++e)
            speeds[e] = einfeldt(sample(u, e, basis.right), sample(u, (e + 1) % cells, basis.left), example.gas,
                                 c.safety);
        for (int e = 0; e < cells; ++e) {
            const Speed &l = speeds[(e + cells - 1) % cells], &r = speeds[e];
            double alpha = std::max({l.a, l.b, r.a, r.b});
            for (int q = 0; q < basis.points; ++q) alpha = std::max(alpha, example.gas.spectral_radius(point(u, e, q)));
            rate = std::max(rate, alpha / h);
        }
        require(std::isfinite(rate) && rate > 0, "Invalid CFL rate");
        return rate;
    }
    void volume_residual(const Vec &u, Vec &r) const {  // :102-112
        r.resize(size());
        for (int k = 0; k < size(); ++k) r[k] = mass[k] * (u[k] - old[k]) / step_size;
        if (nodes == 1) return;
        Mat flux(basis.points, NV);
        for (int e = 0; e < cells; ++e) {
            for (int q = 0; q < basis.points; ++q) {
                const State f = example.gas.flux(point(u, e, q));
                for (int v = 0; v < NV; ++v) flux(q, v) = f[v];
            }
            for (int v = 0; v < NV; ++v)
                for (int i = 0; i < nodes; ++i) {
                    double s = 0;
                    for (int q = 0; q < basis.points; ++q) s += basis.weighted_derivatives(q, i) * flux(q, v);
                    r[index(e, v, i)] -= s;
                }
        }
    }
    void dg_residual(const Vec &u, Vec &r) const {  // :113-127
        volume_residual(u, r);
        for (int e = 0; e < cells; ++e) {
            const int next = (e + 1) % cells;
            const Speed s = speeds[e];
            const State l = sample(u, e, basis.right), rr = sample(u, next, basis.left);
            const State fl = example.gas.flux(l), fr = example.gas.flux(rr);
            State f;
            for (int v = 0; v < NV; ++v) f[v] = (s.b * fl[v] + s.a * fr[v] + s.a * s.b * (l[v] - rr[v])) / (s.a + s.b);
            for (int v = 0; v < NV; ++v)
                for (int i = 0; i < nodes; ++i) {
                    r[index(e, v, i)] += basis.right[i] * f[v];
                    r[index(next, v, i)] -= basis.left[i] * f[v];
                }
        }
    }
    bool admissible(const Vec &u) const {  // :128-144
        for (double z : u)
            if (!std::isfinite(z)) return false;
        auto ok = [&](const State &z) { return z[0] > 1e-13 && example.gas.pressure(z) > 1e-13; };
        for (int e = 0; e < cells; ++e) {
            for (int i = 0; i < nodes; ++i)
                if (!ok({u[index(e, 0, i)], u[index(e, 1, i)], u[index(e, 2, i)]})) return false;
            for (int q = 0; q < basis.points; ++q)
                if (!ok(point(u, e, q))) return false;
            if (!ok(sample(u, e, basis.left)) || !ok(sample(u, e, basis.right))) return false;
        }
        return true;
    }
    State integral(const Vec &u) const {  // :148-155
        State t{0, 0, 0};
        for (int e = 0; e < cells; ++e)
            for (int v = 0; v < NV; ++v)
                for (int i = 0; i < nodes; ++i) t[v] += .5 * h * basis.w[i] * u[index(e, v, i)];
        return t;
    }
    State error(const Vec &u, double t) const {  // :156-168
        State sq{0, 0, 0};
        for (int e = 0; e < cells; ++e)
            for (size_t q = 0; q < basis.error_x.size(); ++q) {
                State z{0, 0, 0};
                for (int v = 0; v < NV; ++v)
                    for (int i = 0; i < nodes; ++i)
                        z[v] += basis.error_values(static_cast<int>(q), i) * u[index(e, v, i)];
                const State x = example.exact(-1 + h * (e + .5) + .5 * h * basis.error_x[q], t);
                for (int v = 0; v < NV; ++v) sq[v] += .5 * h * basis.error_w[q] * (z[v] - x[v]) * (z[v] - x[v]);
            }
        for (double &s : sq) s = std::sqrt(s);
        return sq;
    }
    double dual_norm(const Vec &x) const {  // time_integrator.cpp:38-40
        double s = 0;
        for (size_t k = 0; k < x.size(); ++k) s += x[k] * x[k] * inverse_mass[k];
        return std::sqrt(s);
    }
    double state_norm(const Vec &x) const {  // time_integrator.cpp:41-43
        double s = 0;
        for (size_t k = 0; k < x.size(); ++k) s += x[k] * x[k] * mass[k];
        return std::sqrt(s);
    }
};

// volume_jacobian.cpp:3-132 + dg_operator.cpp:3-79: the DG Newton Jacobian
// from F'(U) at the volume points and the two HLL face derivatives per face.
class DGOperator {
  public:
    explicit DGOperator(const Problem &p)
        : p_(p), coef_(static_cast<size_t>(p.cells) * p.basis.points * NV * NV), left_(p.cells), right_(p.cells) {}
    void refresh(const Vec &u) {
        require(p_.step_size > 0, "Set the time step before refreshing the volume Jacobian");
        inverse_step_ = 1 / p_.step_size;
        if (p_.nodes > 1)
            for (int e = 0; e < p_.cells; ++e)
                for (int q = 0; q < p_.basis.points; ++q) {
                    const Jacobian J = p_.example.gas.jacobian(p_.point(u, e, q));
                    for (int k = 0; k < NV * NV; ++k) coef(e, q)[k] = J[k];
                }
        for (int f = 0; f < p_.cells; ++f) {
            const int next = (f + 1) % p_.cells;
            const Speed s = p_.speeds[f];
            const Jacobian L = p_.example.gas.jacobian(p_.sample(u, f, p_.basis.right));
            const Jacobian R = p_.example.gas.jacobian(p_.sample(u, next, p_.basis.left));
            for (int k = 0; k < NV * NV; ++k) {
                const double id = (k % 4 == 0) ? 1.0 : 0.0;
                left_[f][k] = s.b / (s.a + s.b) * (L[k] + s.a * id);
                right_[f][k] = s.a / (s.a + s.b) * (R[k] - s.b * id);
            }
        }
        ready_ = true;
    }
    void apply(const Vec &x, Vec &y) const {
        require(ready_, "Refresh the DG Jacobian before applying it");
        const int n = p_.nodes, Q = p_.basis.points;
        y.resize(x.size());
        for (size_t k = 0; k < x.size(); ++k) y[k] = p_.mass[k] * x[k] * inverse_step_;
        if (n > 1) {
            Mat dir(Q, NV), fd(Q, NV);
            for (int e = 0; e < p_.cells; ++e) {
                for (int q = 0; q < Q; ++q) {
                    const State z = p_.point(x, e, q);
                    for (int v = 0; v < NV; ++v) dir(q, v) = z[v];
                }
                for (int v = 0; v < NV; ++v) {
                    for (int q = 0; q < Q; ++q) {
                        double s = 0;
                        for (int w = 0; w < NV; ++w) s += coef(e, q)[v + NV * w] * dir(q, w);
                        fd(q, v) = s;
                    }
                    for (int i = 0; i < n; ++i) {
                        double s = 0;
                        for (int q = 0; q < Q; ++q) s += p_.basis.weighted_derivatives(q, i) * fd(q, v);
                        y[p_.index(e, v, i)] -= s;
                    }
                }
            }
        }
        for (int f = 0; f < p_.cells; ++f) {
            const int next = (f + 1) % p_.cells;
            const State l = p_.sample(x, f, p_.basis.right), r = p_.sample(x, next, p_.basis.left);
            State flux;
            for (int v = 0; v < NV; ++v) {
                double a = 0, b = 0;
                for (int w = 0; w < NV; ++w) {
                    a += left_[f][v + NV * w] * l[w];
                    b += right_[f][v + NV * w] * r[w];
                }
                flux[v] = a + b;
            }
            for (int v = 0; v < NV; ++v)
                for (int i = 0; i < n; ++i) {
                    y[p_.index(f, v, i)] += p_.basis.right[i] * flux[v];
                    y[p_.index(next, v, i)] -= p_.basis.left[i] * flux[v];
                }
        }
    }
    // volume_jacobian.cpp:105-132: the volume part of block(row, row) plus M / tau
    void volume_block(int row, Mat &out) const {
        const int n = p_.nodes, Q = p_.basis.points, m = p_.m;
        out = Mat(m, m);
        if (n > 1)
            for (int w = 0; w < NV; ++w)
                for (int j = 0; j < n; ++j)
                    for (int v = 0; v < NV; ++v)
                        for (int i = 0; i < n; ++i) {
                            double value;
                            if (Q == n) {
                                value = -coef(row, j)[v + NV * w] * p_.basis.weighted_derivatives(j, i);
                            } else {
                                double s = 0;
                                for (int q = 0; q < Q; ++q)
                                    s += p_.basis.weighted_derivatives(q, i) *
                                         (coef(row, q)[v + NV * w] * p_.basis.values(q, j));
                                value = -s;
                            }
                            out(v * n + i, w * n + j) = value;
                        }
        for (int k = 0; k < m; ++k) out(k, k) += p_.mass[static_cast<size_t>(row) * m + k] * inverse_step_;
    }
    // dg_operator.cpp:50-79 with volume_jacobian.cpp:105-132; column-major m x m,
    // row/column index v * n + i
    void block(int row, int col, Mat &out) const {
        require(ready_, "Refresh the DG Jacobian before assembling a block");
        const int n = p_.nodes, Q = p_.basis.points, m = p_.m;
        out = Mat(m, m);
        if (row == col) {
            if (n > 1)
                for (int w = 0; w < NV; ++w)
                    for (int j = 0; j < n; ++j)
                        for (int v = 0; v < NV; ++v)
                            for (int i = 0; i < n; ++i) {
                                double value;
                                if (Q == n) {
                                    value = -coef(row, j)[v + NV * w] * p_.basis.weighted_derivatives(j, i);
                                } else {
                                    double s = 0;
                                    for (int q = 0; q < Q; ++q)
                                        s += p_.basis.weighted_derivatives(q, i) *
                                             (coef(row, q)[v + NV * w] * p_.basis.values(q, j));
                                    value = -s;
                                }
                                out(v * n + i, w * n + j) = value;
                            }
            for (int k = 0; k < m; ++k) out(k, k) += p_.mass[static_cast<size_t>(row) * m + k] * inverse_step_;
        }
        for (int side = 0; side < 2; ++side) {
            const int face = side == 0 ? (row + p_.cells - 1) % p_.cells : row, next = (face + 1) % p_.cells;
            if (col != face && col != next) continue;
            const Vec &test = side == 0 ? p_.basis.left : p_.basis.right;
            const double sign = side == 0 ? -1.0 : 1.0;
            for (int trace_side = 0; trace_side < 2; ++trace_side) {
                if (col != (trace_side == 0 ? face : next)) continue;
                const Vec &trial = trace_side == 0 ? p_.basis.right : p_.basis.left;
                const Jacobian &d = trace_side == 0 ? left_[face] : right_[face];
                for (int w = 0; w < NV; ++w)
                    for (int j = 0; j < n; ++j)
                        for (int v = 0; v < NV; ++v)
                            for (int i = 0; i < n; ++i) out(v * n + i, w * n + j) += (sign * trial[j] * d[v + NV * w]) * test[i];
            }
        }
    }

  private:
    const Problem &p_;
    Vec coef_;
    std::vector<Jacobian> left_, right_;
    double inverse_step_ = 0;
    bool ready_ = false;
    double *coef(int e, int q) { return coef_.data() + (static_cast<size_t>(e) * p_.basis.points + q) * NV * NV; }
    const double *coef(int e, int q) const {
        return coef_.data() + (static_cast<size_t>(e) * p_.basis.points + q) * NV * NV;
    }
};

// hybrid.cpp:3-139 (1D): static condensation of each cell onto its two face
// traces (3 values per face). Per cell A = volume block, B (m x 6), C (6 x m);
// X = A^-1 B, y = A^-1 r, S = -C X; the Schur operator is -(a+b) I per face
// plus the cells' S; recover du = -y - X dl.
class Hybrid {
  public:
    explicit Hybrid(const Problem &p)
        : p_(p), dg_(p), m_(p.m), X_(p.cells), C_(p.cells), S_(p.cells), B_(p.cells), y_(p.cells),
          rhs_(static_cast<size_t>(NV) * p.cells) {}
    int face(int cell, int side) const { return side == 0 ? (cell + p_.cells - 1) % p_.cells : cell; }
    void refresh_projected(const Vec &u, const Vec &r) {
        dg_.refresh(u);  // the volume Jacobian; the face part below replaces the DG faces
        std::fill(rhs_.begin(), rhs_.end(), 0.0);
        const int n = p_.nodes, m = m_;
        for (int cell = 0; cell < p_.cells; ++cell) {
            Mat A;
            dg_.volume_block(cell, A);
            Mat B(m, 2 * NV), C(2 * NV, m);
            for (int side = 0; side < 2; ++side) {
                const Speed sp = p_.speeds[face(cell, side)];
                const Vec &shape = side == 0 ? p_.basis.left : p_.basis.right;
                const double sign = side == 0 ? -1.0 : 1.0, st = side == 0 ? sp.b : sp.a;
                const Jacobian J = p_.example.gas.jacobian(p_.sample(u, cell, shape));
                Jacobian d;
                for (int k = 0; k < NV * NV; ++k) d[k] = sign * J[k] + st * ((k % 4 == 0) ? 1.0 : 0.0);
                for (int v = 0; v < NV; ++v)
                    for (int i = 0; i < n; ++i) {
                        const int row = v * n + i;
                        B(row, side * NV + v) = -shape[i] * st;
                        for (int w = 0; w < NV; ++w) {
                            C(side * NV + w, row) = shape[i] * d[w + NV * v];
                            for (int j = 0; j < n; ++j) A(row, w * n + j) += shape[i] * shape[j] * d[v + NV * w];
                        }
                    }
            }
            oracle::DenseLU lu;
            lu.compute(A);
            lu.solve_multiple(B, 2 * NV, X_[cell]);
            Vec y(r.begin() + static_cast<long>(cell) * m, r.begin() + static_cast<long>(cell + 1) * m);
            lu.solve_inplace(y.data());
            y_[cell] = y;
            Mat S(2 * NV, 2 * NV);
            for (int j = 0; j < 2 * NV; ++j)
                for (int a = 0; a < 2 * NV; ++a) {
                    double acc = 0;
                    for (int k = 0; k < m; ++k) acc += C(a, k) * X_[cell](k, j);
                    S(a, j) = -acc;
                }
            S_[cell] = S;
            C_[cell] = C;
            B_[cell] = B;
            for (int side = 0; side < 2; ++side)
                for (int w = 0; w < NV; ++w) {
                    double acc = 0;
                    for (int k = 0; k < m; ++k) acc += C(side * NV + w, k) * y[k];
                    rhs_[static_cast<size_t>(face(cell, side)) * NV + w] += acc;
                }
        }
        ready_ = true;
    }
    const Vec &rhs() const { return rhs_; }
    void apply(const Vec &x, Vec &y) const {  // hybrid.cpp:82-101
        require(ready_, "Refresh the hybrid operator first");
        y.assign(x.size(), 0.0);
        for (int f = 0; f < p_.cells; ++f)
            for (int v = 0; v < NV; ++v) y[f * NV + v] = -(p_.speeds[f].a + p_.speeds[f].b) * x[f * NV + v];
        for (int cell = 0; cell < p_.cells; ++cell) {
            double g[2 * NV];
            for (int side = 0; side < 2; ++side)
                for (int v = 0; v < NV; ++v) g[side * NV + v] = x[face(cell, side) * NV + v];
            for (int side = 0; side < 2; ++side)
                for (int w = 0; w < NV; ++w) {
                    double acc = 0;
                    for (int k = 0; k < 2 * NV; ++k) acc += S_[cell](side * NV + w, k) * g[k];
                    y[face(cell, side) * NV + w] += acc;
                }
        }
    }
    void block(int row, int col, Mat &out) const {  // hybrid.cpp:102-115
        out = Mat(NV, NV);
        if (row == col)
            for (int v = 0; v < NV; ++v) out(v, v) = -(p_.speeds[row].a + p_.speeds[row].b);
        for (int adjacent = 0; adjacent < 2; ++adjacent) {
            const int cell = (row + adjacent) % p_.cells, row_side = adjacent == 0 ? 1 : 0;
            for (int col_side = 0; col_side < 2; ++col_side)
                if (face(cell, col_side) == col)
                    for (int a = 0; a < NV; ++a)
                        for (int b = 0; b < NV; ++b) out(a, b) += S_[cell](row_side * NV + a, col_side * NV + b);
        }
    }
    void recover(const Vec &dl, Vec &du) const {  // hybrid.cpp:128-137
        du.assign(static_cast<size_t>(p_.size()), 0.0);
        for (int cell = 0; cell < p_.cells; ++cell)
            for (int k = 0; k < m_; ++k) {
                double acc = 0;
                for (int side = 0; side < 2; ++side)
                    for (int v = 0; v < NV; ++v) acc += X_[cell](k, side * NV + v) * dl[face(cell, side) * NV + v];
                du[static_cast<size_t>(cell) * m_ + k] = -y_[cell][k] - acc;
            }
    }
    void recovered_residual(const Vec &dl, Vec &out) const {  // hybrid.cpp:116-127
        Vec e;
        apply(dl, e);
        for (size_t k = 0; k < e.size(); ++k) e[k] -= rhs_[k];
        out.assign(static_cast<size_t>(p_.size()), 0.0);
        for (int cell = 0; cell < p_.cells; ++cell) {
            double g[2 * NV];
            for (int side = 0; side < 2; ++side) {
                const Speed sp = p_.speeds[face(cell, side)];
                for (int v = 0; v < NV; ++v) g[side * NV + v] = e[face(cell, side) * NV + v] / -(sp.a + sp.b);
            }
            for (int k = 0; k < m_; ++k) {
                double acc = 0;
                for (int j = 0; j < 2 * NV; ++j) acc += B_[cell](k, j) * g[j];
                out[static_cast<size_t>(cell) * m_ + k] = -acc;
            }
        }
    }

  private:
    const Problem &p_;
    DGOperator dg_;
    int m_;
    std::vector<Mat> X_, C_, S_, B_;
    std::vector<Vec> y_;
    Vec rhs_;
    bool ready_ = false;
};

struct Update {
    int step, stage, iteration, krylov;
    double time, stage_time, dt, cfl, rms, damping, linear_relative, dg_linear_relative;
};
struct Run {
    double final_time = 0, conservation_drift = 0, max_linear = 0, max_dg_linear = 0, min_step = 0, max_step = 0;
    int steps = 0, newton = 0, krylov = 0;
    std::vector<Update> updates;
};

// time_integrator.cpp:45-207 with GMRES(restart) + block Jacobi: DG on the
// element blocks (metric M^-1), or condensed HDG on the 3x3 face blocks
// (unweighted metric, time_integrator.cpp:36) with the 4-attempt Schur
// refinement of :137-153
Run advance(Problem &p, Vec &solution, bool hdg = false) {
    const Config &c = p.c;
    require(static_cast<int>(solution.size()) == p.size() && p.admissible(solution), "Invalid initial solution");
    require(c.steps > 0 || c.final_time < p.example.shock_time(), "Final time reaches simple-wave shock");
    constexpr double alpha = .24169426078821, beta = alpha / 4., eta = .12915286960590;
    constexpr double A[4][4] = {{alpha, 0, 0, 0}, {-alpha, alpha, 0, 0}, {0, 1 - alpha, alpha, 0},
                                {beta, eta, .5 - beta - eta - alpha, alpha}};
    constexpr double B[4] = {0, 1. / 6, 1. / 6, 2. / 3}, stage_c[4] = {alpha, 0, 1, .5};
    const int stages = c.integrator == "be" ? 1 : 4;
    const int D = p.size();
    Run stats;
    DGOperator dg(p);
    BlockJacobi pc(p.cells, p.m);
    GMRES gmres(D, c.restart, p.inverse_mass);
    Hybrid hybrid(p);
    BlockJacobi face_pc(p.cells, NV);
    GMRES trace_gmres(NV * p.cells, c.restart, Vec(static_cast<size_t>(NV) * p.cells, 1.0));
    Vec trace_direction(static_cast<size_t>(NV) * p.cells);
    p.freeze(solution);
    const State initial_integral = p.integral(solution);
    Vec state = solution, trial(D), residual(D), trial_residual(D), direction(D), rhs(D), step_state(D),
        stage_base(D), equation_error(D);
    std::vector<Vec> slopes(4, Vec(D));
    while (c.steps > 0 ? stats.steps < c.steps : stats.final_time < c.final_time) {
        const double rate = p.freeze(solution);
        double step = c.cfl / rate;
        if (c.steps == 0) {
            const double remaining = c.final_time - stats.final_time;
            step = std::min(step, remaining);
            if (remaining - step <= 16 * std::numeric_limits<double>::epsilon() * std::max(remaining, c.final_time))
                step = remaining;
        }
        require(std::isfinite(step) && step > 0 && stats.final_time + step > stats.final_time, "Invalid CFL step");
        require(stats.final_time + step < p.example.shock_time(), "Step reaches simple-wave shock");
        step_state = solution;
        state = solution;
        for (int stage = 0; stage < stages; ++stage) {
            const double diagonal = stages == 1 ? 1. : A[stage][stage];
            stage_base = step_state;
            for (int j = 0; j < stage; ++j)
                for (int k = 0; k < D; ++k) stage_base[k] += step * A[stage][j] * slopes[j][k];
            if (stage > 0) p.freeze(state);
            p.step_size = diagonal * step;
            p.old = stage_base;
            const double tau = p.step_size, old_norm = p.state_norm(stage_base);
            const int limit = c.newton_steps > 0 ? c.newton_steps : c.max_newton;
            double initial_rms = -1.;
            bool converged = false, residual_ready = false;
            for (int it = 0; it <= limit; ++it) {
                if (!residual_ready) p.dg_residual(state, residual);
                const double norm = p.dual_norm(residual), rms = norm / std::sqrt(2.);
                if (initial_rms < 0) initial_rms = rms;
                if (c.newton_steps == 0 && rms <= c.nonlinear_tolerance * std::max(1., initial_rms)) {
                    converged = true;
                    break;
                }
                if (it == limit) break;
                const double allowance = 128 * std::numeric_limits<double>::epsilon() *
                                         ((p.state_norm(state) + old_norm) / tau + norm);
                const double target = std::max(10 * c.linear_tolerance * norm, allowance);
                LinearResult linear{0, 0};
                if (!hdg) {
                    dg.refresh(state);
                    for (int k = 0; k < D; ++k) rhs[k] = -residual[k];
                    pc.refresh([&](int r, int cl, Mat &blk) { dg.block(r, cl, blk); });
                    linear = gmres.solve([&](const Vec &a, Vec &b) { dg.apply(a, b); },
                                         [&](const Vec &a, Vec &b) { pc.apply(a, b); }, rhs, direction,
                                         c.max_linear, c.linear_tolerance);
                    const Vec &gr = gmres.residual();
                    for (int k = 0; k < D; ++k) equation_error[k] = -gr[k];
                } else {
                    hybrid.refresh_projected(state, residual);
                    face_pc.refresh([&](int r, int cl, Mat &blk) { hybrid.block(r, cl, blk); });
                    double tolerance = c.linear_tolerance;
                    bool valid = false;
                    for (int attempt = 0; attempt < 4; ++attempt) {
                        const LinearResult cur =
                            trace_gmres.solve([&](const Vec &a, Vec &b) { hybrid.apply(a, b); },
                                              [&](const Vec &a, Vec &b) { face_pc.apply(a, b); }, hybrid.rhs(),
                                              trace_direction, c.max_linear, tolerance);
                        linear.iterations += cur.iterations;
                        linear.relative_residual = cur.relative_residual;
                        hybrid.recover(trace_direction, direction);
                        hybrid.recovered_residual(trace_direction, equation_error);
                        valid = p.dual_norm(equation_error) <= target;
                        if (valid) break;
                        tolerance *= .01;
                    }
                    require(valid, "Recovered DG residual criterion failed after Schur refinement");
                }
                const double error_norm = p.dual_norm(equation_error);
                require(error_norm <= target, "DG Newton linear residual too large");
                double damping = 1.;
                bool accepted = false;
                for (int search = 0; search < 30; ++search) {
                    for (int k = 0; k < D; ++k) trial[k] = state[k] + damping * direction[k];
                    if (p.admissible(trial)) {
                        p.dg_residual(trial, trial_residual);
                        if (p.dual_norm(trial_residual) <= (1 - 1e-4 * damping) * norm + allowance) {
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
                stats.max_linear = std::max(stats.max_linear, linear.relative_residual);
                const double dgrel = norm > 0 ? error_norm / norm : 0.;
                stats.max_dg_linear = std::max(stats.max_dg_linear, dgrel);
                stats.krylov += linear.iterations;
                ++stats.newton;
                stats.updates.push_back({stats.steps, stage, it, linear.iterations, stats.final_time + step,
                                         stats.final_time + (stages == 1 ? 1. : stage_c[stage]) * step, step,
                                         step * rate, rms, damping, linear.relative_residual, dgrel});
            }
            require(converged || c.newton_steps > 0, "Newton iteration limit reached");
            for (int k = 0; k < D; ++k) slopes[stage][k] = (state[k] - stage_base[k]) / tau;
        }
        if (stages == 1) solution = state;
        else {
            solution = step_state;
            for (int stage = 0; stage < stages; ++stage)
                for (int k = 0; k < D; ++k) solution[k] += step * B[stage] * slopes[stage][k];
            require(p.admissible(solution), "DIRK3 step produced an inadmissible state");
        }
        stats.final_time =
            c.steps == 0 && step >= c.final_time - stats.final_time ? c.final_time : stats.final_time + step;
        stats.min_step = stats.steps == 0 ? step : std::min(stats.min_step, step);
        stats.max_step = std::max(stats.max_step, step);
        ++stats.steps;
        const State now = p.integral(solution);
        double drift = 0;
        for (int v = 0; v < NV; ++v) drift = std::max(drift, std::abs(now[v] - initial_integral[v]));
        stats.conservation_drift = std::max(stats.conservation_drift, drift);
    }
    return stats;
}

struct Context {
    Problem problem;
    DGOperator dg;
    BlockJacobi pc;
    Hybrid hybrid;
    BlockJacobi face_pc;
    Run last;
    std::string error;
    explicit Context(const Config &c)
        : problem(c), dg(problem), pc(c.cells, NV * (c.degree + 1)), hybrid(problem), face_pc(c.cells, NV) {}
};
}  // namespace oracle1d

// ------------------------------------------------------------------ C API
// Vectors are in the reference layout: index cell * 3n + v * n + i.
using oracle1d::Context;
using oracle1d::Vec;
namespace {
thread_local std::string g_error1;
template <class F> int guarded(Context *c, F &&f) {
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        if (c) c->error = e.what();
        g_error1 = e.what();
        return 1;
    }
}
Vec in(const Context *c, const double *x) { return Vec(x, x + c->problem.size()); }
void out(const Vec &v, double *y) { std::memcpy(y, v.data(), v.size() * sizeof(double)); }
}  // namespace

extern "C" {
void *orc1_create(int argc, const char *const *argv) {
    try {
        return new Context(oracle1d::parse(argc, argv));
    } catch (const std::exception &e) {
        g_error1 = e.what();
        return nullptr;
    }
}
void orc1_destroy(void *c) { delete static_cast<Context *>(c); }
const char *orc1_last_error() { return g_error1.c_str(); }
int orc1_size(void *c) { return static_cast<Context *>(c)->problem.size(); }
int orc1_nodes(void *c) { return static_cast<Context *>(c)->problem.nodes; }
double orc1_h(void *c) { return static_cast<Context *>(c)->problem.h; }
int orc1_initial(void *h, double *u) {
    auto *c = static_cast<Context *>(h);
    return guarded(c, [&] {
        Vec v;
        c->problem.initial(v);
        out(v, u);
    });
}
int orc1_freeze(void *h, const double *u, double *rate, double *speeds) {
    auto *c = static_cast<Context *>(h);
    return guarded(c, [&] {
        *rate = c->problem.freeze(in(c, u));
        for (int f = 0; f < c->problem.cells; ++f) {
            speeds[2 * f] = c->problem.speeds[f].a;
            speeds[2 * f + 1] = c->problem.speeds[f].b;
        }
    });
}
int orc1_set_step(void *h, double tau, const double *old) {
    auto *c = static_cast<Context *>(h);
    return guarded(c, [&] {
        c->problem.step_size = tau;
        c->problem.old = in(c, old);
    });
}
int orc1_residual(void *h, const double *u, double *r) {
    auto *c = static_cast<Context *>(h);
    return guarded(c, [&] {
        Vec v;
        c->problem.dg_residual(in(c, u), v);
        out(v, r);
    });
}
int orc1_admissible(void *h, const double *u) {
    auto *c = static_cast<Context *>(h);
    return c->problem.admissible(in(c, u)) ? 1 : 0;
}
int orc1_refresh(void *h, const double *u) {
    auto *c = static_cast<Context *>(h);
    return guarded(c, [&] { c->dg.refresh(in(c, u)); });
}
int orc1_apply(void *h, const double *x, double *y) {
    auto *c = static_cast<Context *>(h);
    return guarded(c, [&] {
        Vec v;
        c->dg.apply(in(c, x), v);
        out(v, y);
    });
}
int orc1_block(void *h, int row, int col, double *outm) {
    auto *c = static_cast<Context *>(h);
    return guarded(c, [&] {
        oracle::Mat b;
        c->dg.block(row, col, b);
        out(b.a, outm);
    });
}
int orc1_bj_refresh(void *h) {
    auto *c = static_cast<Context *>(h);
    return guarded(c, [&] { c->pc.refresh([&](int r, int cl, oracle::Mat &b) { c->dg.block(r, cl, b); }); });
}
int orc1_bj_apply(void *h, const double *r, double *z) {
    auto *c = static_cast<Context *>(h);
    return guarded(c, [&] {
        Vec v;
        c->pc.apply(in(c, r), v);
        out(v, z);
    });
}
int orc1_gmres(void *h, const double *rhs, double *x, int restart, int limit, double tol, int *its, double *rel) {
    auto *c = static_cast<Context *>(h);
    return guarded(c, [&] {
        oracle1d::GMRES g(c->problem.size(), restart, c->problem.inverse_mass);
        Vec sol;
        const auto res = g.solve([&](const Vec &a, Vec &b) { c->dg.apply(a, b); },
                                 [&](const Vec &a, Vec &b) { c->pc.apply(a, b); }, in(c, rhs), sol, limit, tol);
        out(sol, x);
        *its = res.iterations;
        *rel = res.relative_residual;
    });
}
int orc1_error(void *h, const double *u, double t, double *err3) {
    auto *c = static_cast<Context *>(h);
    return guarded(c, [&] {
        const auto e = c->problem.error(in(c, u), t);
        for (int v = 0; v < 3; ++v) err3[v] = e[v];
    });
}
int orc1_integral(void *h, const double *u, double *out3) {
    auto *c = static_cast<Context *>(h);
    return guarded(c, [&] {
        const auto e = c->problem.integral(in(c, u));
        for (int v = 0; v < 3; ++v) out3[v] = e[v];
    });
}
// Whole run from u (overwritten by the final state). summary: final_time,
// steps, newton, krylov, drift, max_linear, max_dg_linear, min_step, max_step.
// updates: up to cap rows of (step, stage, iteration, krylov, time, stage_time,
// dt, cfl, rms, damping, linear_relative, dg_linear_relative).
// condensed HDG (hybrid.cpp, 1D): trace vectors are face * 3 + v
int orc1_hdg_refresh(void *h, const double *u, const double *r) {
    auto *c = static_cast<Context *>(h);
    return guarded(c, [&] {
        c->hybrid.refresh_projected(in(c, u), in(c, r));
        c->face_pc.refresh([&](int a, int b, oracle::Mat &blk) { c->hybrid.block(a, b, blk); });
    });
}
int orc1_hdg_rhs(void *h, double *out3c) {
    auto *c = static_cast<Context *>(h);
    return guarded(c, [&] { out(c->hybrid.rhs(), out3c); });
}
int orc1_hdg_apply(void *h, const double *x, double *y) {
    auto *c = static_cast<Context *>(h);
    return guarded(c, [&] {
        Vec v;
        c->hybrid.apply(Vec(x, x + 3 * c->problem.cells), v);
        out(v, y);
    });
}
int orc1_hdg_block(void *h, int row, int col, double *out9) {
    auto *c = static_cast<Context *>(h);
    return guarded(c, [&] {
        oracle::Mat b;
        c->hybrid.block(row, col, b);
        out(b.a, out9);
    });
}
int orc1_hdg_bj_apply(void *h, const double *r, double *z) {
    auto *c = static_cast<Context *>(h);
    return guarded(c, [&] {
        Vec v;
        c->face_pc.apply(Vec(r, r + 3 * c->problem.cells), v);
        out(v, z);
    });
}
int orc1_hdg_recover(void *h, const double *dl, double *du) {
    auto *c = static_cast<Context *>(h);
    return guarded(c, [&] {
        Vec v;
        c->hybrid.recover(Vec(dl, dl + 3 * c->problem.cells), v);
        out(v, du);
    });
}
int orc1_hdg_recovered_residual(void *h, const double *dl, double *res) {
    auto *c = static_cast<Context *>(h);
    return guarded(c, [&] {
        Vec v;
        c->hybrid.recovered_residual(Vec(dl, dl + 3 * c->problem.cells), v);
        out(v, res);
    });
}
int orc1_advance(void *h, int hdg, double *u, double *summary, double *updates, int cap, int *count) {
    auto *c = static_cast<Context *>(h);
    return guarded(c, [&] {
        Vec v = in(c, u);
        c->last = oracle1d::advance(c->problem, v, hdg != 0);
        out(v, u);
        const auto &s = c->last;
        const double sm[9] = {s.final_time, double(s.steps), double(s.newton), double(s.krylov), s.conservation_drift,
                              s.max_linear, s.max_dg_linear, s.min_step, s.max_step};
        std::memcpy(summary, sm, sizeof(sm));
        *count = static_cast<int>(s.updates.size());
        for (int k = 0; k < std::min(cap, *count); ++k) {
            const auto &w = s.updates[k];
            const double row[12] = {double(w.step), double(w.stage), double(w.iteration), double(w.krylov), w.time,
                                    w.stage_time, w.dt, w.cfl, w.rms, w.damping, w.linear_relative,
                                    w.dg_linear_relative};
            std::memcpy(updates + 12 * k, row, sizeof(row));
        }
    });
}
}
