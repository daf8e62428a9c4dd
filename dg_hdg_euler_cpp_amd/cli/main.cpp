// dg_hdg_2d: the reference command line (dg_hdg_2d/src/main.cpp:7-107) on the
// device path. Same flags, same four CSV schemas (_summary, _iterations,
// _solution, _difference), same "ERROR: <message>" / exit 1 contract; the
// solver runs through the C ABI of libdgh_device.so.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dgh/dgh_device.h"

namespace {
void check(int rc, const dgh_ctx *ctx) {
    if (rc != DGH_OK) throw std::runtime_error(dgh_last_error(ctx));
}
void hip(hipError_t e) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e));
}
struct DeviceVector {
    double *p = nullptr;
    explicit DeviceVector(size_t n) { hip(hipMalloc(&p, n * sizeof(double))); }
    ~DeviceVector() { (void)hipFree(p); }
};
const char *integrator_name(int i) { return i == 0 ? "be" : "ssp-dirk3"; }

// example.cpp:20-39 at time t (simple wave) and :12-16 (entropy wave), for the
// _solution.csv exact columns.
void exact(const dgh_config &c, double lx, double ly, double x, double y, double t, double out[4]) {
    (void)lx;
    (void)ly;
    const double g = 3.0, pi = std::acos(-1.);
    if (c.testcase == DGH_SEEDED) {
        for (int v = 0; v < 4; ++v) out[v] = std::numeric_limits<double>::quiet_NaN();
        return;
    }
    if (c.testcase == DGH_WAVE) {
        const double rho = 1 + .2 * std::sin(pi * (x - t)) * std::sin(pi * (y - .5 * t));
        out[0] = rho; out[1] = rho; out[2] = .5 * rho; out[3] = 1 / (g - 1) + .625 * rho;
        return;
    }
    const double cosine = std::cos(45. * pi / 180.), sine = std::sin(45. * pi / 180.);
    const bool diagonal = c.testcase == DGH_DIAGONAL;
    if (diagonal) x = cosine * x + sine * y;
    const double a = (3 - g) / 2, b = (g + 1) / 4;
    double lo = x - t * (a + b), hi = x - t * (a - b), xi = x - t * a;
    for (int it = 0; it < 80; ++it) {
        const double f = xi + t * (a + b * std::sin(pi * xi)) - x;
        if (std::abs(f) < 2e-15) break;
        if (f > 0) hi = xi; else lo = xi;
        const double z = xi - f / (1 + t * b * pi * std::cos(pi * xi));
        xi = z > lo && z < hi ? z : .5 * (lo + hi);
    }
    const double jm = std::sin(pi * xi), velocity = .5 * (2 + jm), sound = (g - 1) * (2 - jm) / 4;
    const double rho = std::pow(sound * sound / g, 1 / (g - 1));
    out[0] = rho;
    out[1] = rho * velocity * (diagonal ? cosine : 1.);
    out[2] = diagonal ? rho * velocity * sine : 0.;
    out[3] = std::pow(rho, g) / (g - 1) + .5 * rho * velocity * velocity;
}
}  // namespace

int main(int argc, char **argv) {
    try {
        dgh_config config;
        char err[512] = {0};
        for (int i = 1; i < argc; ++i)
            if (std::string(argv[i]) == "--help") {
                dgh_config_parse(2, const_cast<const char *const *>(argv + i - 1), &config, err, sizeof(err));
                return 0;
            }
        if (dgh_config_parse(argc, argv, &config, err, sizeof(err)) != DGH_OK) throw std::runtime_error(err);
        const std::string prefix = config.prefix;
        const auto path = std::filesystem::path(prefix);
        if (path.has_parent_path()) std::filesystem::create_directories(path.parent_path());
        std::ofstream summary(prefix + "_summary.csv"), updates(prefix + "_iterations.csv");
        if (!summary || !updates) throw std::runtime_error("Cannot open output files");
        summary << std::setprecision(17)
                << "method,nx,ny,k,quad,cfl,integrator,linear,preconditioner,ordering,ntol,ltol,safety,"
                   "fixed_newton_steps,warmup_updates,steps,final_time,newton_iterations,krylov_iterations,"
                   "max_linear_relative_residual,max_dg_linear_relative_residual,min_step,max_step,mean_dt,"
                   "time_loop_s,measured_updates,mean_newton_s,volume_dofs,trace_dofs,"
                   "rho_L2,momentum_x_L2,momentum_y_L2,energy_L2,max_conservation_drift,lx,ly,jacobi_local\n";
        updates << std::setprecision(17)
                << "method,step,stage,iteration,time,stage_time,dt,cfl,residual_physical_rms,damping,"
                   "krylov_iterations,linear_relative_residual,dg_linear_relative_residual,iteration_s,measured\n";
        std::ofstream solution_file;
        if (!config.timing_only && config.save_solution) {
            solution_file.open(prefix + "_solution.csv");
            if (!solution_file) throw std::runtime_error("Cannot open solution output");
            solution_file << std::setprecision(17)
                          << "method,cell,node,x,y,rho,momentum_x,momentum_y,energy,rho_exact,momentum_x_exact,"
                             "momentum_y_exact,energy_exact\n";
        }
        std::vector<double> dg_solution;
        bool have_dg = false;
        for (int method : {DGH_METHOD_DG, DGH_METHOD_HDG}) {
            const std::string name = method == DGH_METHOD_DG ? "dg" : "hdg";
            if (config.method != DGH_METHOD_BOTH && config.method != method) continue;
            dgh_config run_cfg = config;
            run_cfg.method = method;
            dgh_ctx *ctx = nullptr;
            check(dgh_create(&run_cfg, &ctx), nullptr);
            struct Guard { dgh_ctx *c; ~Guard() { dgh_destroy(c); } } guard{ctx};
            int64_t dims[10];
            double geo[5];
            dgh_dims(ctx, dims);
            dgh_geometry(ctx, geo);
            const size_t size = static_cast<size_t>(dims[0]);
            const int n = static_cast<int>(dims[3]), cells = static_cast<int>(dims[2]), plane = n * n;
            DeviceVector u(size);
            check(dgh_initial(ctx, u.p), ctx);
            dgh_run_stats stats{};
            const int capacity = 1 << 16;
            std::vector<dgh_update_stats> list(capacity);
            int count = 0;
            check(dgh_advance(ctx, u.p, &stats, list.data(), capacity, &count), ctx);
            double error[4];
            if (config.timing_only)
                for (double &e : error) e = std::numeric_limits<double>::quiet_NaN();
            else
                check(dgh_error_l2(ctx, u.p, stats.final_time, error), ctx);
            summary << name << ',' << config.nx << ',' << config.ny << ',' << config.degree << ','
                    << (config.quadrature ? config.quadrature : config.degree + 1) << ',' << config.cfl << ','
                    << integrator_name(config.integrator) << ',' << (config.direct ? "direct" : "gmres") << ','
                    << (config.direct ? "none" : config.precond ? "block-ilu0" : "block-jacobi") << ','
                    << (config.direct ? "none" : config.ordering == 2 ? "redblack" : "natural") << ','
                    << config.nonlinear_tolerance << ',' << config.linear_tolerance << ',' << config.safety << ','
                    << config.newton_steps << ',' << config.warmup_updates << ',' << stats.steps << ','
                    << stats.final_time << ',' << stats.newton_iterations << ',' << stats.krylov_iterations << ','
                    << stats.max_linear_residual << ',' << stats.max_dg_linear_residual << ',' << stats.minimum_step
                    << ',' << stats.maximum_step << ',' << (stats.steps ? stats.final_time / stats.steps : 0.) << ','
                    << stats.seconds << ',' << stats.measured_updates << ','
                    << (stats.measured_updates ? stats.measured_newton_seconds / stats.measured_updates : 0.) << ','
                    << dims[0] << ',' << dims[1];
            for (double e : error) summary << ',' << e;
            summary << ',' << stats.conservation_drift << ',' << geo[2] << ',' << geo[3] << ','
                    << (method == DGH_METHOD_DG ? "dense" : "face-dense") << '\n';
            for (int i = 0; i < count && i < capacity; ++i) {
                const auto &s = list[i];
                updates << name << ',' << s.step << ',' << s.stage << ',' << s.iteration << ',' << s.time << ','
                        << s.stage_time << ',' << s.dt << ',' << s.cfl << ',' << s.residual_physical_rms << ','
                        << s.damping << ',' << s.krylov_iterations << ',' << s.linear_relative_residual << ','
                        << s.dg_linear_relative_residual << ',' << s.iteration_s << ','
                        << (i >= config.warmup_updates) << '\n';
            }
            std::vector<double> host(size), xs(n), ws(n);
            hip(hipMemcpy(host.data(), u.p, size * sizeof(double), hipMemcpyDeviceToHost));
            check(dgh_basis(ctx, xs.data(), ws.data()), ctx);
            if (!config.timing_only && config.save_solution) {
                for (int cell = 0; cell < cells; ++cell)
                    for (int i = 0; i < plane; ++i) {
                        const double x = -.5 * geo[2] + geo[0] * (cell % config.nx + .5) + .5 * geo[0] * xs[i % n];
                        const double y = -.5 * geo[3] + geo[1] * (cell / config.nx + .5) + .5 * geo[1] * xs[i / n];
                        double ex[4];
                        exact(config, geo[2], geo[3], x, y, stats.final_time, ex);
                        solution_file << name << ',' << cell << ',' << i << ',' << x << ',' << y;
                        for (int v = 0; v < 4; ++v) solution_file << ',' << host[static_cast<size_t>(cell) * 4 * plane + v * plane + i];
                        for (int v = 0; v < 4; ++v) solution_file << ',' << ex[v];
                        solution_file << '\n';
                    }
            }
            std::cout << name << ": steps=" << stats.steps << ", Newton=" << stats.newton_iterations
                      << ", Krylov=" << stats.krylov_iterations << ", time=" << stats.seconds << " s, mean Newton="
                      << (stats.measured_updates ? stats.measured_newton_seconds / stats.measured_updates : 0.)
                      << " s, conservation drift=" << stats.conservation_drift << '\n';
            if (method == DGH_METHOD_DG && config.method == DGH_METHOD_BOTH) {
                dg_solution = host;
                have_dg = true;
            } else if (have_dg) {
                double maximum = 0, joint = 0;
                std::ofstream comparison(prefix + "_difference.csv");
                if (!comparison) throw std::runtime_error("Cannot open comparison output");
                comparison << std::setprecision(17) << "rho_L2,momentum_x_L2,momentum_y_L2,energy_L2,joint_L2,max_nodal\n";
                for (size_t i = 0; i < size; ++i) maximum = std::max(maximum, std::abs(dg_solution[i] - host[i]));
                for (int v = 0; v < 4; ++v) {
                    double squared = 0;
                    for (int cell = 0; cell < cells; ++cell)
                        for (int k = 0; k < plane; ++k) {
                            const size_t idx = static_cast<size_t>(cell) * 4 * plane + v * plane + k;
                            const double d = dg_solution[idx] - host[idx];
                            squared += d * d * (.25 * geo[0] * geo[1] * ws[k % n] * ws[k / n]);
                        }
                    comparison << std::sqrt(squared) << ',';
                    joint += squared;
                }
                comparison << std::sqrt(joint) << ',' << maximum << '\n';
                std::cout << "DG/HDG max difference=" << maximum << '\n';
                if (!(maximum < 1e-8)) throw std::runtime_error("DG/HDG solutions disagree");
            }
        }
    } catch (const std::exception &error) {
        std::cerr << "ERROR: " << error.what() << '\n';
        return 1;
    }
    return 0;
}
