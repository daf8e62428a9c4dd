// dg_hdg_1d: the reference's 1D command line (dg_hdg_1d/src/main.cpp:7-107) on
// the device path. Same flags, the same _summary / _iterations / _solution CSV
// schemas (plus _difference.csv for --method both) and the same
// "ERROR: <message>" / exit 1 contract; the solver runs through the 1D C ABI
// of libdgh_device.so (include/dgh/dgh1d_device.h), one context per method.
#include <hip/hip_runtime.h>

#include <cmath>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dgh/dgh1d_device.h"

namespace {
void check(int rc, const dgh1_ctx *ctx) {
    if (rc != DGH_OK) throw std::runtime_error(ctx ? dgh1_last_error(ctx) : dgh1_create_error());
}
void hip(hipError_t e) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e));
}
// simple_wave.cpp:3-25 at (x, t), gamma = 3, for the _solution.csv exact columns
void exact(double x, double t, double out[3]) {
    const double g = 3.0, pi = std::acos(-1.), a = (3 - g) / 2, b = (g + 1) / 4;
    double lo = x - t * (a + b), hi = x - t * (a - b), xi = x - t * a;
    for (int it = 0; it < 80; ++it) {
        const double f = xi + t * (a + b * std::sin(pi * xi)) - x;
        if (std::abs(f) < 2e-15) break;
        if (f > 0) hi = xi;
        else lo = xi;
        const double z = xi - f / (1 + t * b * pi * std::cos(pi * xi));
        xi = z > lo && z < hi ? z : .5 * (lo + hi);
    }
    const double jm = std::sin(pi * xi), v = .5 * (2 + jm), c = (g - 1) * (2 - jm) / 4;
    const double rho = std::pow(c * c / g, 1 / (g - 1)), p = std::pow(rho, g);
    out[0] = rho;
    out[1] = rho * v;
    out[2] = p / (g - 1) + .5 * rho * v * v;
}
}  // namespace

int main(int argc, char **argv) {
    try {
        for (int i = 1; i < argc; ++i)
            if (std::string(argv[i]) == "--help") {
                std::cout << "1D periodic Euler DG / condensed HDG on the device\n"
                             "--cells N --k K --quad Q --cfl C (--T T | --steps N)\n"
                             "--method dg|hdg|both --linear gmres --precond block-jacobi --ordering natural\n"
                             "--restart N --max-linear N --max-newton N --ntol R --ltol R --safety S --prefix PATH\n"
                             "--integrator ssp-dirk3|be --newton-steps N --warmup-updates N --timing-only\n";
                return 0;
            }
        dgh1_config config;
        check(dgh1_parse_config(argc - 1, argv + 1, &config), nullptr);
        const std::string prefix = config.prefix;
        const auto path = std::filesystem::path(prefix);
        if (path.has_parent_path()) std::filesystem::create_directories(path.parent_path());
        std::ofstream summary(prefix + "_summary.csv"), updates(prefix + "_iterations.csv");
        if (!summary || !updates) throw std::runtime_error("Cannot open output files");
        summary << std::setprecision(17)
                << "method,cells,k,quad,cfl,integrator,linear,preconditioner,ordering,ntol,ltol,safety,"
                   "fixed_newton_steps,warmup_updates,steps,final_time,newton_iterations,krylov_iterations,"
                   "max_linear_relative_residual,max_dg_linear_relative_residual,min_step,max_step,mean_dt,"
                   "time_loop_s,measured_updates,mean_newton_s,volume_dofs,trace_dofs,"
                   "rho_L2,momentum_L2,energy_L2,max_conservation_drift\n";
        updates << std::setprecision(17)
                << "method,step,stage,iteration,time,stage_time,dt,cfl,residual_physical_rms,damping,"
                   "krylov_iterations,linear_relative_residual,dg_linear_relative_residual,iteration_s,measured\n";
        std::ofstream solution;
        if (!config.timing_only) {
            solution.open(prefix + "_solution.csv");
            if (!solution) throw std::runtime_error("Cannot open solution output");
            solution << std::setprecision(17)
                     << "method,cell,node,x,rho,momentum,energy,rho_exact,momentum_exact,energy_exact\n";
        }
        const int n = config.degree + 1, cells = config.cells;
        const double h = 2.0 / cells;
        std::vector<double> dg_solution, weights;
        bool have_dg = false;
        for (int method : {DGH_METHOD_DG, DGH_METHOD_HDG}) {
            if (config.method != DGH_METHOD_BOTH && config.method != method) continue;
            const std::string name = method == DGH_METHOD_DG ? "dg" : "hdg";
            dgh1_config run_cfg = config;
            run_cfg.method = method;
            dgh1_ctx *ctx = dgh1_create(&run_cfg);
            check(ctx ? DGH_OK : DGH_ERROR, nullptr);
            struct Guard {
                dgh1_ctx *c;
                ~Guard() { dgh1_destroy(c); }
            } guard{ctx};
            const size_t size = static_cast<size_t>(dgh1_size(ctx));
            double *u = nullptr;
            hip(hipMalloc(&u, size * sizeof(double)));
            struct Free {
                double *p;
                ~Free() { (void)hipFree(p); }
            } free_u{u};
            check(dgh1_initial(ctx, u), ctx);
            dgh_run_stats stats{};
            const int capacity = 1 << 16;
            std::vector<dgh_update_stats> list(capacity);
            int count = 0;
            check(dgh1_advance(ctx, u, &stats, list.data(), capacity, &count), ctx);
            double error[3];
            if (config.timing_only)
                for (double &e : error) e = std::numeric_limits<double>::quiet_NaN();
            else
                check(dgh1_error(ctx, u, stats.final_time, error), ctx);
            summary << name << ',' << cells << ',' << config.degree << ',' << (config.quadrature ? config.quadrature : n)
                    << ',' << config.cfl << ',' << (config.integrator == 0 ? "be" : "ssp-dirk3")
                    << ",gmres,block-jacobi,natural," << config.nonlinear_tolerance << ',' << config.linear_tolerance
                    << ',' << config.safety << ',' << config.newton_steps << ',' << config.warmup_updates << ','
                    << stats.steps << ',' << stats.final_time << ',' << stats.newton_iterations << ','
                    << stats.krylov_iterations << ',' << stats.max_linear_residual << ','
                    << stats.max_dg_linear_residual << ',' << stats.minimum_step << ',' << stats.maximum_step << ','
                    << (stats.steps ? stats.final_time / stats.steps : 0.) << ',' << stats.seconds << ','
                    << stats.measured_updates << ','
                    << (stats.measured_updates ? stats.measured_newton_seconds / stats.measured_updates : 0.) << ','
                    << size << ',' << cells * 3;
            for (double e : error) summary << ',' << e;
            summary << ',' << stats.conservation_drift << '\n';
            for (int i = 0; i < count && i < capacity; ++i) {
                const auto &s = list[i];
                updates << name << ',' << s.step << ',' << s.stage << ',' << s.iteration << ',' << s.time << ','
                        << s.stage_time << ',' << s.dt << ',' << s.cfl << ',' << s.residual_physical_rms << ','
                        << s.damping << ',' << s.krylov_iterations << ',' << s.linear_relative_residual << ','
                        << s.dg_linear_relative_residual << ',' << s.iteration_s << ','
                        << (i >= config.warmup_updates) << '\n';
            }
            std::vector<double> host(size), xs(n);
            hip(hipMemcpy(host.data(), u, size * sizeof(double), hipMemcpyDeviceToHost));
            check(dgh1_nodes(ctx, xs.data()), ctx);
            if (!config.timing_only)
                for (int cell = 0; cell < cells; ++cell)
                    for (int i = 0; i < n; ++i) {
                        const double x = -1 + h * (cell + .5) + .5 * h * xs[i];
                        double ex[3];
                        exact(x, stats.final_time, ex);
                        solution << name << ',' << cell << ',' << i << ',' << x;
                        for (int v = 0; v < 3; ++v) solution << ',' << host[static_cast<size_t>(cell) * 3 * n + v * n + i];
                        for (int v = 0; v < 3; ++v) solution << ',' << ex[v];
                        solution << '\n';
                    }
            std::cout << name << ": steps=" << stats.steps << ", Newton=" << stats.newton_iterations
                      << ", Krylov=" << stats.krylov_iterations << ", time=" << stats.seconds << " s, mean Newton="
                      << (stats.measured_updates ? stats.measured_newton_seconds / stats.measured_updates : 0.)
                      << " s, conservation drift=" << stats.conservation_drift << '\n';
            if (method == DGH_METHOD_DG) {
                dg_solution = host;
                have_dg = true;
            } else if (have_dg) {
                // main.cpp:78-101: mass-weighted L2 of DG - HDG per variable, joint, max nodal
                std::vector<double> w(n);
                check(dgh1_weights(ctx, w.data()), ctx);
                double maximum = 0, joint = 0;
                std::ofstream comparison(prefix + "_difference.csv");
                if (!comparison) throw std::runtime_error("Cannot open comparison output");
                comparison << std::setprecision(17) << "rho_L2,momentum_L2,energy_L2,joint_L2,max_nodal\n";
                for (size_t k = 0; k < size; ++k) maximum = std::max(maximum, std::abs(dg_solution[k] - host[k]));
                for (int v = 0; v < 3; ++v) {
                    double squared = 0;
                    for (int cell = 0; cell < cells; ++cell)
                        for (int i = 0; i < n; ++i) {
                            const size_t k = static_cast<size_t>(cell) * 3 * n + v * n + i;
                            const double d = dg_solution[k] - host[k];
                            squared += d * d * (.5 * h * w[i]);
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
