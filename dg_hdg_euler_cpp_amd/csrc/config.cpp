// Reference CLI semantics (dg_hdg_2d/src/config.cpp:6-122, config.hpp:7-18)
// plus the device-path flags --device, --precision, --inner-restart, --seed.
#include <cmath>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>

#include "dgh_host.hpp"

namespace dgh {
namespace {
void require(bool condition, const char *message) {
    if (!condition) throw std::runtime_error(message);
}
}  // namespace

dgh_config default_config() {
    dgh_config c;
    std::memset(&c, 0, sizeof(c));
    c.nx = 8;
    c.ny = 8;
    c.degree = 3;
    c.quadrature = 0;
    c.restart = 30;
    c.max_linear = 1500;
    c.max_newton = 30;
    c.steps = 0;
    c.newton_steps = 0;
    c.warmup_updates = 0;
    c.cfl = 2.;
    c.final_time = .2;
    c.safety = 1.1;
    c.nonlinear_tolerance = 1e-12;
    c.linear_tolerance = 1e-12;
    c.save_solution = 1;
    c.integrator = 1;
    c.method = DGH_METHOD_BOTH;
    c.testcase = DGH_DIAGONAL;
    c.precision = DGH_PRECISION_FP64;
    c.inner_restart = 30;
    c.inner_max = 300;
    c.inner_tolerance = 1e-3;
    c.seed = 2025;
    c.graph = 1;
    std::strncpy(c.prefix, "output/run_2d", sizeof(c.prefix) - 1);
    return c;
}

void validate_config(const dgh_config &c) {
    require(c.nx >= 3 && c.ny >= 3 && c.degree >= 0 && c.degree <= 8 && c.quadrature >= 0, "Invalid mesh or degree");
    require(c.testcase >= DGH_DIAGONAL && c.testcase <= DGH_SEEDED, "Invalid testcase");
    require(c.quadrature == 0 || c.quadrature >= c.degree + 1, "Quadrature must be >= degree+1");
    require(c.restart > 0 && c.max_linear > 0 && c.max_newton > 0 && c.steps >= 0, "Invalid iteration limit");
    require(c.newton_steps >= 0 && c.warmup_updates >= 0, "Invalid timing update count");
    require(c.integrator == 0 || c.integrator == 1, "Invalid integrator");
    require(std::isfinite(c.cfl) && c.cfl > 0 && std::isfinite(c.final_time) && c.final_time >= 0,
            "CFL must be positive and final time nonnegative");
    require(std::isfinite(c.safety) && c.safety >= 1 && std::isfinite(c.nonlinear_tolerance) &&
                c.nonlinear_tolerance > 0 && std::isfinite(c.linear_tolerance) && c.linear_tolerance > 0 &&
                c.linear_tolerance < 1,
            "Invalid solver tolerance or wave-speed safety");
    require(c.method >= DGH_METHOD_DG && c.method <= DGH_METHOD_BOTH, "Method must be dg, hdg, or both");
    require(c.precision >= DGH_PRECISION_FP64 && c.precision <= DGH_PRECISION_MIXED_FP16, "Invalid precision");
    require(c.inner_restart > 0 && c.inner_restart <= 31 && c.inner_max > 0, "Invalid iteration limit");
    require(std::isfinite(c.inner_tolerance) && c.inner_tolerance > 0 && c.inner_tolerance < 1,
            "Invalid solver tolerance or wave-speed safety");
    require(c.precond == 0 || c.precond == 1, "Invalid block preconditioner");
    require(c.ordering >= 0 && c.ordering <= 2, "Invalid ordering");
    if (c.precond == 1) {
        // The device block ILU(0) (block_ilu.cpp) is built for the DG element
        // graph in red-black order, where all its updates are diagonal.
        require(c.ordering == 2, "Block ILU(0) on the device runs in red-black order: use --ordering redblack");
        require(c.nx % 2 == 0 && c.ny % 2 == 0 && c.nx >= 4 && c.ny >= 4,
                "Block ILU(0) on the device runs in red-black order: nx and ny must be even and >= 4");
        require(c.method == DGH_METHOD_DG, "Block ILU(0) on the device is built for --method dg");
        require(c.precision != DGH_PRECISION_MIXED_BF16 && c.precision != DGH_PRECISION_MIXED_FP16,
                "Block ILU(0) on the device is built for --precision fp64 and mixed");
        require(!c.direct, "Block ILU(0) preconditions GMRES; --linear direct takes no preconditioner");
    } else {
        require(c.ordering != 2, "Red-black ordering applies to --precond block-ilu0");
    }
}

dgh_config parse_config(int argc, const char *const *argv) {
    dgh_config c = default_config();
    bool time_seen = false, steps_seen = false;
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
        else if (option == "--testcase") {
            auto name = value();
            if (name == "diagonal") c.testcase = DGH_DIAGONAL;
            else if (name == "wave") c.testcase = DGH_WAVE;
            else if (name == "axis") c.testcase = DGH_AXIS;
            else if (name == "seeded") c.testcase = DGH_SEEDED;
            else throw std::runtime_error("Invalid testcase");
        } else if (option == "--no-save-solution") c.save_solution = 0;
        else if (option == "--newton-steps") c.newton_steps = integer();
        else if (option == "--warmup-updates") c.warmup_updates = integer();
        else if (option == "--integrator") {
            auto name = value();
            require(name == "be" || name == "ssp-dirk3", "Invalid integrator");
            c.integrator = name == "be" ? 0 : 1;
        } else if (option == "--timing-only") c.timing_only = 1;
        else if (option == "--check-update") c.check_update = 1;
        else if (option == "--k") c.degree = integer();
        else if (option == "--quad") c.quadrature = integer();
        else if (option == "--cfl") c.cfl = real();
        else if (option == "--T") { c.final_time = real(); time_seen = true; }
        else if (option == "--steps") { c.steps = integer(); steps_seen = true; }
        else if (option == "--safety") c.safety = real();
        else if (option == "--method") {
            auto name = value();
            require(name == "dg" || name == "hdg" || name == "both", "Method must be dg, hdg, or both");
            c.method = name == "dg" ? DGH_METHOD_DG : name == "hdg" ? DGH_METHOD_HDG : DGH_METHOD_BOTH;
        } else if (option == "--prefix") {
            auto p = value();
            require(p.size() < sizeof(c.prefix), "Prefix too long");
            std::memset(c.prefix, 0, sizeof(c.prefix));
            std::strncpy(c.prefix, p.c_str(), sizeof(c.prefix) - 1);
        } else if (option == "--ntol") c.nonlinear_tolerance = real();
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
            c.ordering = order == "natural" ? 0 : order == "colamd" ? 1 : 2;
        } else if (option == "--precond") {
            auto name = value();
            require(name == "block-jacobi" || name == "block-ilu0", "Invalid block preconditioner");
            c.precond = name == "block-ilu0";
        } else if (option == "--device") {
            auto name = value();
            require(name == "gpu", "Only --device gpu is built; the CPU path is the reference executable");
        } else if (option == "--precision") {
            auto name = value();
            if (name == "fp64") c.precision = DGH_PRECISION_FP64;
            else if (name == "mixed") c.precision = DGH_PRECISION_MIXED;
            else if (name == "mixed-bf16") c.precision = DGH_PRECISION_MIXED_BF16;
            else if (name == "mixed-fp16") c.precision = DGH_PRECISION_MIXED_FP16;
            else throw std::runtime_error("Invalid precision");
        } else if (option == "--inner-restart") c.inner_restart = integer();
        else if (option == "--inner-max") c.inner_max = integer();
        else if (option == "--inner-tol") c.inner_tolerance = real();
        else if (option == "--graph") c.graph = 1;
        else if (option == "--no-graph") c.graph = 0;
        else if (option == "--seed") {
            auto text = value();
            size_t end;
            unsigned long long s = std::stoull(text, &end);
            require(end == text.size(), "Invalid integer");
            c.seed = s;
        } else if (option == "--help") {
            std::cout << "2D periodic tensor-product Euler DG / condensed HDG (MI355X device path)\n"
                         "--nx NX --ny NY --testcase diagonal|wave|axis|seeded --k K --quad Q --cfl C (--T T | --steps N)\n"
                         "--method dg|hdg|both --linear direct|gmres --precond block-jacobi|block-ilu0\n"
                         "--ordering natural|colamd|redblack --restart N --max-linear N --max-newton N\n"
                         "--ntol R --ltol R --safety S --prefix PATH\n"
                         "--integrator ssp-dirk3|be --newton-steps N --warmup-updates N\n"
                         "--timing-only --check-update --no-save-solution\n"
                         "--device gpu --precision fp64|mixed|mixed-bf16|mixed-fp16 --inner-restart N --inner-max N\n"
                         "--inner-tol R --seed S --graph | --no-graph\n";
            throw std::runtime_error("help");
        } else
            throw std::runtime_error("Unknown option: " + option);
    }
    require(!(time_seen && steps_seen), "Choose either --T or --steps");
    require(!steps_seen || c.steps > 0, "steps must be positive");
    validate_config(c);
    return c;
}
}  // namespace dgh
