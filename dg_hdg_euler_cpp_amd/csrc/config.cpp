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

namespace {
// The flag table: each row names a flag, how its argument is read and what it
// sets. The messages are the reference's (config.cpp:26-45, 100-122), since
// its harness and ours match on them.
enum class Arg { None, Int, Real, Word, Count };
struct Value {
    int i = 0;
    double d = 0;
    unsigned long long u = 0;
    std::string word;
};
struct Flag {
    const char *name;
    Arg arg;
    void (*set)(dgh_config &, const Value &);
};
struct Name {
    const char *word;
    int code;
};
template <size_t K>
int named(const std::string &word, const Name (&names)[K], const char *message) {
    for (const Name &n : names)
        if (word == n.word) return n.code;
    throw std::runtime_error(message);
}
const Name kTestcases[] = {{"diagonal", DGH_DIAGONAL}, {"wave", DGH_WAVE}, {"axis", DGH_AXIS}, {"seeded", DGH_SEEDED}};
const Name kIntegrators[] = {{"be", 0}, {"ssp-dirk3", 1}};
const Name kMethods[] = {{"dg", DGH_METHOD_DG}, {"hdg", DGH_METHOD_HDG}, {"both", DGH_METHOD_BOTH}};
const Name kLinear[] = {{"gmres", 0}, {"direct", 1}};
const Name kOrderings[] = {{"natural", 0}, {"colamd", 1}, {"redblack", 2}};
const Name kPreconds[] = {{"block-jacobi", 0}, {"block-ilu0", 1}};
const Name kPrecisions[] = {{"fp64", DGH_PRECISION_FP64}, {"mixed", DGH_PRECISION_MIXED},
                            {"mixed-bf16", DGH_PRECISION_MIXED_BF16}, {"mixed-fp16", DGH_PRECISION_MIXED_FP16}};

const char kUsage[] =
    "2D periodic tensor-product Euler DG / condensed HDG (MI355X device path)\n"
    "--nx NX --ny NY --testcase diagonal|wave|axis|seeded --k K --quad Q --cfl C (--T T | --steps N)\n"
    "--method dg|hdg|both --linear direct|gmres --precond block-jacobi|block-ilu0\n"
    "--ordering natural|colamd|redblack --restart N --max-linear N --max-newton N\n"
    "--ntol R --ltol R --safety S --prefix PATH\n"
    "--integrator ssp-dirk3|be --newton-steps N --warmup-updates N\n"
    "--timing-only --check-update --no-save-solution\n"
    "--device gpu --precision fp64|mixed|mixed-bf16|mixed-fp16 --inner-restart N --inner-max N\n"
    "--inner-tol R --seed S --graph | --no-graph\n";

// --T and --steps are exclusive; the parser records which one it saw here.
constexpr int kSawTime = 1, kSawSteps = 2;
thread_local int g_seen = 0;

const Flag kFlags[] = {
    {"--nx", Arg::Int, [](dgh_config &c, const Value &v) { c.nx = v.i; }},
    {"--ny", Arg::Int, [](dgh_config &c, const Value &v) { c.ny = v.i; }},
    {"--k", Arg::Int, [](dgh_config &c, const Value &v) { c.degree = v.i; }},
    {"--quad", Arg::Int, [](dgh_config &c, const Value &v) { c.quadrature = v.i; }},
    {"--testcase", Arg::Word,
     [](dgh_config &c, const Value &v) { c.testcase = named(v.word, kTestcases, "Invalid testcase"); }},
    {"--integrator", Arg::Word,
     [](dgh_config &c, const Value &v) { c.integrator = named(v.word, kIntegrators, "Invalid integrator"); }},
    {"--method", Arg::Word,
     [](dgh_config &c, const Value &v) { c.method = named(v.word, kMethods, "Method must be dg, hdg, or both"); }},
    {"--linear", Arg::Word,
     [](dgh_config &c, const Value &v) { c.direct = named(v.word, kLinear, "Invalid linear solver"); }},
    {"--ordering", Arg::Word,
     [](dgh_config &c, const Value &v) {
         c.ordering = named(v.word, kOrderings, "Invalid ordering");
         c.colamd = c.ordering == 1;
     }},
    {"--precond", Arg::Word,
     [](dgh_config &c, const Value &v) { c.precond = named(v.word, kPreconds, "Invalid block preconditioner"); }},
    {"--precision", Arg::Word,
     [](dgh_config &c, const Value &v) { c.precision = named(v.word, kPrecisions, "Invalid precision"); }},
    {"--device", Arg::Word,
     [](dgh_config &, const Value &v) {
         require(v.word == "gpu", "Only --device gpu is built; the CPU path is the reference executable");
     }},
    {"--prefix", Arg::Word,
     [](dgh_config &c, const Value &v) {
         require(v.word.size() < sizeof(c.prefix), "Prefix too long");
         std::memset(c.prefix, 0, sizeof(c.prefix));
         std::strncpy(c.prefix, v.word.c_str(), sizeof(c.prefix) - 1);
     }},
    {"--cfl", Arg::Real, [](dgh_config &c, const Value &v) { c.cfl = v.d; }},
    {"--T", Arg::Real,
     [](dgh_config &c, const Value &v) {
         c.final_time = v.d;
         g_seen |= kSawTime;
     }},
    {"--steps", Arg::Int,
     [](dgh_config &c, const Value &v) {
         c.steps = v.i;
         g_seen |= kSawSteps;
     }},
    {"--safety", Arg::Real, [](dgh_config &c, const Value &v) { c.safety = v.d; }},
    {"--ntol", Arg::Real, [](dgh_config &c, const Value &v) { c.nonlinear_tolerance = v.d; }},
    {"--ltol", Arg::Real, [](dgh_config &c, const Value &v) { c.linear_tolerance = v.d; }},
    {"--restart", Arg::Int, [](dgh_config &c, const Value &v) { c.restart = v.i; }},
    {"--max-linear", Arg::Int, [](dgh_config &c, const Value &v) { c.max_linear = v.i; }},
    {"--max-newton", Arg::Int, [](dgh_config &c, const Value &v) { c.max_newton = v.i; }},
    {"--newton-steps", Arg::Int, [](dgh_config &c, const Value &v) { c.newton_steps = v.i; }},
    {"--warmup-updates", Arg::Int, [](dgh_config &c, const Value &v) { c.warmup_updates = v.i; }},
    {"--inner-restart", Arg::Int, [](dgh_config &c, const Value &v) { c.inner_restart = v.i; }},
    {"--inner-max", Arg::Int, [](dgh_config &c, const Value &v) { c.inner_max = v.i; }},
    {"--inner-tol", Arg::Real, [](dgh_config &c, const Value &v) { c.inner_tolerance = v.d; }},
    {"--seed", Arg::Count, [](dgh_config &c, const Value &v) { c.seed = v.u; }},
    {"--timing-only", Arg::None, [](dgh_config &c, const Value &) { c.timing_only = 1; }},
    {"--check-update", Arg::None, [](dgh_config &c, const Value &) { c.check_update = 1; }},
    {"--no-save-solution", Arg::None, [](dgh_config &c, const Value &) { c.save_solution = 0; }},
    {"--graph", Arg::None, [](dgh_config &c, const Value &) { c.graph = 1; }},
    {"--no-graph", Arg::None, [](dgh_config &c, const Value &) { c.graph = 0; }},
    {"--help", Arg::None,
     [](dgh_config &, const Value &) {
         std::cout << kUsage;
         throw std::runtime_error("help");
     }},
};

// The argument text of a flag, converted whole or rejected.
Value read_argument(Arg arg, const char *text) {
    Value v;
    size_t end = 0;
    const std::string word = text;
    switch (arg) {
    case Arg::Int:
        v.i = std::stoi(word, &end);
        require(end == word.size(), "Invalid integer");
        break;
    case Arg::Count:
        v.u = std::stoull(word, &end);
        require(end == word.size(), "Invalid integer");
        break;
    case Arg::Real:
        v.d = std::stod(word, &end);
        require(end == word.size(), "Invalid number");
        break;
    default:
        v.word = word;
        break;
    }
    return v;
}
}  // namespace

dgh_config parse_config(int argc, const char *const *argv) {
    dgh_config c = default_config();
    g_seen = 0;
    int i = 1;
    while (i < argc) {
        const std::string option = argv[i++];
        const Flag *flag = nullptr;
        for (const Flag &f : kFlags)
            if (option == f.name) flag = &f;
        if (!flag) throw std::runtime_error("Unknown option: " + option);
        Value v;
        if (flag->arg != Arg::None) {
            require(i < argc, "Missing option value");
            v = read_argument(flag->arg, argv[i++]);
        }
        flag->set(c, v);
    }
    require(g_seen != (kSawTime | kSawSteps), "Choose either --T or --steps");
    require(!(g_seen & kSawSteps) || c.steps > 0, "steps must be positive");
    validate_config(c);
    return c;
}
}  // namespace dgh
