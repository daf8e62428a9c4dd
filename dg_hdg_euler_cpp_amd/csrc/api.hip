// Host side of libdgh_device.so: context set-up, the device-resident Newton /
// GMRES drivers and the extern "C" entry points of include/dgh/dgh_device.h.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dgh/dgh_device.h"
#include "dgh_internal.hpp"
#include "dgh_host.hpp"

using namespace dgh;

// Gram-Schmidt passes per Arnoldi step of the inner FP32 GMRES of the mixed
// solve. One pass: the inner solve only has to reach 1e-3 and the FP64 outer
// FGMRES keeps the acceptance test, and at 256^2, k=8 (FP16 factors) the
// inner iteration counts and residuals match the two-pass form while an update
// takes 1.26 s instead of 1.34 s. The FP64 solver keeps two passes.
#ifndef DGH_INNER_GS_PASSES
#define DGH_INNER_GS_PASSES 1
#endif

namespace {
thread_local std::string g_create_error;

void require(bool condition, const char *message) {
    if (!condition) throw std::runtime_error(message);
}
// An environment switch that selects an alternative kernel for measurement.
bool env_flag(const char *name) {
    const char *e = std::getenv(name);
    return e && e[0] == '1';
}
void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}
#define HIPCHECK(call) hip_check((call), #call)

using Clock = std::chrono::steady_clock;
double elapsed(Clock::time_point start) { return std::chrono::duration<double>(Clock::now() - start).count(); }

// ------------------------------------------------------------- device memory
// Bumped whenever a device buffer is (re)allocated or freed: captured graphs
// record buffer addresses and stay valid while it is unchanged.
uint64_t &alloc_epoch() {
    static uint64_t epoch = 0;
    return epoch;
}
template <class T> struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    void alloc(size_t count, size_t &counter) {
        if (n >= count && p) return;
        release(counter);
        HIPCHECK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
        n = count;
        counter += count * sizeof(T);
        ++alloc_epoch();
    }
    void release(size_t &counter) {
        if (p) {
            (void)hipFree(p);
            counter -= n * sizeof(T);
            ++alloc_epoch();
        }
        p = nullptr;
        n = 0;
    }
};

// One executable graph per Arnoldi step index of a Krylov solver, kept while
// the key (solver, operators, allocation epoch, step size, restart) is
// unchanged, i.e. across the solves of one stage, and captured again after.
struct StepGraphs {
    std::vector<hipGraphExec_t> execs;
    std::vector<uint64_t> key;
    void drop() {
        for (auto &e : execs)
            if (e) (void)hipGraphExecDestroy(e);
        execs.clear();
        key.clear();
    }
    void prepare(const std::vector<uint64_t> &k, int count) {
        if (k == key && static_cast<int>(execs.size()) == count) return;
        drop();
        key = k;
        execs.assign(count, nullptr);
    }
};
uint64_t bits_of(double v) {
    uint64_t b;
    std::memcpy(&b, &v, sizeof b);
    return b;
}

struct Timer {
    double ms = 0;
    long long launches = 0;
};

// ------------------------------------------------ fixed scratch (no allocation)
// Device memory of dgh_hdg_apply_noalloc: the skeleton-diagonal product D x
// (2 cells 4n doubles, padded to 256 bytes) followed by the per-element Schur
// products S_K gather(x) (cells 16n doubles). Every GPU the library runs on
// holds its own copy of this array until the process ends. In host code the
// name `scratch` is not a device address: its per-device address comes from
// hipGetSymbolAddress (scratch_address) and only that address is passed on.
constexpr size_t SCRATCH_DOUBLES = size_t(8) << 20;  // 64 MiB
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// Address of `scratch` on the current device, looked up once per device.
double *scratch_address() {
    int device = 0;
    HIPCHECK(hipGetDevice(&device));
    require(device >= 0 && device < SCRATCH_MAX_DEVICES, "Device ordinal beyond the fixed scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        require(p != nullptr, "hipGetSymbolAddress returned no address for the fixed scratch");
        g_scratch_ptr[device] = static_cast<double *>(p);
    }
    return g_scratch_ptr[device];
}

}  // namespace

// ----------------------------------------------------------------- context
struct dgh_ctx {
    dgh_config cfg{};
    Geo geo{};
    Basis hbasis{};
    int n = 0, P = 0, m = 0, cells = 0, faces = 0, fw = 0;
    long long D = 0, T = 0;
    double tau = 0, inv_step = 0;
    bool refreshed = false, frozen = false, bj_ready = false;
    hipStream_t stream = nullptr;
    std::string error;
    size_t bytes = 0;
    bool timing = false;
    std::map<std::string, Timer> timers;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    DevBuf<Basis> basis;
    DevBuf<double> old, speeds, tr_u, tr_x, fq, gtr, vcache, fcache, partials, scalars;
    DevBuf<int> flags;
    DevBuf<unsigned long long> rate_bits;
    DevBuf<double> factors;
    DevBuf<float> factors32;
    DevBuf<int> rowsrc;
    // GMRES
    DevBuf<double> kbasis, kw, kz, kwork, kres, ky;
    // mixed precision: FP32 / BF16 factors, FP32 inner Krylov space, FGMRES Z
    DevBuf<float> factors_f32, ibasis, iw, it, is, ires, ib, zbasis;
    DevBuf<unsigned short> factors_bf16;  // 2-byte factors (BF16 or FP16) in the pair layout
    DevBuf<float> factor_scale;           // FP16: 2^-e per block (U' = sigma U)
    DevBuf<float> factor_rdiag;           // 2-byte factors: FP32 1 / (sigma U_ii) per row
    DevBuf<float> factor_umax;            // FP16: max |U| per block, recorded by the LU
    DevBuf<int> rowsrc32;
    DevBuf<double> check_dir;
    bool bj32_ready = false;
    // HDG: X [cells][m x t], y [cells][m], S_K [cells][t x t], per-element face
    // results, trace work vectors, face block-Jacobi factors
    DevBuf<double> hX, hY, hS, hfres, hloc, hdiag, hrhs, hdinv, itm, face_factors, trace_dir, hlambda, indep;
    DevBuf<int> face_rowsrc;
    bool hdg_ready = false, face_bj_ready = false;
    // Newton
    DevBuf<double> state, trial, residual, trial_residual, direction, rhs, equation_error, step_state, stage_base,
        slopes;
    // Discrete adjoint: the recorded trajectory (converged stage states and the
    // wave speeds each stage froze), the step sizes, and the reverse-sweep work
    enum Trajectory { PLAIN, RECORD, REPLAY };
    Trajectory trajectory = PLAIN;
    DevBuf<double> ck_states, ck_speeds, adj_kbar, adj_lambda, adj_rhs, adj_base;
    std::vector<double> ck_step, ck_rate;
    // Graph-captured GMRES (cfg.graph): one executable graph per Arnoldi step
    // index, captured within a linear solve on first use; the Hessenberg
    // matrix, rotations and rotated right-hand side live on the device and
    // three scalars per step come back through pinned, mapped host memory.
    StepGraphs step_graphs, inner_graphs;
    // Block ILU(0) in red-black order (cfg.precond == 1): compact red factors
    // in factors[0, cells/2), black ones in factors[cells/2, cells); the red
    // couplings S [red cell][side][4n x 4n]; apply work vectors
    DevBuf<double> ilu_Y, ilu_S, ilu_z, ilu_w;
    bool ilu_ready = false;
    // staging for the in-place packing of LU blocks (launch_pack)
    DevBuf<double> pack_scratch;
    // mixed precision: FP32 factors in factors_f32 (same split), FP32 Y and work
    DevBuf<float> ilu_Y32, ilu_z32, ilu_w32;
    bool ilu32_ready = false;
    DevBuf<double> gH, gcs, gsn, gg, gnext, igH, igcs, igsn, igg, ignext;
    double *gout_host = nullptr, *gout_dev = nullptr;  // [0, 4): FP64 solver, [4, 8): inner FP32 solver
    int ck_stages = 0;
    bool adj_bj_ready = false;

    explicit dgh_ctx(const dgh_config &c) : cfg(c) {
        n = c.degree + 1;
        P = n * n;
        m = 4 * P;
        cells = c.nx * c.ny;
        faces = 2 * cells;
        fw = 4 * n;
        D = static_cast<long long>(cells) * m;
        T = static_cast<long long>(faces) * fw;
        const double pi = std::acos(-1.0);
        const double cosine = std::cos(45. * pi / 180.), sine = std::sin(45. * pi / 180.);
        geo.nx = c.nx;
        geo.ny = c.ny;
        geo.n = n;
        geo.cells = cells;
        geo.eq = std::max(20, 2 * n + 3);
        geo.nq = c.quadrature > n ? c.quadrature : n;
        geo.testcase = c.testcase;
        geo.lx = c.testcase == DGH_DIAGONAL ? 2. / cosine : 2.;
        geo.ly = c.testcase == DGH_DIAGONAL ? 2. / sine : 2.;
        geo.hx = geo.lx / c.nx;
        geo.hy = geo.ly / c.ny;
        geo.gamma = 3.0;
        geo.safety = c.safety;
        geo.cosine = cosine;
        geo.sine = sine;
        hbasis = make_basis(n, geo.hx, geo.hy, geo.nq);
        HIPCHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HIPCHECK(hipEventCreate(&ev0));
        HIPCHECK(hipEventCreate(&ev1));
        basis.alloc(1, bytes);
        HIPCHECK(hipMemcpy(basis.p, &hbasis, sizeof(Basis), hipMemcpyHostToDevice));
        old.alloc(D, bytes);
        speeds.alloc(2LL * faces * n, bytes);
        tr_u.alloc(16LL * cells * n, bytes);
        tr_x.alloc(16LL * cells * n, bytes);
        fq.alloc(static_cast<size_t>(T), bytes);
        gtr.alloc(2 * static_cast<size_t>(T), bytes);
        vcache.alloc(3LL * cells * geo.nq * geo.nq, bytes);
        fcache.alloc(static_cast<size_t>(F_NFIELDS) * faces * n, bytes);
        partials.alloc(REDUCE_BLOCKS * 32, bytes);
        scalars.alloc(256, bytes);
        flags.alloc(8, bytes);
        rate_bits.alloc(1, bytes);
        HIPCHECK(hipMemsetAsync(old.p, 0, D * sizeof(double), stream));
        HIPCHECK(hipStreamSynchronize(stream));
    }
    ~dgh_ctx() {
        DevBuf<double> *bufs[] = {&old, &speeds, &tr_u, &tr_x, &fq, &gtr, &vcache, &fcache, &partials, &scalars,
                                  &factors, &kbasis, &kw, &kz, &kwork, &kres, &ky, &state, &trial, &residual,
                                  &trial_residual, &direction, &rhs, &equation_error, &step_state, &stage_base,
                                  &slopes, &check_dir, &hX, &hY, &hS, &hfres, &hloc, &hdiag, &hrhs, &hdinv, &itm,
                                  &face_factors, &trace_dir, &hlambda, &indep, &ck_states, &ck_speeds,
                                  &adj_kbar, &adj_lambda, &adj_rhs, &adj_base};
        for (auto *b : bufs) b->release(bytes);
        DevBuf<float> *fbufs[] = {&factors_f32, &ibasis, &iw, &it, &is, &ires, &ib, &zbasis};
        for (auto *b : fbufs) b->release(bytes);
        factors_bf16.release(bytes);
        factor_scale.release(bytes);
        factor_rdiag.release(bytes);
        factor_umax.release(bytes);
        rowsrc32.release(bytes);
        face_rowsrc.release(bytes);
        factors32.release(bytes);
        rowsrc.release(bytes);
        flags.release(bytes);
        rate_bits.release(bytes);
        basis.release(bytes);
        drop_step_graphs();
        if (gout_host) (void)hipHostFree(gout_host);
        DevBuf<double> *gbufs[] = {&gH,  &gcs,  &gsn,   &gg,    &gnext, &igH,  &igcs,        &igsn,
                                   &igg, &ignext, &ilu_Y, &ilu_S, &ilu_z, &ilu_w, &pack_scratch};
        for (auto *b : gbufs) b->release(bytes);
        DevBuf<float> *ibufs[] = {&ilu_Y32, &ilu_z32, &ilu_w32};
        for (auto *b : ibufs) b->release(bytes);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void drop_step_graphs() {
        step_graphs.drop();
        inner_graphs.drop();
    }
    // Pinned, mapped host memory the captured Givens steps report to.
    void graph_out() {
        if (gout_host) return;
        HIPCHECK(hipHostMalloc(reinterpret_cast<void **>(&gout_host), 8 * sizeof(double), hipHostMallocMapped));
        HIPCHECK(hipHostGetDevicePointer(reinterpret_cast<void **>(&gout_dev), gout_host, 0));
    }
    // Capture fn() on the context stream as one executable graph in *exec.
    template <class F> void capture(hipGraphExec_t *exec, F &&fn) {
        hipGraph_t graph = nullptr;
        HIPCHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        try {
            fn();
        } catch (...) {
            // leave the stream out of capture mode before reporting
            if (hipStreamEndCapture(stream, &graph) == hipSuccess && graph) (void)hipGraphDestroy(graph);
            throw;
        }
        HIPCHECK(hipStreamEndCapture(stream, &graph));
        const hipError_t e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPCHECK(e);
    }

    // ---- timing of kernel groups on the context stream
    template <class F> void timed(const char *name, F &&launch) {
        if (!timing) {
            launch();
            return;
        }
        HIPCHECK(hipEventRecord(ev0, stream));
        launch();
        HIPCHECK(hipEventRecord(ev1, stream));
        HIPCHECK(hipEventSynchronize(ev1));
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, ev0, ev1));
        auto &t = timers[name];
        t.ms += ms;
        t.launches += 1;
    }
    void sync() { HIPCHECK(hipStreamSynchronize(stream)); }
    // LU blocks of width w into the packed layout the block solve reads
    template <typename T> void pack_blocks(int w, T *blocks, int count) {
        const int chunk = std::min(count, 2048);
        pack_scratch.alloc(static_cast<size_t>(chunk) * w * w, bytes);
        timed("pack", [&] {
            HIPCHECK(launch_pack(w, blocks, count, reinterpret_cast<T *>(pack_scratch.p), chunk, stream));
        });
    }
    bool oi() const { return geo.nq > n; }  // over-integrated volume rule (--quad Q > k+1)

    // ---- reductions
    double reduce_sq(const double *x, long long count, int mode, const double *weights = nullptr) {
        HIPCHECK(launch_weighted_sq(n, x, count, mode, weights, partials.p, basis.p, stream));
        HIPCHECK(launch_finish_sum(partials.p, REDUCE_BLOCKS, 1, scalars.p + 200, stream));
        double out = 0;
        HIPCHECK(hipMemcpyAsync(&out, scalars.p + 200, sizeof(double), hipMemcpyDeviceToHost, stream));
        sync();
        return out;
    }
    double dual_norm(const double *x) { return std::sqrt(reduce_sq(x, D, 0)); }
    double state_norm(const double *x) { return std::sqrt(reduce_sq(x, D, 1)); }
    double gmres_norm(const double *x) { return std::sqrt(std::max(0.0, reduce_sq(x, D, 0))); }
    std::array<double, 4> integral(const double *u) {
        HIPCHECK(launch_integral(n, u, cells, partials.p, basis.p, stream));
        HIPCHECK(launch_finish_sum(partials.p, REDUCE_BLOCKS, 4, scalars.p + 200, stream));
        std::array<double, 4> out{};
        HIPCHECK(hipMemcpyAsync(out.data(), scalars.p + 200, 4 * sizeof(double), hipMemcpyDeviceToHost, stream));
        sync();
        return out;
    }
    int read_flag(int idx) {
        int f = 0;
        HIPCHECK(hipMemcpyAsync(&f, flags.p + idx, sizeof(int), hipMemcpyDeviceToHost, stream));
        sync();
        return f;
    }
    void set_flag(int idx, int value) { HIPCHECK(hipMemsetD32Async(flags.p + idx, value, 1, stream)); }
    void copy(double *dst, const double *src, long long count) {
        HIPCHECK(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToDevice, stream));
    }
    void axpby(double a, const double *x, double b, const double *y, double *out, long long count) {
        HIPCHECK(launch_axpby(a, x, b, y, out, count, stream));
    }

    // ---- Problem
    void initial(double *u) {
        set_flag(3, 0);
        if (cfg.testcase == DGH_SEEDED) HIPCHECK(launch_seeded(n, u, cfg.seed, flags.p + 3, basis.p, geo, stream));
        else HIPCHECK(launch_initial(n, u, flags.p + 3, basis.p, geo, stream));
        require(read_flag(3) == 0, "Nonphysical cell average");
        require(admissible(u), "Invalid projected initial state");
    }
    void seeded(double *u, unsigned long long seed) {
        set_flag(3, 0);
        HIPCHECK(launch_seeded(n, u, seed, flags.p + 3, basis.p, geo, stream));
        require(read_flag(3) == 0, "Nonphysical cell average");
        require(admissible(u), "Invalid projected initial state");
    }
    double freeze(const double *u) {
        set_flag(0, 0);
        HIPCHECK(hipMemsetAsync(rate_bits.p, 0, sizeof(unsigned long long), stream));
        HIPCHECK(launch_restrict(n, u, tr_u.p, basis.p, geo, stream));
        HIPCHECK(launch_freeze_speeds(n, tr_u.p, speeds.p, flags.p, geo, stream));
        if (oi()) HIPCHECK(launch_oi_freeze_rate(n, u, speeds.p, rate_bits.p, flags.p, basis.p, geo, stream));
        else HIPCHECK(launch_freeze_rate(n, u, speeds.p, rate_bits.p, flags.p, geo, stream));
        copy(old.p, u, D);
        unsigned long long bits = 0;
        HIPCHECK(hipMemcpyAsync(&bits, rate_bits.p, sizeof(bits), hipMemcpyDeviceToHost, stream));
        const int f = read_flag(0);
        require(!(f & 1), "Invalid frozen face state");
        require(!(f & 2), "Invalid CFL volume state");
        frozen = true;
        refreshed = false;
        double rate;
        std::memcpy(&rate, &bits, sizeof(rate));
        return rate;
    }
    void residual_into(const double *u, double *r) {
        require(frozen, "Freeze the wave speeds before evaluating the residual");
        timed("residual", [&] {
            HIPCHECK(launch_restrict(n, u, tr_u.p, basis.p, geo, stream));
            HIPCHECK(launch_residual_flux(n, tr_u.p, speeds.p, fq.p, basis.p, geo, stream));
            if (oi()) HIPCHECK(launch_oi_element_residual(n, u, old.p, fq.p, r, tau, basis.p, geo, stream));
            else HIPCHECK(launch_element_residual(n, u, old.p, fq.p, r, tau, basis.p, geo, stream));
        });
    }
    bool admissible(const double *u) {
        set_flag(1, 1);
        HIPCHECK(launch_admissible(n, u, flags.p + 1, basis.p, geo, stream));
        if (oi()) HIPCHECK(launch_oi_admissible(n, u, flags.p + 1, basis.p, geo, stream));
        return read_flag(1) != 0;
    }
    std::array<double, 4> error_l2(const double *u, double time) {
        HIPCHECK(launch_error_partials(n, u, time, partials.p, REDUCE_BLOCKS, basis.p, geo, stream));
        HIPCHECK(launch_finish_sum(partials.p, REDUCE_BLOCKS, 4, scalars.p + 200, stream));
        std::array<double, 4> out{};
        HIPCHECK(hipMemcpyAsync(out.data(), scalars.p + 200, 4 * sizeof(double), hipMemcpyDeviceToHost, stream));
        sync();
        for (auto &v : out) v = std::sqrt(v);
        return out;
    }

    // ---- DG linearisation
    void dg_refresh(const double *u) {
        require(tau > 0, "Set the time step before refreshing the Jacobian");
        require(frozen, "Freeze the wave speeds before refreshing the Jacobian");
        inv_step = 1 / tau;
        if (oi()) HIPCHECK(launch_oi_volume_cache(n, u, vcache.p, basis.p, geo, stream));
        else HIPCHECK(launch_volume_cache(n, u, vcache.p, geo, stream));
        HIPCHECK(launch_restrict(n, u, tr_u.p, basis.p, geo, stream));
        HIPCHECK(launch_face_cache(n, tr_u.p, speeds.p, fcache.p, basis.p, geo, stream));
        refreshed = true;
        bj_ready = false;
        ilu_ready = false;
        ilu32_ready = false;
        adj_bj_ready = false;
        bj32_ready = false;
        hdg_ready = false;
        face_bj_ready = false;
    }
    void dg_apply(const double *x, double *y) {
        require(refreshed, "Refresh the volume Jacobian before applying it");
        timed("jv", [&] {
            HIPCHECK(launch_restrict(n, x, tr_x.p, basis.p, geo, stream));
            HIPCHECK(launch_jv_flux(n, tr_x.p, fcache.p, fq.p, geo, stream));
            if (oi()) HIPCHECK(launch_oi_element_jv(n, x, vcache.p, fq.p, y, inv_step, basis.p, geo, stream));
            else HIPCHECK(launch_element_jv(n, x, vcache.p, fq.p, y, inv_step, basis.p, geo, stream));
        });
    }
    void dg_apply_transpose(const double *w, double *z) {
        require(refreshed, "Refresh the volume Jacobian before applying it");
        timed("jtv", [&] {
            HIPCHECK(launch_restrict(n, w, tr_x.p, basis.p, geo, stream));
            HIPCHECK(launch_adjoint_face(n, tr_x.p, fcache.p, gtr.p, geo, stream));
            if (oi()) HIPCHECK(launch_oi_element_jtv(n, w, vcache.p, gtr.p, z, inv_step, basis.p, geo, stream));
            else HIPCHECK(launch_element_jtv(n, w, vcache.p, gtr.p, z, inv_step, basis.p, geo, stream));
        });
    }
    void bj_refresh() {
        require(refreshed, "Refresh the volume Jacobian before assembling a block");
        factors.alloc(static_cast<size_t>(cells) * m * m, bytes);
        rowsrc.alloc(static_cast<size_t>(cells) * m, bytes);
        set_flag(2, 0);
        timed("assemble", [&] {
            HIPCHECK(launch_assemble_blocks(n, vcache.p, fcache.p, factors.p, inv_step, basis.p, geo, 0, cells, stream));
        });
        timed("lu", [&] { HIPCHECK(launch_lu(m, factors.p, rowsrc.p, cells, flags.p + 2, stream)); });
        require(read_flag(2) == 0, "Singular block Jacobi factor");
        pack_blocks(m, factors.p, cells);
        bj_ready = true;
        ilu_ready = false;
    }
    // Block ILU(0) of the element graph in red-black order (block_ilu.cpp:48-61
    // with red cells first): red LU, red couplings, black diagonal update,
    // black LU. Shares the factor storage with bj_refresh.
    void ilu_refresh() {
        require(refreshed, "Refresh the volume Jacobian before assembling a block");
        require(cfg.nx % 2 == 0 && cfg.ny % 2 == 0 && cfg.nx >= 4 && cfg.ny >= 4,
                "Block ILU(0) on the device runs in red-black order: nx and ny must be even and >= 4");
        const int half = cells / 2;
        const size_t mm = static_cast<size_t>(m) * m;
        factors.alloc(static_cast<size_t>(cells) * mm, bytes);
        rowsrc.alloc(static_cast<size_t>(cells) * m, bytes);
        ilu_S.alloc(static_cast<size_t>(half) * 4 * 16 * n * n, bytes);
        ilu_Y.alloc(static_cast<size_t>(half) * 16 * n * m, bytes);
        if (!ilu_z.p) {
            ilu_z.alloc(D, bytes);
            ilu_w.alloc(D, bytes);
            HIPCHECK(hipMemsetAsync(ilu_z.p, 0, D * sizeof(double), stream));
        }
        bj_ready = false;
        adj_bj_ready = false;
        ilu_ready = false;
        set_flag(2, 0);
        timed("assemble", [&] {
            HIPCHECK(launch_assemble_colour(n, vcache.p, fcache.p, factors.p, inv_step, basis.p, geo, 0, stream));
            HIPCHECK(launch_assemble_colour(n, vcache.p, fcache.p, factors.p + half * mm, inv_step, basis.p, geo, 1,
                                            stream));
        });
        timed("lu", [&] { HIPCHECK(launch_lu(m, factors.p, rowsrc.p, half, flags.p + 2, stream)); });
        timed("ilu_coupling", [&] { HIPCHECK(launch_ilu_coupling(n, factors.p, rowsrc.p, ilu_Y.p, ilu_S.p, basis.p, geo, stream)); });
        pack_blocks(m, factors.p, half);
        timed("ilu_update", [&] {
            HIPCHECK(launch_ilu_update(n, factors.p + half * mm, ilu_S.p, fcache.p, basis.p, geo, stream));
        });
        timed("lu", [&] {
            HIPCHECK(launch_lu(m, factors.p + half * mm, rowsrc.p + static_cast<size_t>(half) * m, half, flags.p + 2,
                               stream));
        });
        require(read_flag(2) == 0, "Singular or nonfinite LU factors");
        pack_blocks(m, factors.p + half * mm, half);
        ilu_ready = true;
    }
    // block_ilu.cpp:62-84 in red-black order: z0_r = A_rr^-1 r_r;
    // w_b = r_b - sum A_br z0_r; z_b = A_bb'^-1 w_b;
    // z_r = A_rr^-1 (r_r - sum A_rb z_b) = z0_r - sum Y_r J R_b z_b
    void ilu_apply(const double *r, double *z) {
        require(ilu_ready, "Refresh the block ILU(0) preconditioner first");
        require(r != z, "The block ILU(0) apply needs distinct input and output vectors");
        const int half = cells / 2;
        const size_t mm = static_cast<size_t>(m) * m;
        timed("ilu_apply", [&] {
            HIPCHECK(launch_bj_apply_colour(m, factors.p, rowsrc.p, r, ilu_z.p, half, 0, cfg.nx, stream));
            HIPCHECK(launch_restrict(n, ilu_z.p, tr_x.p, basis.p, geo, stream));
            HIPCHECK(launch_ilu_offdiag(n, 1, tr_x.p, fcache.p, r, ilu_w.p, basis.p, geo, stream));
            HIPCHECK(launch_bj_apply_colour(m, factors.p + half * mm, rowsrc.p + static_cast<size_t>(half) * m,
                                            ilu_w.p, z, half, 1, cfg.nx, stream));
            HIPCHECK(launch_restrict(n, z, tr_x.p, basis.p, geo, stream));
            HIPCHECK(launch_ilu_red_back(n, ilu_Y.p, tr_x.p, fcache.p, ilu_z.p, z, geo, stream));
        });
    }
    // The same preconditioner with FP32 factors for the inner FP32 GMRES of the
    // mixed-precision solve: blocks assembled from the FP64 linearisation into
    // FP32, FP32 LU and block solves, couplings accumulated in FP64.
    void ilu_refresh_mixed() {
        require(refreshed, "Refresh the volume Jacobian before assembling a block");
        require(cfg.nx % 2 == 0 && cfg.ny % 2 == 0 && cfg.nx >= 4 && cfg.ny >= 4,
                "Block ILU(0) on the device runs in red-black order: nx and ny must be even and >= 4");
        const int half = cells / 2;
        const size_t mm = static_cast<size_t>(m) * m;
        factors_f32.alloc(static_cast<size_t>(cells) * mm, bytes);
        rowsrc32.alloc(static_cast<size_t>(cells) * m, bytes);
        ilu_S.alloc(static_cast<size_t>(half) * 4 * 16 * n * n, bytes);
        ilu_Y32.alloc(static_cast<size_t>(half) * 16 * n * m, bytes);
        if (!ilu_z32.p) {
            ilu_z32.alloc(D, bytes);
            ilu_w32.alloc(D, bytes);
            HIPCHECK(hipMemsetAsync(ilu_z32.p, 0, D * sizeof(float), stream));
        }
        bj32_ready = false;
        ilu32_ready = false;
        set_flag(2, 0);
        timed("assemble", [&] {
            HIPCHECK(launch_assemble_colour(n, vcache.p, fcache.p, factors_f32.p, inv_step, basis.p, geo, 0, stream));
            HIPCHECK(launch_assemble_colour(n, vcache.p, fcache.p, factors_f32.p + half * mm, inv_step, basis.p, geo,
                                            1, stream));
        });
        timed("lu", [&] { HIPCHECK(launch_lu(m, factors_f32.p, rowsrc32.p, half, flags.p + 2, stream)); });
        timed("ilu_coupling", [&] {
            HIPCHECK(launch_ilu_coupling(n, factors_f32.p, rowsrc32.p, ilu_Y32.p, ilu_S.p, basis.p, geo, stream));
        });
        pack_blocks(m, factors_f32.p, half);
        timed("ilu_update", [&] {
            HIPCHECK(launch_ilu_update(n, factors_f32.p + half * mm, ilu_S.p, fcache.p, basis.p, geo, stream));
        });
        timed("lu", [&] {
            HIPCHECK(launch_lu(m, factors_f32.p + half * mm, rowsrc32.p + static_cast<size_t>(half) * m, half,
                               flags.p + 2, stream));
        });
        require(read_flag(2) == 0, "Singular or nonfinite LU factors");
        pack_blocks(m, factors_f32.p + half * mm, half);
        ilu32_ready = true;
    }
    void ilu_apply32(const float *r, float *z) {
        require(ilu32_ready, "Refresh the block ILU(0) preconditioner first");
        require(r != z, "The block ILU(0) apply needs distinct input and output vectors");
        const int half = cells / 2;
        const size_t mm = static_cast<size_t>(m) * m;
        timed("ilu_apply", [&] {
            HIPCHECK(launch_bj_apply_colour(m, factors_f32.p, rowsrc32.p, r, ilu_z32.p, half, 0, cfg.nx, stream));
            HIPCHECK(launch_restrict(n, ilu_z32.p, tr_x.p, basis.p, geo, stream));
            HIPCHECK(launch_ilu_offdiag(n, 1, tr_x.p, fcache.p, r, ilu_w32.p, basis.p, geo, stream));
            HIPCHECK(launch_bj_apply_colour(m, factors_f32.p + half * mm, rowsrc32.p + static_cast<size_t>(half) * m,
                                            ilu_w32.p, z, half, 1, cfg.nx, stream));
            HIPCHECK(launch_restrict(n, z, tr_x.p, basis.p, geo, stream));
            HIPCHECK(launch_ilu_red_back(n, ilu_Y32.p, tr_x.p, fcache.p, ilu_z32.p, z, geo, stream));
        });
    }
    void precond_refresh() {
        if (cfg.precond == 1) ilu_refresh();
        else bj_refresh();
    }
    void precond_refresh_mixed() {
        if (cfg.precond == 1) ilu_refresh_mixed();
        else bj_refresh_mixed();
    }
    void precond_apply32(const float *r, float *z) {
        if (cfg.precond == 1) ilu_apply32(r, z);
        else bj_apply32(r, z);
    }
    void precond_apply(const double *r, double *z) {
        if (cfg.precond == 1) ilu_apply(r, z);
        else bj_apply(r, z);
    }
    void bj_apply(const double *r, double *z) {
        require(bj_ready, "Refresh the block Jacobi preconditioner first");
        timed("bj_apply", [&] { HIPCHECK(launch_bj_apply(m, factors.p, rowsrc.p, r, z, cells, stream)); });
    }
    // Adjoint block Jacobi: LU of each transposed diagonal block B_i^T, so
    // that the apply returns B_i^-T r with the same factor/apply kernels.
    // Shares the factor storage with bj_refresh.
    void bj_refresh_transposed() {
        require(refreshed, "Refresh the volume Jacobian before assembling a block");
        factors.alloc(static_cast<size_t>(cells) * m * m, bytes);
        rowsrc.alloc(static_cast<size_t>(cells) * m, bytes);
        bj_ready = false;
        set_flag(2, 0);
        timed("assemble", [&] {
            HIPCHECK(launch_assemble_blocks_transposed(n, vcache.p, fcache.p, factors.p, inv_step, basis.p, geo,
                                                       stream));
        });
        timed("lu", [&] { HIPCHECK(launch_lu(m, factors.p, rowsrc.p, cells, flags.p + 2, stream)); });
        require(read_flag(2) == 0, "Singular block Jacobi factor");
        pack_blocks(m, factors.p, cells);
        adj_bj_ready = true;
        ilu_ready = false;
    }
    void bj_apply_transposed(const double *r, double *z) {
        require(adj_bj_ready, "Refresh the transposed block Jacobi preconditioner first");
        timed("bj_apply", [&] { HIPCHECK(launch_bj_apply(m, factors.p, rowsrc.p, r, z, cells, stream)); });
    }

    // ---- GMRES (gmres.cpp:17-86): restarted, right-preconditioned, weighted
    //      CGS2, Givens, true residual after each cycle. Vectors of length N
    //      with the DG metric M^-1 (weights == nullptr) or a weight vector.
    struct Linear { int iterations; double relative; };
    using Op = std::function<void(const double *, double *)>;
    Linear gmres(const double *rhs, double *x, int limit, double tolerance) {
        return gmres_generic(1, D, nullptr, [&](const double *a, double *b) { dg_apply(a, b); },
                             [&](const double *a, double *b) { precond_apply(a, b); }, rhs, x, limit, tolerance);
    }
    // GMRES on J^T lambda = rhs, right-preconditioned by the transposed block
    // Jacobi; the residual of J^T lambda is a weak-form vector like the
    // primal one, so the same M^-1 metric applies
    Linear gmres_transposed(const double *rhs, double *x, int limit, double tolerance) {
        return gmres_generic(2, D, nullptr, [&](const double *a, double *b) { dg_apply_transpose(a, b); },
                             [&](const double *a, double *b) { bj_apply_transposed(a, b); }, rhs, x, limit,
                             tolerance);
    }
    double vec_norm(const double *x, long long N, const double *weights) {
        return std::sqrt(std::max(0.0, weights ? reduce_sq(x, N, 3, weights) : reduce_sq(x, N, 0)));
    }
    // solver: 1 DG, 2 transposed DG, 3 HDG Schur system (the operator pair,
    // part of the key of the captured step graphs)
    Linear gmres_generic(int solver, long long D, const double *weights, const Op &A, const Op &M, const double *rhs,
                         double *x, int limit, double tolerance) {
        const int restart = static_cast<int>(std::min<long long>(D, cfg.restart));
        require(restart + 1 <= 32, "GMRES restart above 31 is not supported on the device");
        kbasis.alloc(static_cast<size_t>(restart + 1) * D, bytes);
        kw.alloc(D, bytes);
        kz.alloc(D, bytes);
        kwork.alloc(D, bytes);
        kres.alloc(D, bytes);
        ky.alloc(64, bytes);
        std::vector<const double *> V(restart + 1);
        for (int k = 0; k <= restart; ++k) V[k] = kbasis.p + static_cast<size_t>(k) * D;
        auto Vk = [&](int k) { return kbasis.p + static_cast<size_t>(k) * D; };
        HIPCHECK(hipMemsetAsync(x, 0, D * sizeof(double), stream));
        copy(kres.p, rhs, D);
        const double rhs_norm = vec_norm(rhs, D, weights);
        if (rhs_norm == 0) return {0, 0};
        std::vector<double> H((restart + 1) * restart), cs(restart), sn(restart), g(restart + 1), y(restart);
        auto Hm = [&](int i, int j) -> double & { return H[i + (restart + 1) * j]; };
        int iterations = 0;
        double host_scalars[65];
        // A captured step records the operators' kernel arguments (buffers,
        // caches, step size): its graph is replayed while none of them moved.
        const bool use_graph = cfg.graph != 0 && !timing;
        if (use_graph) {
            gH.alloc(static_cast<size_t>(restart + 1) * restart, bytes);
            gcs.alloc(restart, bytes);
            gsn.alloc(restart, bytes);
            gg.alloc(restart + 1, bytes);
            gnext.alloc(1, bytes);
            graph_out();
            auto key = [&] {
                return std::vector<uint64_t>{static_cast<uint64_t>(solver), alloc_epoch(), bits_of(inv_step),
                                             static_cast<uint64_t>(restart), static_cast<uint64_t>(D),
                                             reinterpret_cast<uintptr_t>(weights)};
            };
            // before new captures, one eager application of both operators, so
            // that any buffer they allocate on first use exists
            if (key() != step_graphs.key) {
                M(rhs, kz.p);
                A(kz.p, kw.p);
            }
            step_graphs.prepare(key(), restart);
        }
        auto arnoldi_ortho = [&](int j) {
            HIPCHECK(launch_multi_dot(n, V.data(), j + 1, kw.p, weights, D, partials.p, basis.p, stream));
            HIPCHECK(launch_finish_sum(partials.p, REDUCE_BLOCKS, j + 1, scalars.p, stream));
            HIPCHECK(launch_update_dot(n, V.data(), j + 1, scalars.p, kw.p, weights, D, partials.p, 0, basis.p,
                                       stream));
            HIPCHECK(launch_finish_sum(partials.p, REDUCE_BLOCKS, j + 1, scalars.p + 32, stream));
            HIPCHECK(launch_update_dot(n, V.data(), j + 1, scalars.p + 32, kw.p, weights, D, partials.p, 1,
                                       basis.p, stream));
            HIPCHECK(launch_finish_sum(partials.p, REDUCE_BLOCKS, 1, scalars.p + 64, stream));
        };
        while (iterations < limit) {
            const double beta = vec_norm(kres.p, D, weights);
            if (beta <= tolerance * rhs_norm) return {iterations, beta / rhs_norm};
            HIPCHECK(launch_scale_div(kres.p, beta, Vk(0), D, stream));  // basis_.col(0) = residual_ / beta
            std::fill(H.begin(), H.end(), 0.0);
            std::fill(g.begin(), g.end(), 0.0);
            g[0] = beta;
            int count = 0;
            if (use_graph) {
                HIPCHECK(hipMemcpyAsync(gg.p, g.data(), (restart + 1) * sizeof(double), hipMemcpyHostToDevice, stream));
                for (int j = 0; j < restart && iterations < limit; ++j) {
                    if (!step_graphs.execs[j]) {
                        capture(&step_graphs.execs[j], [&] {
                            M(Vk(j), kz.p);
                            A(kz.p, kw.p);
                            arnoldi_ortho(j);
                            HIPCHECK(launch_arnoldi_givens(j, restart, scalars.p, gH.p, gcs.p, gsn.p, gg.p, gnext.p,
                                                           gout_dev, 2, stream));
                            HIPCHECK(launch_scale_div_dev(kw.p, gnext.p, Vk(j + 1), D, stream));
                        });
                    }
                    HIPCHECK(hipGraphLaunch(step_graphs.execs[j], stream));
                    sync();
                    const double gabs = gout_host[0], next = gout_host[1], d = gout_host[2];
                    require(std::isfinite(d) && d > 0, "GMRES Arnoldi breakdown");
                    ++iterations;
                    count = j + 1;
                    if (gabs <= tolerance * rhs_norm || next <= std::numeric_limits<double>::min()) break;
                }
                HIPCHECK(hipMemcpyAsync(H.data(), gH.p, H.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
                HIPCHECK(hipMemcpyAsync(g.data(), gg.p, g.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
                sync();
            }
            for (int j = 0; !use_graph && j < restart && iterations < limit; ++j) {
                M(Vk(j), kz.p);
                A(kz.p, kw.p);
                timed("gmres_ortho", [&] {
                    arnoldi_ortho(j);
                    HIPCHECK(hipMemcpyAsync(host_scalars, scalars.p, 65 * sizeof(double), hipMemcpyDeviceToHost,
                                            stream));
                    sync();
                });
                for (int k = 0; k <= j; ++k) Hm(k, j) = (0.0 + host_scalars[k]) + host_scalars[32 + k];
                const double next = std::sqrt(std::max(0.0, host_scalars[64]));
                Hm(j + 1, j) = next;
                if (next > 0) HIPCHECK(launch_scale_div(kw.p, next, Vk(j + 1), D, stream));
                for (int i = 0; i < j; ++i) {
                    const double a = Hm(i, j), b = Hm(i + 1, j);
                    Hm(i, j) = cs[i] * a + sn[i] * b;
                    Hm(i + 1, j) = -sn[i] * a + cs[i] * b;
                }
                const double d = std::hypot(Hm(j, j), Hm(j + 1, j));
                require(std::isfinite(d) && d > 0, "GMRES Arnoldi breakdown");
                cs[j] = Hm(j, j) / d;
                sn[j] = Hm(j + 1, j) / d;
                Hm(j, j) = d;
                Hm(j + 1, j) = 0;
                g[j + 1] = -sn[j] * g[j];
                g[j] *= cs[j];
                ++iterations;
                count = j + 1;
                if (std::abs(g[j + 1]) <= tolerance * rhs_norm || next <= std::numeric_limits<double>::min()) break;
            }
            for (int i = count - 1; i >= 0; --i) {
                double value = g[i];
                for (int k = i + 1; k < count; ++k) value -= Hm(i, k) * y[k];
                y[i] = value / Hm(i, i);
            }
            HIPCHECK(hipMemcpyAsync(ky.p, y.data(), count * sizeof(double), hipMemcpyHostToDevice, stream));
            HIPCHECK(launch_combine(V.data(), count, ky.p, kwork.p, D, stream));
            M(kwork.p, kw.p);
            axpby(1.0, x, 1.0, kw.p, x, D);
            A(x, kwork.p);
            axpby(1.0, rhs, -1.0, kwork.p, kres.p, D);
            const double relative = vec_norm(kres.p, D, weights) / rhs_norm;
            require(std::isfinite(relative), "Nonfinite GMRES residual");
            if (relative <= tolerance) return {iterations, relative};
        }
        throw std::runtime_error("GMRES iteration limit reached");
    }

    // ---- HDG (hybrid.cpp:31-121)
    void hdg_refresh_projected(const double *u, const double *r) {
        dg_refresh(u);  // volume + face linearisation caches (hybrid.cpp:57)
        const long long t = 16LL * n;
        hX.alloc(static_cast<size_t>(cells) * m * t, bytes);
        hY.alloc(static_cast<size_t>(D), bytes);
        hS.alloc(static_cast<size_t>(cells) * t * t, bytes);
        hfres.alloc(static_cast<size_t>(cells) * t, bytes);
        hloc.alloc(static_cast<size_t>(cells) * t, bytes);
        hrhs.alloc(static_cast<size_t>(T), bytes);
        // element-streamed: element blocks factorised in batches of at most ~40 GB
        const long long per = static_cast<long long>(m) * m * 8;
        const int batch = static_cast<int>(std::max<long long>(1, std::min<long long>(cells, 40000000000LL / per)));
        factors.alloc(static_cast<size_t>(batch) * m * m, bytes);
        rowsrc.alloc(static_cast<size_t>(batch) * m, bytes);
        const bool mfma_solve = !env_flag("DGH_HDG_SOLVE_VECTOR");
        if (mfma_solve) hdinv.alloc(hdg_solve_workspace(n, batch), bytes);
        bj_ready = false;
        set_flag(2, 0);
        for (int first = 0; first < cells; first += batch) {
            const int count = std::min(batch, cells - first);
            timed("hdg_condense", [&] {
                HIPCHECK(launch_assemble_hdg(n, vcache.p, fcache.p, factors.p, inv_step, basis.p, geo, first, count,
                                             stream));
                HIPCHECK(launch_lu(m, factors.p, rowsrc.p, count, flags.p + 2, stream));
                HIPCHECK(launch_hdg_solve(n, factors.p, rowsrc.p, fcache.p, r, hX.p, hY.p, first, count, basis.p,
                                          geo, stream, mfma_solve ? hdinv.p : nullptr));
            });
        }
        require(read_flag(2) == 0, "Singular or nonfinite LU factors");
        timed("hdg_condense", [&] {
            HIPCHECK(launch_hdg_cx(n, hX.p, hY.p, fcache.p, hS.p, hfres.p, basis.p, geo, stream));
            HIPCHECK(launch_hdg_face_sum(n, nullptr, 0.0, nullptr, hfres.p, hrhs.p, geo, stream));
        });
        hdg_ready = true;
        face_bj_ready = false;
    }
    void hdg_apply(const double *x, double *y) {
        require(hdg_ready, "Refresh the HDG condensation before applying it");
        hdiag.alloc(static_cast<size_t>(T), bytes);
        timed("schur", [&] {
            HIPCHECK(launch_hdg_diag_x(n, fcache.p, x, hdiag.p, basis.p, geo, stream));
            HIPCHECK(launch_hdg_schur(n, hS.p, x, hloc.p, geo, stream));
            HIPCHECK(launch_hdg_face_sum(n, nullptr, 0.0, hdiag.p, hloc.p, y, geo, stream));
        });
    }
    // hdg_apply with its two intermediate vectors in caller-given device
    // memory of `capacity` doubles (the fixed scratch): no allocation.
    void hdg_apply_fixed(const double *x, double *y, double *work, size_t capacity) {
        const size_t diag = (static_cast<size_t>(T) + 31) / 32 * 32;
        const size_t need = diag + static_cast<size_t>(cells) * 16 * n;
        if (need > capacity)
            throw std::runtime_error("dgh_hdg_apply_noalloc: this mesh needs " + std::to_string(need) +
                                     " doubles of scratch, beyond the fixed maximum of " +
                                     std::to_string(capacity) + " (use dgh_hdg_apply)");
        require(hdg_ready, "Refresh the HDG condensation before applying it");
        HIPCHECK(launch_hdg_diag_x(n, fcache.p, x, work, basis.p, geo, stream));
        HIPCHECK(launch_hdg_schur(n, hS.p, x, work + diag, geo, stream));
        HIPCHECK(launch_hdg_face_sum(n, nullptr, 0.0, work, work + diag, y, geo, stream));
    }
    void hdg_bj_refresh() {
        require(hdg_ready, "Refresh the HDG condensation before assembling a block");
        face_factors.alloc(static_cast<size_t>(faces) * fw * fw, bytes);
        face_rowsrc.alloc(static_cast<size_t>(faces) * fw, bytes);
        set_flag(2, 0);
        timed("face_lu", [&] {
            HIPCHECK(launch_hdg_face_blocks(n, hS.p, fcache.p, face_factors.p, basis.p, geo, stream));
            HIPCHECK(launch_lu(fw, face_factors.p, face_rowsrc.p, faces, flags.p + 2, stream));
        });
        require(read_flag(2) == 0, "Singular block Jacobi factor");
        pack_blocks(fw, face_factors.p, faces);
        face_bj_ready = true;
    }
    void hdg_bj_apply(const double *r, double *z) {
        require(face_bj_ready, "Refresh the block Jacobi preconditioner first");
        timed("face_bj_apply", [&] { HIPCHECK(launch_bj_apply(fw, face_factors.p, face_rowsrc.p, r, z, faces, stream)); });
    }
    void hdg_recover(const double *dl, double *du) {
        require(hdg_ready, "Refresh the HDG condensation before applying it");
        HIPCHECK(launch_hdg_recover(n, hX.p, hY.p, dl, du, geo, stream));
    }
    void hdg_recovered_residual(const double *e, double *out) {
        require(hdg_ready, "Refresh the HDG condensation before applying it");
        HIPCHECK(launch_hdg_recovered_residual(n, e, fcache.p, out, basis.p, geo, stream));
    }
    const double *inverse_trace_mass() {
        if (!itm.p) {
            itm.alloc(static_cast<size_t>(T), bytes);
            HIPCHECK(launch_inverse_trace_mass(n, itm.p, basis.p, geo, stream));
        }
        return itm.p;
    }
    Linear hdg_gmres(const double *rhs, double *x, int limit, double tolerance) {
        return gmres_generic(3, T, inverse_trace_mass(), [&](const double *a, double *b) { hdg_apply(a, b); },
                             [&](const double *a, double *b) { hdg_bj_apply(a, b); }, rhs, x, limit, tolerance);
    }
    void scale_div(const double *src, double d, double *dst) { HIPCHECK(launch_scale_div(src, d, dst, D, stream)); }

    // ---- mixed precision: FP32 operator, FP32/BF16 block Jacobi, FP32 inner
    //      GMRES inside FP64 flexible GMRES (no reference counterpart)
    bool pair_rows = false;  // layout of factors_bf16: row pairs (comparison path) or column pairs
    bool two_byte_factors() const {
        return cfg.precision == DGH_PRECISION_MIXED_BF16 || cfg.precision == DGH_PRECISION_MIXED_FP16;
    }
    void bj_refresh_mixed() {
        require(refreshed, "Refresh the volume Jacobian before assembling a block");
        factors_f32.alloc(static_cast<size_t>(cells) * m * m, bytes);
        rowsrc32.alloc(static_cast<size_t>(cells) * m, bytes);
        set_flag(2, 0);
        timed("assemble", [&] {
            HIPCHECK(launch_assemble_blocks(n, vcache.p, fcache.p, factors_f32.p, inv_step, basis.p, geo, 0, cells,
                                            stream));
        });
        const bool fp16 = cfg.precision == DGH_PRECISION_MIXED_FP16;
        if (fp16) factor_umax.alloc(cells, bytes);
        timed("lu", [&] {
            HIPCHECK(launch_lu(m, factors_f32.p, rowsrc32.p, cells, flags.p + 2, stream, fp16 ? factor_umax.p : nullptr));
        });
        require(read_flag(2) == 0, "Singular block Jacobi factor");
        ilu32_ready = false;
        // the 2-byte forms convert the square LU factors directly; only the
        // FP32 apply reads the packed layout
        if (two_byte_factors()) {
            pair_rows = two_byte_rowpairs_requested();  // fixed until the next refresh
            factors_bf16.alloc(static_cast<size_t>(cells) * pair_factor_stride(m, pair_rows), bytes);
            factor_scale.alloc(cells, bytes);
            factor_rdiag.alloc(static_cast<size_t>(cells) * m, bytes);
            set_flag(2, 0);
            timed("pack", [&] {
                HIPCHECK(launch_to_pairs(m, factors_f32.p, factors_bf16.p, factor_scale.p, factor_rdiag.p,
                                         factor_umax.p, cells, fp16, pair_rows, flags.p + 2, stream));
            });
            require(read_flag(2) == 0, "Singular block Jacobi factor");
        } else {
            pack_blocks(m, factors_f32.p, cells);
        }
        bj32_ready = true;
    }
    void bj_apply32(const float *r, float *z) {
        require(bj32_ready, "Refresh the block Jacobi preconditioner first");
        timed("bj_apply", [&] {
            if (two_byte_factors())
                HIPCHECK(launch_bj_apply_pairs(m, factors_bf16.p, factor_scale.p, factor_rdiag.p, rowsrc32.p, r, z,
                                               cells, cfg.precision == DGH_PRECISION_MIXED_FP16, pair_rows, stream));
            else
                HIPCHECK(launch_bj_apply(m, factors_f32.p, rowsrc32.p, r, z, cells, stream));
        });
    }
    void dg_apply32(const float *x, float *y) {
        require(refreshed, "Refresh the volume Jacobian before applying it");
        timed("jv", [&] {
            HIPCHECK(launch_restrict(n, x, tr_x.p, basis.p, geo, stream));
            HIPCHECK(launch_jv_flux(n, tr_x.p, fcache.p, fq.p, geo, stream));
            if (oi()) HIPCHECK(launch_oi_element_jv(n, x, vcache.p, fq.p, y, inv_step, basis.p, geo, stream));
            else HIPCHECK(launch_element_jv(n, x, vcache.p, fq.p, y, inv_step, basis.p, geo, stream));
        });
    }
    void dg_apply_mixed(const float *x, double *y) {
        require(refreshed, "Refresh the volume Jacobian before applying it");
        timed("jv", [&] {
            HIPCHECK(launch_restrict(n, x, tr_x.p, basis.p, geo, stream));
            HIPCHECK(launch_jv_flux(n, tr_x.p, fcache.p, fq.p, geo, stream));
            if (oi()) HIPCHECK(launch_oi_element_jv(n, x, vcache.p, fq.p, y, inv_step, basis.p, geo, stream));
            else HIPCHECK(launch_element_jv(n, x, vcache.p, fq.p, y, inv_step, basis.p, geo, stream));
        });
    }
    double dot32(const float *a) {
        const float *v[1] = {a};
        HIPCHECK(launch_multi_dot(n, v, 1, a, nullptr, D, partials.p, basis.p, stream));
        HIPCHECK(launch_finish_sum(partials.p, REDUCE_BLOCKS, 1, scalars.p + 200, stream));
        double out = 0;
        HIPCHECK(hipMemcpyAsync(&out, scalars.p + 200, sizeof(double), hipMemcpyDeviceToHost, stream));
        sync();
        return std::sqrt(std::max(0.0, out));
    }
    // Right-preconditioned FP32 GMRES(inner_restart) on J z = b with the FP32
    // block Jacobi, the M^-1 metric and FP64 dot accumulation; stops at the
    // Givens residual estimate tol * ||b|| or after max_its iterations.
    // Inner FP32 GMRES: classical Gram-Schmidt with DGH_INNER_GS_PASSES
    // passes per Arnoldi step (2: re-orthogonalised, as the FP64 solver; 1:
    // one projection, then the update fused with the norm).
    int gmres_inner(const float *b, float *z, double tol, int max_its) {
        const int restart = std::min<int>(cfg.inner_restart, static_cast<int>(std::min<long long>(D, 31)));
        ibasis.alloc(static_cast<size_t>(restart + 1) * D, bytes);
        iw.alloc(D, bytes);
        it.alloc(D, bytes);
        is.alloc(D, bytes);
        ires.alloc(D, bytes);
        std::vector<const float *> V(restart + 1);
        for (int k = 0; k <= restart; ++k) V[k] = ibasis.p + static_cast<size_t>(k) * D;
        auto Vk = [&](int k) { return ibasis.p + static_cast<size_t>(k) * D; };
        HIPCHECK(hipMemsetAsync(z, 0, D * sizeof(float), stream));
        HIPCHECK(hipMemcpyAsync(ires.p, b, D * sizeof(float), hipMemcpyDeviceToDevice, stream));
        const double rhs_norm = dot32(b);
        if (rhs_norm == 0) return 0;
        std::vector<double> H((restart + 1) * restart), cs(restart), sn(restart), g(restart + 1), y(restart);
        auto Hm = [&](int i, int j) -> double & { return H[i + (restart + 1) * j]; };
        int iterations = 0;
        double host_scalars[65];
        auto ortho = [&](int j) {
            HIPCHECK(launch_multi_dot(n, V.data(), j + 1, iw.p, nullptr, D, partials.p, basis.p, stream));
            HIPCHECK(launch_finish_sum(partials.p, REDUCE_BLOCKS, j + 1, scalars.p, stream));
            if (DGH_INNER_GS_PASSES == 2) {
                HIPCHECK(launch_update_dot(n, V.data(), j + 1, scalars.p, iw.p, nullptr, D, partials.p, 0, basis.p,
                                           stream));
                HIPCHECK(launch_finish_sum(partials.p, REDUCE_BLOCKS, j + 1, scalars.p + 32, stream));
            }
            HIPCHECK(launch_update_dot(n, V.data(), j + 1, scalars.p + (DGH_INNER_GS_PASSES == 2 ? 32 : 0), iw.p,
                                       nullptr, D, partials.p, 1, basis.p, stream));
            HIPCHECK(launch_finish_sum(partials.p, REDUCE_BLOCKS, 1, scalars.p + 64, stream));
        };
        // Captured Arnoldi steps as in gmres_generic (the Givens update on the
        // device); the graphs are kept while the key is unchanged.
        const bool use_graph = cfg.graph != 0 && !timing;
        if (use_graph) {
            igH.alloc(static_cast<size_t>(restart + 1) * restart, bytes);
            igcs.alloc(restart, bytes);
            igsn.alloc(restart, bytes);
            igg.alloc(restart + 1, bytes);
            ignext.alloc(1, bytes);
            graph_out();
            auto key = [&] {
                return std::vector<uint64_t>{10, alloc_epoch(), bits_of(inv_step), static_cast<uint64_t>(restart),
                                             static_cast<uint64_t>(D)};
            };
            if (key() != inner_graphs.key) {
                precond_apply32(b, it.p);
                dg_apply32(it.p, iw.p);
            }
            inner_graphs.prepare(key(), restart);
        }
        while (iterations < max_its) {
            const double beta = dot32(ires.p);
            if (beta <= tol * rhs_norm) break;
            HIPCHECK(launch_scale_div(ires.p, beta, Vk(0), D, stream));
            std::fill(H.begin(), H.end(), 0.0);
            std::fill(g.begin(), g.end(), 0.0);
            g[0] = beta;
            int count = 0;
            bool done = false;
            if (use_graph) {
                HIPCHECK(hipMemcpyAsync(igg.p, g.data(), (restart + 1) * sizeof(double), hipMemcpyHostToDevice, stream));
                for (int j = 0; j < restart && iterations < max_its; ++j) {
                    if (!inner_graphs.execs[j]) {
                        capture(&inner_graphs.execs[j], [&] {
                            precond_apply32(Vk(j), it.p);
                            dg_apply32(it.p, iw.p);
                            ortho(j);
                            HIPCHECK(launch_arnoldi_givens(j, restart, scalars.p, igH.p, igcs.p, igsn.p, igg.p,
                                                           ignext.p, gout_dev + 4, DGH_INNER_GS_PASSES, stream));
                            HIPCHECK(launch_scale_div_dev(iw.p, ignext.p, Vk(j + 1), D, stream));
                        });
                    }
                    HIPCHECK(hipGraphLaunch(inner_graphs.execs[j], stream));
                    sync();
                    const double gabs = gout_host[4], next = gout_host[5], d = gout_host[6];
                    require(std::isfinite(d) && d > 0, "GMRES Arnoldi breakdown");
                    ++iterations;
                    count = j + 1;
                    if (gabs <= tol * rhs_norm || next <= std::numeric_limits<double>::min()) {
                        done = true;
                        break;
                    }
                }
                HIPCHECK(hipMemcpyAsync(H.data(), igH.p, H.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
                HIPCHECK(hipMemcpyAsync(g.data(), igg.p, g.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
                sync();
            }
            for (int j = 0; !use_graph && j < restart && iterations < max_its; ++j) {
                precond_apply32(Vk(j), it.p);
                dg_apply32(it.p, iw.p);
                timed("gmres_ortho", [&] {
                    ortho(j);
                    HIPCHECK(hipMemcpyAsync(host_scalars, scalars.p, 65 * sizeof(double), hipMemcpyDeviceToHost,
                                            stream));
                    sync();
                });
                for (int k = 0; k <= j; ++k)
                    Hm(k, j) = (0.0 + host_scalars[k]) + (DGH_INNER_GS_PASSES == 2 ? host_scalars[32 + k] : 0.0);
                const double next = std::sqrt(std::max(0.0, host_scalars[64]));
                Hm(j + 1, j) = next;
                if (next > 0) HIPCHECK(launch_scale_div(iw.p, next, Vk(j + 1), D, stream));
                for (int i = 0; i < j; ++i) {
                    const double a = Hm(i, j), bb = Hm(i + 1, j);
                    Hm(i, j) = cs[i] * a + sn[i] * bb;
                    Hm(i + 1, j) = -sn[i] * a + cs[i] * bb;
                }
                const double d = std::hypot(Hm(j, j), Hm(j + 1, j));
                require(std::isfinite(d) && d > 0, "GMRES Arnoldi breakdown");
                cs[j] = Hm(j, j) / d;
                sn[j] = Hm(j + 1, j) / d;
                Hm(j, j) = d;
                Hm(j + 1, j) = 0;
                g[j + 1] = -sn[j] * g[j];
                g[j] *= cs[j];
                ++iterations;
                count = j + 1;
                if (std::abs(g[j + 1]) <= tol * rhs_norm || next <= std::numeric_limits<double>::min()) {
                    done = true;
                    break;
                }
            }
            for (int i = count - 1; i >= 0; --i) {
                double value = g[i];
                for (int k = i + 1; k < count; ++k) value -= Hm(i, k) * y[k];
                y[i] = value / Hm(i, i);
            }
            HIPCHECK(hipMemcpyAsync(ky.p, y.data(), count * sizeof(double), hipMemcpyHostToDevice, stream));
            HIPCHECK(launch_combine(V.data(), count, ky.p, it.p, D, stream));
            precond_apply32(it.p, is.p);
            HIPCHECK(launch_axpby(1.0, z, 1.0, is.p, z, D, stream));
            if (done || iterations >= max_its) break;
            dg_apply32(z, iw.p);
            HIPCHECK(launch_axpby(1.0, b, -1.0, iw.p, ires.p, D, stream));
        }
        return iterations;
    }
    // Flexible GMRES(restart) in FP64 (weighted CGS2, Givens, true residual
    // after each cycle as gmres.cpp:17-86) whose variable right
    // preconditioner is one FP32 inner solve per outer iteration.
    Linear fgmres(const double *rhs, double *x, int limit, double tolerance, int &outer) {
        const int restart = static_cast<int>(std::min<long long>(D, cfg.restart));
        require(restart + 1 <= 32, "GMRES restart above 31 is not supported on the device");
        kbasis.alloc(static_cast<size_t>(restart + 1) * D, bytes);
        zbasis.alloc(static_cast<size_t>(restart) * D, bytes);
        kw.alloc(D, bytes);
        kwork.alloc(D, bytes);
        kres.alloc(D, bytes);
        ky.alloc(64, bytes);
        ib.alloc(D, bytes);
        std::vector<const double *> V(restart + 1);
        std::vector<const float *> Z(restart);
        for (int k = 0; k <= restart; ++k) V[k] = kbasis.p + static_cast<size_t>(k) * D;
        for (int k = 0; k < restart; ++k) Z[k] = zbasis.p + static_cast<size_t>(k) * D;
        auto Vk = [&](int k) { return kbasis.p + static_cast<size_t>(k) * D; };
        auto Zk = [&](int k) { return zbasis.p + static_cast<size_t>(k) * D; };
        HIPCHECK(hipMemsetAsync(x, 0, D * sizeof(double), stream));
        copy(kres.p, rhs, D);
        outer = 0;
        const double rhs_norm = gmres_norm(rhs);
        if (rhs_norm == 0) return {0, 0};
        std::vector<double> H((restart + 1) * restart), cs(restart), sn(restart), g(restart + 1), y(restart);
        auto Hm = [&](int i, int j) -> double & { return H[i + (restart + 1) * j]; };
        int inner_total = 0;
        double host_scalars[65];
        while (inner_total < limit) {
            const double beta = gmres_norm(kres.p);
            if (beta <= tolerance * rhs_norm) return {inner_total, beta / rhs_norm};
            scale_div(kres.p, beta, Vk(0));
            std::fill(H.begin(), H.end(), 0.0);
            std::fill(g.begin(), g.end(), 0.0);
            g[0] = beta;
            int count = 0;
            for (int j = 0; j < restart && inner_total < limit; ++j) {
                HIPCHECK(launch_scale_div(Vk(j), 1.0, ib.p, D, stream));
                inner_total += gmres_inner(ib.p, Zk(j), cfg.inner_tolerance,
                                           std::max(1, std::min(cfg.inner_max, limit - inner_total)));
                dg_apply_mixed(Zk(j), kw.p);
                timed("gmres_ortho", [&] {
                    HIPCHECK(launch_multi_dot(n, V.data(), j + 1, kw.p, nullptr, D, partials.p, basis.p, stream));
                    HIPCHECK(launch_finish_sum(partials.p, REDUCE_BLOCKS, j + 1, scalars.p, stream));
                    HIPCHECK(launch_update_dot(n, V.data(), j + 1, scalars.p, kw.p, nullptr, D, partials.p, 0,
                                               basis.p, stream));
                    HIPCHECK(launch_finish_sum(partials.p, REDUCE_BLOCKS, j + 1, scalars.p + 32, stream));
                    HIPCHECK(launch_update_dot(n, V.data(), j + 1, scalars.p + 32, kw.p, nullptr, D, partials.p, 1,
                                               basis.p, stream));
                    HIPCHECK(launch_finish_sum(partials.p, REDUCE_BLOCKS, 1, scalars.p + 64, stream));
                    HIPCHECK(hipMemcpyAsync(host_scalars, scalars.p, 65 * sizeof(double), hipMemcpyDeviceToHost,
                                            stream));
                    sync();
                });
                for (int k = 0; k <= j; ++k) Hm(k, j) = (0.0 + host_scalars[k]) + host_scalars[32 + k];
                const double next = std::sqrt(std::max(0.0, host_scalars[64]));
                Hm(j + 1, j) = next;
                if (next > 0) scale_div(kw.p, next, Vk(j + 1));
                for (int i = 0; i < j; ++i) {
                    const double a = Hm(i, j), b = Hm(i + 1, j);
                    Hm(i, j) = cs[i] * a + sn[i] * b;
                    Hm(i + 1, j) = -sn[i] * a + cs[i] * b;
                }
                const double d = std::hypot(Hm(j, j), Hm(j + 1, j));
                require(std::isfinite(d) && d > 0, "GMRES Arnoldi breakdown");
                cs[j] = Hm(j, j) / d;
                sn[j] = Hm(j + 1, j) / d;
                Hm(j, j) = d;
                Hm(j + 1, j) = 0;
                g[j + 1] = -sn[j] * g[j];
                g[j] *= cs[j];
                ++outer;
                count = j + 1;
                if (std::abs(g[j + 1]) <= tolerance * rhs_norm || next <= std::numeric_limits<double>::min()) break;
            }
            for (int i = count - 1; i >= 0; --i) {
                double value = g[i];
                for (int k = i + 1; k < count; ++k) value -= Hm(i, k) * y[k];
                y[i] = value / Hm(i, i);
            }
            HIPCHECK(hipMemcpyAsync(ky.p, y.data(), count * sizeof(double), hipMemcpyHostToDevice, stream));
            HIPCHECK(launch_combine_add(Z.data(), count, ky.p, x, D, stream));
            dg_apply(x, kwork.p);
            axpby(1.0, rhs, -1.0, kwork.p, kres.p, D);
            const double relative = gmres_norm(kres.p) / rhs_norm;
            require(std::isfinite(relative), "Nonfinite GMRES residual");
            if (relative <= tolerance) return {inner_total, relative};
        }
        throw std::runtime_error("GMRES iteration limit reached");
    }

    // One damped Newton update of the current stage (time_integrator.cpp:104-179):
    // state.p holds the iterate and residual.p its residual, whose dual norm is
    // `norm`; tau, old and the frozen speeds are the stage's, old_norm = ||old||_M.
    // Solves for the direction (DG or condensed HDG; FP64, or mixed precision
    // inside FP64 FGMRES), checks the linear residual against the reference's
    // target, and line-searches along the direction; the accepted iterate is
    // left in state.p and its residual in residual.p.
    struct UpdateResult {
        Linear linear;
        int outer;
        double precision_error, error_norm, damping;
    };
    UpdateResult newton_update(double norm, double old_norm) {
        const auto &c = cfg;
        const double allowance = 128 * std::numeric_limits<double>::epsilon() *
                                 ((state_norm(state.p) + old_norm) / tau + norm);
        const double target = std::max(10 * c.linear_tolerance * norm, allowance);
        Linear linear{0, 0};
        int outer = 0;
        double precision_error = -1;
        if (c.method == DGH_METHOD_HDG) {
            // time_integrator.cpp:164-191: projected condensation, face
            // block Jacobi, Schur GMRES with up to four tolerance cuts,
            // acceptance on the recovered DG residual
            hdg_refresh_projected(state.p, residual.p);
            hdg_bj_refresh();
            trace_dir.alloc(static_cast<size_t>(T), bytes);
            hlambda.alloc(static_cast<size_t>(T), bytes);
            double tolerance = c.linear_tolerance;
            bool valid = false;
            for (int attempt = 0; attempt < 4; ++attempt) {
                const Linear current = hdg_gmres(hrhs.p, trace_dir.p, c.max_linear, tolerance);
                linear.iterations += current.iterations;
                linear.relative = current.relative;
                hdg_recover(trace_dir.p, direction.p);
                axpby(-1.0, kres.p, 0.0, nullptr, hlambda.p, T);
                hdg_recovered_residual(hlambda.p, equation_error.p);
                valid = dual_norm(equation_error.p) <= target;
                if (valid) break;
                tolerance *= .01;
            }
            require(valid, "Recovered DG residual criterion failed after Schur refinement");
        } else {
            dg_refresh(state.p);
            axpby(-1.0, residual.p, 0.0, nullptr, rhs.p, D);
            if (c.precision == DGH_PRECISION_FP64) {
                precond_refresh();
                linear = gmres(rhs.p, direction.p, c.max_linear, c.linear_tolerance);
            } else {
                precond_refresh_mixed();
                linear = fgmres(rhs.p, direction.p, c.max_linear, c.linear_tolerance, outer);
            }
            axpby(-1.0, kres.p, 0.0, nullptr, equation_error.p, D);
            if (c.precision != DGH_PRECISION_FP64 && c.check_update) {
                // The same Newton system through the FP64 path. The two
                // factor pools are never resident together (at 512^2, k=8
                // they are 221 GB + 111 GB): the FP32/BF16 factors are
                // released first and the FP64 ones after the check.
                sync();
                factors_f32.release(bytes);
                factors_bf16.release(bytes);
                factor_scale.release(bytes);
                factor_rdiag.release(bytes);
                ilu_Y32.release(bytes);
                bj32_ready = false;
                ilu32_ready = false;
                check_dir.alloc(D, bytes);
                precond_refresh();
                gmres(rhs.p, check_dir.p, c.max_linear, c.linear_tolerance);
                axpby(1.0, direction.p, -1.0, check_dir.p, kwork.p, D);
                precision_error = state_norm(kwork.p) / state_norm(check_dir.p);
                factors.release(bytes);
                ilu_Y.release(bytes);
                bj_ready = false;
                ilu_ready = false;
            }
        }
        const double error_norm = dual_norm(equation_error.p);
        require(error_norm <= target, "DG Newton linear residual too large");
        if (c.method == DGH_METHOD_HDG && c.check_update) {
            // time_integrator.cpp:194-203: independent DG check of the recovered update
            indep.alloc(D, bytes);
            dg_apply(direction.p, indep.p);
            axpby(1.0, indep.p, 1.0, residual.p, indep.p, D);
            require(dual_norm(indep.p) <= std::max(4 * target, 1e-12),
                    "Independent recovered DG equation check failed");
            axpby(1.0, indep.p, -1.0, equation_error.p, indep.p, D);
            require(dual_norm(indep.p) <= std::max(4 * allowance, 1e-12), "Recovered residual identity failed");
        }
        double damping = 1.;
        bool accepted = false;
        for (int search = 0; search < 30; ++search) {
            axpby(1.0, state.p, damping, direction.p, trial.p, D);
            if (admissible(trial.p)) {
                residual_into(trial.p, trial_residual.p);
                if (dual_norm(trial_residual.p) <= (1 - 1e-4 * damping) * norm + allowance) {
                    accepted = true;
                    break;
                }
            }
            damping *= .5;
        }
        require(accepted, "Newton line search failed");
        std::swap(state, trial);
        std::swap(residual, trial_residual);
        return {linear, outer, precision_error, error_norm, damping};
    }
    void newton_buffers() {
        state.alloc(D, bytes);
        trial.alloc(D, bytes);
        residual.alloc(D, bytes);
        trial_residual.alloc(D, bytes);
        direction.alloc(D, bytes);
        rhs.alloc(D, bytes);
        equation_error.alloc(D, bytes);
    }

    // ---- TimeIntegrator::advance (time_integrator.cpp:48-213), DG method
    void advance(double *solution, dgh_run_stats &stats, std::vector<dgh_update_stats> &updates) {
        const auto &c = cfg;
        require(c.method != DGH_METHOD_HDG || c.precision == DGH_PRECISION_FP64,
                "Mixed precision is available for the DG method only");
        require(!c.direct, "--linear direct is a host reference and is not available on the device path");
        require(admissible(solution), "Invalid initial solution");
        const double shock = (c.testcase == DGH_WAVE || c.testcase == DGH_SEEDED)
                                 ? std::numeric_limits<double>::infinity()
                                 : 4 / ((geo.gamma + 1) * std::acos(-1.));
        require(c.steps > 0 || c.final_time < shock, "Final time reaches simple-wave shock");
        constexpr double alpha = .24169426078821, beta = alpha / 4., eta = .12915286960590;
        constexpr double a[4][4] = {{alpha, 0, 0, 0}, {-alpha, alpha, 0, 0}, {0, 1 - alpha, alpha, 0},
                                    {beta, eta, .5 - beta - eta - alpha, alpha}};
        constexpr double b[4] = {0, 1. / 6, 1. / 6, 2. / 3}, stage_c[4] = {alpha, 0, 1, .5};
        const int stages = c.integrator == 0 ? 1 : 4;
        std::memset(&stats, 0, sizeof(stats));
        freeze(solution);
        const auto initial_integral = integral(solution);
        newton_buffers();
        step_state.alloc(D, bytes);
        stage_base.alloc(D, bytes);
        if (stages > 1) slopes.alloc(static_cast<size_t>(4) * D, bytes);
        auto slope = [&](int s) { return slopes.p + static_cast<size_t>(s) * D; };
        const long long speed_count = 2LL * faces * n;
        if (trajectory == RECORD) {
            require(c.steps > 0, "Recording a trajectory for the adjoint needs a fixed step count (--steps N)");
            const size_t records = static_cast<size_t>(c.steps) * stages;
            ck_states.alloc(records * D, bytes);
            ck_speeds.alloc(records * speed_count, bytes);
            ck_step.clear();
            ck_rate.clear();
            ck_stages = stages;
        }
        if (trajectory == REPLAY)
            require(!ck_step.empty() && ck_stages == stages, "Record a trajectory before replaying it");
        const int fixed_steps = trajectory == REPLAY ? static_cast<int>(ck_step.size()) : c.steps;
        const auto begin = Clock::now();
        while (fixed_steps > 0 ? stats.steps < fixed_steps : stats.final_time < c.final_time) {
            // a replay takes the recorded step and frozen speeds instead of freezing again
            const double rate = trajectory == REPLAY ? (load_frozen(stats.steps * stages), ck_rate[stats.steps])
                                                     : freeze(solution);
            double step = trajectory == REPLAY ? ck_step[stats.steps] : c.cfl / rate;
            if (fixed_steps == 0) {
                const double remaining = c.final_time - stats.final_time;
                step = std::min(step, remaining);
                if (remaining - step <= 16 * std::numeric_limits<double>::epsilon() * std::max(remaining, c.final_time))
                    step = remaining;
            }
            require(std::isfinite(step) && step > 0 && stats.final_time + step > stats.final_time, "Invalid CFL step");
            require(stats.final_time + step < shock, "Step reaches simple-wave shock");
            copy(step_state.p, solution, D);
            copy(state.p, solution, D);
            for (int stage = 0; stage < stages; ++stage) {
                const double diagonal = stages == 1 ? 1. : a[stage][stage];
                copy(stage_base.p, step_state.p, D);
                for (int j = 0; j < stage; ++j) axpby(1.0, stage_base.p, step * a[stage][j], slope(j), stage_base.p, D);
                const size_t record = static_cast<size_t>(stats.steps) * stages + stage;
                if (stage > 0) {
                    if (trajectory == REPLAY) load_frozen(record);
                    else freeze(state.p);
                }
                tau = diagonal * step;
                copy(old.p, stage_base.p, D);
                refreshed = false;
                bj_ready = false;
                const double old_norm = state_norm(stage_base.p);
                const int limit = c.newton_steps > 0 ? c.newton_steps : c.max_newton;
                double initial_rms = -1.;
                bool converged = false, residual_ready = false;
                for (int it = 0; it <= limit; ++it) {
                    const auto update_start = Clock::now();
                    if (!residual_ready) residual_into(state.p, residual.p);
                    const double norm = dual_norm(residual.p), rms = norm / std::sqrt(geo.lx * geo.ly);
                    if (initial_rms < 0) initial_rms = rms;
                    if (c.newton_steps == 0 && rms <= c.nonlinear_tolerance * std::max(1., initial_rms)) {
                        converged = true;
                        break;
                    }
                    if (it == limit) break;
                    const UpdateResult up = newton_update(norm, old_norm);
                    const Linear &linear = up.linear;
                    const int outer = up.outer;
                    const double precision_error = up.precision_error, error_norm = up.error_norm,
                                 damping = up.damping;
                    residual_ready = true;
                    sync();
                    const double duration = elapsed(update_start);
                    stats.max_linear_residual = std::max(stats.max_linear_residual, linear.relative);
                    const double dgrel = norm > 0 ? error_norm / norm : 0.;
                    stats.max_dg_linear_residual = std::max(stats.max_dg_linear_residual, dgrel);
                    stats.krylov_iterations += linear.iterations;
                    if (stats.newton_iterations >= c.warmup_updates) {
                        stats.measured_newton_seconds += duration;
                        ++stats.measured_updates;
                    }
                    ++stats.newton_iterations;
                    dgh_update_stats u{};
                    u.step = stats.steps;
                    u.stage = stage;
                    u.iteration = it;
                    u.krylov_iterations = linear.iterations;
                    u.time = stats.final_time + step;
                    u.stage_time = stats.final_time + (stages == 1 ? 1. : stage_c[stage]) * step;
                    u.dt = step;
                    u.cfl = step * rate;
                    u.residual_physical_rms = rms;
                    u.damping = damping;
                    u.linear_relative_residual = linear.relative;
                    u.dg_linear_relative_residual = dgrel;
                    u.iteration_s = duration;
                    u.inner_iterations = linear.iterations;
                    u.outer_iterations = outer;
                    u.precision_error = precision_error;
                    updates.push_back(u);
                }
                require(converged || c.newton_steps > 0, "Newton iteration limit reached");
                if (trajectory == RECORD) {
                    copy(ck_states.p + record * D, state.p, D);
                    copy(ck_speeds.p + record * speed_count, speeds.p, speed_count);
                }
                if (stages > 1) HIPCHECK(launch_sub_div(state.p, stage_base.p, tau, slope(stage), D, stream));
            }
            if (trajectory == RECORD) {
                ck_step.push_back(step);
                ck_rate.push_back(rate);
            }
            if (stages == 1) copy(solution, state.p, D);
            else {
                copy(solution, step_state.p, D);
                for (int s = 0; s < stages; ++s) axpby(1.0, solution, step * b[s], slope(s), solution, D);
                require(admissible(solution), "DIRK3 step produced an inadmissible state");
            }
            stats.final_time = c.steps == 0 && step >= c.final_time - stats.final_time ? c.final_time
                                                                                       : stats.final_time + step;
            stats.minimum_step = stats.steps == 0 ? step : std::min(stats.minimum_step, step);
            stats.maximum_step = std::max(stats.maximum_step, step);
            ++stats.steps;
            const auto now = integral(solution);
            double drift = 0;
            for (int v = 0; v < 4; ++v) drift = std::max(drift, std::abs(now[v] - initial_integral[v]));
            stats.conservation_drift = std::max(stats.conservation_drift, drift);
        }
        sync();
        stats.seconds = elapsed(begin);
    }

    // ---- Discrete adjoint of BE / DIRK3 on a recorded trajectory
    //
    // Frozen schedule: the HLL speeds and the step sizes of the recorded run
    // are constants of the map u_0 -> u_N (the reference never differentiates
    // the speeds: its Jacobians use the frozen ones), so the gradient is that
    // of a replay with the same schedule. Stage s of a step of size h solves
    //   M k_s + F_s(U_s) = 0,  U_s = base_s + tau_s k_s,  tau_s = a_ss h,
    //   base_s = u_n + h sum_{j<s} a_sj k_j,  u_{n+1} = u_n + h sum_s b_s k_s
    // (BE: one stage, a = b = 1). With A_s = M / tau_s + dF_s/dU (the Newton
    // operator at the checkpointed U_s) the reverse sweep is
    //   kbar_s += h b_s ubar;  A_s^T lambda = kbar_s / tau_s;
    //   basebar = (M lambda - kbar_s) / tau_s;  ubar += basebar;
    //   kbar_j += h a_sj basebar (j < s)
    // with each A_s rebuilt from its checkpoint (recompute, not storage).
    void load_frozen(size_t record) {
        const long long speed_count = 2LL * faces * n;
        require(record < ck_step.size() * static_cast<size_t>(ck_stages), "Trajectory record out of range");
        copy(speeds.p, ck_speeds.p + record * speed_count, speed_count);
        frozen = true;
        refreshed = false;
    }
    struct AdjointSweep { int krylov; int solves; double max_relative; };
    AdjointSweep adjoint_backward(const double *final_gradient, double *gradient) {
        require(!ck_step.empty() && ck_stages > 0, "Record a trajectory before the adjoint sweep");
        constexpr double alpha = .24169426078821, beta = alpha / 4., eta = .12915286960590;
        constexpr double dirk_a[4][4] = {{alpha, 0, 0, 0}, {-alpha, alpha, 0, 0}, {0, 1 - alpha, alpha, 0},
                                         {beta, eta, .5 - beta - eta - alpha, alpha}};
        constexpr double dirk_b[4] = {0, 1. / 6, 1. / 6, 2. / 3};
        const int stages = ck_stages;
        auto a = [&](int s, int j) { return stages == 1 ? 1.0 : dirk_a[s][j]; };
        auto b = [&](int s) { return stages == 1 ? 1.0 : dirk_b[s]; };
        adj_kbar.alloc(static_cast<size_t>(stages) * D, bytes);
        adj_lambda.alloc(D, bytes);
        adj_rhs.alloc(D, bytes);
        adj_base.alloc(D, bytes);
        auto kbar = [&](int s) { return adj_kbar.p + static_cast<size_t>(s) * D; };
        AdjointSweep out{0, 0, 0};
        copy(gradient, final_gradient, D);
        for (int step = static_cast<int>(ck_step.size()) - 1; step >= 0; --step) {
            const double h = ck_step[step];
            for (int s = 0; s < stages; ++s) axpby(h * b(s), gradient, 0.0, nullptr, kbar(s), D);
            for (int s = stages - 1; s >= 0; --s) {
                const size_t record = static_cast<size_t>(step) * stages + s;
                load_frozen(record);
                tau = a(s, s) * h;
                dg_refresh(ck_states.p + record * D);
                bj_refresh_transposed();
                axpby(1.0 / tau, kbar(s), 0.0, nullptr, adj_rhs.p, D);
                const Linear lin = gmres_transposed(adj_rhs.p, adj_lambda.p, cfg.max_linear, cfg.linear_tolerance);
                out.krylov += lin.iterations;
                out.solves += 1;
                out.max_relative = std::max(out.max_relative, lin.relative);
                HIPCHECK(launch_mass_combine(n, adj_lambda.p, kbar(s), tau, adj_base.p, D, basis.p, stream));
                axpby(1.0, gradient, 1.0, adj_base.p, gradient, D);
                for (int j = 0; j < s; ++j)
                    if (a(s, j) != 0) axpby(1.0, kbar(j), h * a(s, j), adj_base.p, kbar(j), D);
            }
        }
        sync();
        return out;
    }
    // density integral over the quadrant x > 0, y > 0 and its gradient
    double quadrant_density(const double *u, double *gradient) {
        adj_rhs.alloc(D, bytes);
        double *g = gradient ? gradient : adj_rhs.p;
        HIPCHECK(launch_quadrant_weights(n, g, basis.p, geo, stream));
        HIPCHECK(launch_dot(g, u, D, partials.p, stream));
        HIPCHECK(launch_finish_sum(partials.p, REDUCE_BLOCKS, 1, scalars.p + 200, stream));
        double out = 0;
        HIPCHECK(hipMemcpyAsync(&out, scalars.p + 200, sizeof(double), hipMemcpyDeviceToHost, stream));
        sync();
        return out;
    }
};

// ------------------------------------------------------------------ C ABI
namespace {
template <class F> int guarded(dgh_ctx *ctx, F &&f) {
    try {
        f();
        return DGH_OK;
    } catch (const std::exception &e) {
        if (ctx) ctx->error = e.what();
        else g_create_error = e.what();
        return DGH_ERROR;
    }
}
void check_ctx(dgh_ctx *ctx) { require(ctx != nullptr, "Null context"); }
}  // namespace

extern "C" {

size_t dgh_config_size(void) { return sizeof(dgh_config); }

int dgh_config_default(dgh_config *out) {
    if (!out) return DGH_ERROR;
    *out = default_config();
    return DGH_OK;
}

int dgh_config_parse(int argc, const char *const *argv, dgh_config *out, char *err, size_t errlen) {
    try {
        *out = parse_config(argc, argv);
        return DGH_OK;
    } catch (const std::exception &e) {
        if (err && errlen) {
            std::strncpy(err, e.what(), errlen - 1);
            err[errlen - 1] = 0;
        }
        g_create_error = e.what();
        return DGH_ERROR;
    }
}

int dgh_create(const dgh_config *cfg, dgh_ctx **out) {
    return guarded(nullptr, [&] {
        require(cfg != nullptr && out != nullptr, "Null argument");
        validate_config(*cfg);
        require(cfg->quadrature <= MAXQ, "Over-integration above 16 points per direction is not built on the device");
        *out = new dgh_ctx(*cfg);
    });
}

void dgh_destroy(dgh_ctx *ctx) { delete ctx; }

const char *dgh_last_error(const dgh_ctx *ctx) { return ctx ? ctx->error.c_str() : g_create_error.c_str(); }

int dgh_dims(const dgh_ctx *ctx, int64_t out[10]) {
    if (!ctx) return DGH_ERROR;
    out[0] = ctx->D;
    out[1] = ctx->T;
    out[2] = ctx->cells;
    out[3] = ctx->n;
    out[4] = ctx->faces;
    out[5] = ctx->m;
    out[6] = ctx->fw;
    out[7] = ctx->cfg.nx;
    out[8] = ctx->cfg.ny;
    out[9] = static_cast<int64_t>(ctx->bytes);
    return DGH_OK;
}

int dgh_geometry(const dgh_ctx *ctx, double out[5]) {
    if (!ctx) return DGH_ERROR;
    out[0] = ctx->geo.hx;
    out[1] = ctx->geo.hy;
    out[2] = ctx->geo.lx;
    out[3] = ctx->geo.ly;
    out[4] = ctx->tau;
    return DGH_OK;
}

int dgh_basis(const dgh_ctx *ctx, double *x, double *w) {
    if (!ctx) return DGH_ERROR;
    for (int i = 0; i < ctx->n; ++i) {
        if (x) x[i] = ctx->hbasis.x[i];
        if (w) w[i] = ctx->hbasis.w[i];
    }
    return DGH_OK;
}

int dgh_initial(dgh_ctx *ctx, double *u) {
    return guarded(ctx, [&] { check_ctx(ctx); ctx->initial(u); });
}
int dgh_seeded(dgh_ctx *ctx, uint64_t seed, double *u) {
    return guarded(ctx, [&] { check_ctx(ctx); ctx->seeded(u, seed); });
}
int dgh_freeze(dgh_ctx *ctx, const double *u, double *rate) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        const double r = ctx->freeze(u);
        if (rate) *rate = r;
    });
}
int dgh_set_step(dgh_ctx *ctx, double tau, const double *old) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        ctx->tau = tau;
        if (old) ctx->copy(ctx->old.p, old, ctx->D);
        ctx->refreshed = false;
        ctx->bj_ready = false;
        ctx->sync();
    });
}
int dgh_dg_residual(dgh_ctx *ctx, const double *u, double *r) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        ctx->residual_into(u, r);
        ctx->sync();
    });
}
int dgh_admissible(dgh_ctx *ctx, const double *u, int *ok) {
    return guarded(ctx, [&] { check_ctx(ctx); *ok = ctx->admissible(u) ? 1 : 0; });
}
int dgh_residual_norm(dgh_ctx *ctx, const double *r, double *rms) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        *rms = std::sqrt(ctx->reduce_sq(r, ctx->D, 0) / (ctx->geo.lx * ctx->geo.ly));
    });
}
int dgh_dual_norm(dgh_ctx *ctx, const double *x, double *out) {
    return guarded(ctx, [&] { check_ctx(ctx); *out = ctx->dual_norm(x); });
}
int dgh_state_norm(dgh_ctx *ctx, const double *x, double *out) {
    return guarded(ctx, [&] { check_ctx(ctx); *out = ctx->state_norm(x); });
}
int dgh_integral(dgh_ctx *ctx, const double *u, double out[4]) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        const auto v = ctx->integral(u);
        for (int k = 0; k < 4; ++k) out[k] = v[k];
    });
}
int dgh_error_l2(dgh_ctx *ctx, const double *u, double time, double out[4]) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        const auto v = ctx->error_l2(u, time);
        for (int k = 0; k < 4; ++k) out[k] = v[k];
    });
}
int dgh_speeds(dgh_ctx *ctx, double *out) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        ctx->copy(out, ctx->speeds.p, 2LL * ctx->faces * ctx->n);
        ctx->sync();
    });
}
int dgh_dg_refresh(dgh_ctx *ctx, const double *u) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        ctx->dg_refresh(u);
        ctx->sync();
    });
}
int dgh_dg_apply(dgh_ctx *ctx, const double *x, double *y) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        ctx->dg_apply(x, y);
        ctx->sync();
    });
}
int dgh_dg_apply_transpose(dgh_ctx *ctx, const double *w, double *z) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        ctx->dg_apply_transpose(w, z);
        ctx->sync();
    });
}
int dgh_dg_block(dgh_ctx *ctx, int row, double *out_host) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        require(ctx->refreshed, "Refresh the volume Jacobian before assembling a block");
        require(row >= 0 && row < ctx->cells, "Invalid block coordinates");
        const size_t mm = static_cast<size_t>(ctx->m) * ctx->m;
        DevBuf<double> one;
        one.alloc(mm, ctx->bytes);
        HIPCHECK(launch_assemble_blocks(ctx->n, ctx->vcache.p, ctx->fcache.p, one.p, ctx->inv_step, ctx->basis.p,
                                        ctx->geo, row, 1, ctx->stream));
        HIPCHECK(hipMemcpyAsync(out_host, one.p, mm * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
        one.release(ctx->bytes);
    });
}
int dgh_bj_refresh(dgh_ctx *ctx) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        ctx->bj_refresh();
        ctx->sync();
    });
}
int dgh_bj_apply(dgh_ctx *ctx, const double *r, double *z) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        ctx->bj_apply(r, z);
        ctx->sync();
    });
}
int dgh_ilu_refresh(dgh_ctx *ctx) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        ctx->ilu_refresh();
        ctx->sync();
    });
}
int dgh_ilu_apply(dgh_ctx *ctx, const double *r, double *z) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        ctx->ilu_apply(r, z);
        ctx->sync();
    });
}
int dgh_bj_refresh_mixed(dgh_ctx *ctx) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        ctx->bj_refresh_mixed();
        ctx->sync();
    });
}
int dgh_bj_apply_f32(dgh_ctx *ctx, const float *r, float *z) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        ctx->bj_apply32(r, z);
        ctx->sync();
    });
}
int dgh_dg_apply_f32(dgh_ctx *ctx, const float *x, float *y) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        ctx->dg_apply32(x, y);
        ctx->sync();
    });
}
int dgh_dg_fgmres(dgh_ctx *ctx, const double *rhs, double *x, int max_iterations, double tolerance, int *iterations,
                  int *outer, double *relative_residual) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        int o = 0;
        const auto r = ctx->fgmres(rhs, x, max_iterations, tolerance, o);
        ctx->sync();
        if (iterations) *iterations = r.iterations;
        if (outer) *outer = o;
        if (relative_residual) *relative_residual = r.relative;
    });
}
int dgh_dg_gmres(dgh_ctx *ctx, const double *rhs, double *x, int max_iterations, double tolerance, int *iterations,
                 double *relative_residual, double *residual) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        const auto r = ctx->gmres(rhs, x, max_iterations, tolerance);
        if (residual) ctx->copy(residual, ctx->kres.p, ctx->D);
        ctx->sync();
        if (iterations) *iterations = r.iterations;
        if (relative_residual) *relative_residual = r.relative;
    });
}
int dgh_hdg_refresh_projected(dgh_ctx *ctx, const double *u, const double *r) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        ctx->hdg_refresh_projected(u, r);
        ctx->sync();
    });
}
int dgh_hdg_rhs(dgh_ctx *ctx, double *out) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        require(ctx->hdg_ready, "Refresh the HDG condensation before applying it");
        ctx->copy(out, ctx->hrhs.p, ctx->T);
        ctx->sync();
    });
}
int dgh_hdg_apply(dgh_ctx *ctx, const double *x, double *y) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        ctx->hdg_apply(x, y);
        ctx->sync();
    });
}
int dgh_hdg_apply_noalloc(dgh_ctx *ctx, const double *x, double *y) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        ctx->hdg_apply_fixed(x, y, scratch_address(), SCRATCH_DOUBLES);
        ctx->sync();
    });
}
int dgh_hdg_block(dgh_ctx *ctx, int face, double *out_host) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        require(ctx->hdg_ready, "Refresh the HDG condensation before assembling a block");
        require(face >= 0 && face < ctx->faces, "Invalid block coordinates");
        {
            ctx->face_factors.alloc(static_cast<size_t>(ctx->faces) * ctx->fw * ctx->fw, ctx->bytes);
            HIPCHECK(launch_hdg_face_blocks(ctx->n, ctx->hS.p, ctx->fcache.p, ctx->face_factors.p, ctx->basis.p,
                                            ctx->geo, ctx->stream));
            ctx->face_bj_ready = false;
        }
        const size_t fw2 = static_cast<size_t>(ctx->fw) * ctx->fw;
        HIPCHECK(hipMemcpyAsync(out_host, ctx->face_factors.p + fw2 * face, fw2 * sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
    });
}
int dgh_hdg_bj_refresh(dgh_ctx *ctx) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        ctx->hdg_bj_refresh();
        ctx->sync();
    });
}
int dgh_hdg_bj_apply(dgh_ctx *ctx, const double *r, double *z) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        ctx->hdg_bj_apply(r, z);
        ctx->sync();
    });
}
int dgh_hdg_recover(dgh_ctx *ctx, const double *dlambda, double *du) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        ctx->hdg_recover(dlambda, du);
        ctx->sync();
    });
}
int dgh_hdg_recovered_residual(dgh_ctx *ctx, const double *schur_error, double *out) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        ctx->hdg_recovered_residual(schur_error, out);
        ctx->sync();
    });
}
int dgh_hdg_gmres(dgh_ctx *ctx, const double *rhs, double *x, int max_iterations, double tolerance, int *iterations,
                  double *relative_residual, double *residual) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        const auto r = ctx->hdg_gmres(rhs, x, max_iterations, tolerance);
        if (residual) ctx->copy(residual, ctx->kres.p, ctx->T);
        ctx->sync();
        if (iterations) *iterations = r.iterations;
        if (relative_residual) *relative_residual = r.relative;
    });
}
int dgh_advance(dgh_ctx *ctx, double *u, dgh_run_stats *run, dgh_update_stats *updates, int capacity, int *count) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        dgh_run_stats stats{};
        std::vector<dgh_update_stats> list;
        ctx->advance(u, stats, list);
        if (run) *run = stats;
        if (updates)
            for (int k = 0; k < capacity && k < static_cast<int>(list.size()); ++k) updates[k] = list[k];
        if (count) *count = static_cast<int>(list.size());
    });
}
int dgh_newton_update(dgh_ctx *ctx, double *u, dgh_update_stats *out) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        const auto &c = ctx->cfg;
        require(c.method != DGH_METHOD_HDG || c.precision == DGH_PRECISION_FP64,
                "Mixed precision is available for the DG method only");
        require(ctx->tau > 0, "Set the time step before a Newton update");
        ctx->newton_buffers();
        const auto start = Clock::now();
        ctx->copy(ctx->state.p, u, ctx->D);
        ctx->residual_into(ctx->state.p, ctx->residual.p);
        const double norm = ctx->dual_norm(ctx->residual.p), rms = norm / std::sqrt(ctx->geo.lx * ctx->geo.ly);
        const auto up = ctx->newton_update(norm, ctx->state_norm(ctx->old.p));
        ctx->copy(u, ctx->state.p, ctx->D);
        ctx->sync();
        if (out) {
            dgh_update_stats s{};
            s.krylov_iterations = s.inner_iterations = up.linear.iterations;
            s.outer_iterations = up.outer;
            s.dt = ctx->tau;
            s.residual_physical_rms = rms;
            s.damping = up.damping;
            s.linear_relative_residual = up.linear.relative;
            s.dg_linear_relative_residual = norm > 0 ? up.error_norm / norm : 0.;
            s.iteration_s = elapsed(start);
            s.precision_error = up.precision_error;
            *out = s;
        }
    });
}
int dgh_bj_refresh_transposed(dgh_ctx *ctx) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        ctx->bj_refresh_transposed();
        ctx->sync();
    });
}
int dgh_bj_apply_transposed(dgh_ctx *ctx, const double *r, double *z) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        ctx->bj_apply_transposed(r, z);
        ctx->sync();
    });
}
int dgh_dg_gmres_transposed(dgh_ctx *ctx, const double *rhs, double *x, int max_iterations, double tolerance,
                            int *iterations, double *relative_residual, double *residual) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        const auto r = ctx->gmres_transposed(rhs, x, max_iterations, tolerance);
        if (residual) ctx->copy(residual, ctx->kres.p, ctx->D);
        ctx->sync();
        if (iterations) *iterations = r.iterations;
        if (relative_residual) *relative_residual = r.relative;
    });
}
namespace {
int run_trajectory(dgh_ctx *ctx, dgh_ctx::Trajectory mode, double *u, dgh_run_stats *run, dgh_update_stats *updates,
                   int capacity, int *count) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        dgh_run_stats stats{};
        std::vector<dgh_update_stats> list;
        ctx->trajectory = mode;
        try {
            ctx->advance(u, stats, list);
        } catch (...) {
            ctx->trajectory = dgh_ctx::PLAIN;
            throw;
        }
        ctx->trajectory = dgh_ctx::PLAIN;
        if (run) *run = stats;
        if (updates)
            for (int k = 0; k < capacity && k < static_cast<int>(list.size()); ++k) updates[k] = list[k];
        if (count) *count = static_cast<int>(list.size());
    });
}
}  // namespace
int dgh_trajectory_record(dgh_ctx *ctx, double *u, dgh_run_stats *run, dgh_update_stats *updates, int capacity,
                          int *count) {
    return run_trajectory(ctx, dgh_ctx::RECORD, u, run, updates, capacity, count);
}
int dgh_trajectory_replay(dgh_ctx *ctx, double *u, dgh_run_stats *run, dgh_update_stats *updates, int capacity,
                          int *count) {
    return run_trajectory(ctx, dgh_ctx::REPLAY, u, run, updates, capacity, count);
}
int dgh_adjoint_backward(dgh_ctx *ctx, const double *final_gradient, double *initial_gradient,
                         dgh_adjoint_stats *stats) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        const auto r = ctx->adjoint_backward(final_gradient, initial_gradient);
        if (stats) {
            stats->krylov_iterations = r.krylov;
            stats->linear_solves = r.solves;
            stats->max_linear_residual = r.max_relative;
        }
    });
}
int dgh_quadrant_density(dgh_ctx *ctx, const double *u, double *value, double *gradient) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        const double v = ctx->quadrant_density(u, gradient);
        if (value) *value = v;
    });
}
int dgh_kernel_time(dgh_ctx *ctx, const char *name, double *total_ms, long long *launches) {
    return guarded(ctx, [&] {
        check_ctx(ctx);
        auto it = ctx->timers.find(name);
        *total_ms = it == ctx->timers.end() ? 0.0 : it->second.ms;
        *launches = it == ctx->timers.end() ? 0 : it->second.launches;
    });
}
int dgh_reset_kernel_times(dgh_ctx *ctx) {
    return guarded(ctx, [&] { check_ctx(ctx); ctx->timers.clear(); });
}
int dgh_set_timing(dgh_ctx *ctx, int enabled) {
    return guarded(ctx, [&] { check_ctx(ctx); ctx->timing = enabled != 0; });
}
}
