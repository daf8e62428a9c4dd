"""Newton-update throughput of the 2D DG k=8 256x256 implicit Euler solve on one MI355X.

A step is one damped Newton update of the BE stage: residual, Jacobian
refresh, block-Jacobi assembly + batched 324x324 LU, right-preconditioned
GMRES(20) to the reference's linear target, line search with the positivity
check (dgh_newton_update). The steps follow the reference's speed protocol
(speed_test/2d/benchmark.py:26-36: BE at CFL 5 from the `diagonal` state, four
Newton updates of one step, the first a warm-up, the mean of updates 1-3):
update 0 is run once, untimed; then every step is one of the updates 1, 2, 3,
and every third step starts again from the state update 0 produced, so each
timed update solves a system the reference's protocol solves. Inputs are
resident in HBM before the timed region.

    python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--precision both|fp64|mixed|mixed-fp16]

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
GRAPH = True  # GMRES Arnoldi steps replay captured hipGraphs (the library default); --no-graph: host loop
CYCLE = 3      # updates 1-3 of the reference protocol


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def solver_args(nx, k, init, precision, method="dg", extra=()):
    args = ["--nx", str(nx), "--ny", str(nx), "--k", str(k), "--method", method, "--integrator", "be",
            "--steps", "1", "--newton-steps", "4", "--warmup-updates", "1", "--cfl", "5", "--ntol", "1e-12",
            "--ltol", "1e-12", "--restart", "20", "--max-linear", "1500", "--safety", "1.1", "--timing-only",
            "--testcase", init, "--seed", "2025", "--precision", precision]
    if not GRAPH:
        args += ["--no-graph"]
    return args + list(extra)


class Protocol:
    """The reference speed protocol on one device context: freeze, the BE step
    size CFL 5 / rate, update 0 (untimed), then updates 1-3 repeated."""

    def __init__(self, dev):
        import torch

        self.dev, self.torch = dev, torch
        self.u0 = dev.initial()
        self.rate = dev.freeze(self.u0)
        dev.set_step(5.0 / self.rate, self.u0)
        self.u = self.u0.clone()
        self.first = dev.newton_update(self.u)  # update 0: the protocol's warm-up
        self.u1 = self.u.clone()

    def run(self, count):
        """count updates of the cycle 1, 2, 3, 1, 2, 3, ...; returns their stats."""
        out = []
        for i in range(count):
            if i % CYCLE == 0:
                self.u.copy_(self.u1)
            out.append(self.dev.newton_update(self.u))
        return out

    def timed(self, warmup, steps):
        self.run(warmup)
        self.torch.cuda.synchronize()
        t0 = time.perf_counter()
        stats = self.run(steps)
        self.torch.cuda.synchronize()
        return stats, time.perf_counter() - t0

    def kernel_times(self, groups):
        """One protocol cycle with HIP-event timing of the named kernel groups
        (each launch group bracketed by events on the library's stream)."""
        dev = self.dev
        dev.set_timing(True)
        dev.reset_kernel_times()
        self.run(CYCLE)
        out = {}
        for g in groups:
            ms, n = dev.kernel_time(g)
            if n:
                out[g] = {"launches": n, "avg_ms": ms / n, "total_ms_per_update": ms / CYCLE}
        dev.set_timing(False)
        return out


FACTOR_BYTES = {"fp64": 8, "mixed": 4, "mixed-bf16": 2, "mixed-fp16": 2}
VECTOR_BYTES = {"fp64": 8, "mixed": 4, "mixed-bf16": 4, "mixed-fp16": 4}


def bj_bytes(g, precision="fp64"):
    """Algorithmic bytes of one block-Jacobi apply: LU factors (m^2 per block),
    int32 row permutation, read r and write z."""
    return (g["cells"] * g["m"] * g["m"] * FACTOR_BYTES[precision] + g["cells"] * g["m"] * 4
            + 2 * g["size"] * VECTOR_BYTES[precision])


def jv_bytes(g):
    """Algorithmic bytes of one Jv: read x, the volume cache (u, v, H per node),
    the face cache (9 doubles per face node) and write y."""
    n = g["n"]
    return 2 * g["size"] * 8 + 3 * g["cells"] * n * n * 8 + 9 * g["faces"] * n * 8


def lu_flops(g):
    """(2/3) m^3 per element block."""
    return g["cells"] * 2.0 / 3.0 * g["m"] ** 3


def run_form(args, precision, steps, warmup, method="dg", extra=(), groups=()):
    """Protocol run of one form: `steps` timed updates after `warmup`, then one
    instrumented cycle for kernel times. Returns a summary dict."""
    from dg_hdg_euler_cpp_amd import DeviceContext

    label = f"{method} {precision}{' ' + ' '.join(extra) if extra else ''}"
    with DeviceContext(solver_args(args.nx, args.k, args.init, precision, method, extra)) as dev:
        p = Protocol(dev)
        stats, wall = p.timed(warmup, steps)
        kt = p.kernel_times(groups) if groups else {}
        geom = dict(size=dev.size, cells=dev.cells, m=dev.m, n=dev.n, faces=dev.faces)
        jv = None
        if precision == "fp64" and method == "dg":
            jv = time_jv(dev, geom)
    ups = steps / wall
    log(f"[bench] {label}: {steps} timed updates in {wall:.3f} s -> {ups:.4f} updates/s "
        f"(update 0: rms {p.first['residual_physical_rms']:.6e} krylov {p.first['krylov_iterations']})")
    for i, s in enumerate(stats):
        log(f"   step {i} (update {i % CYCLE + 1}): rms={s['residual_physical_rms']:.6e} "
            f"krylov={s['krylov_iterations']} damping={s['damping']} t={s['iteration_s']:.4f}s")
    cyc = stats[:CYCLE]
    return {"updates_per_s": ups, "ms_per_update": 1e3 * wall / steps, "timed_steps": steps,
            "krylov_per_update": sum(s["krylov_iterations"] for s in stats) / steps,
            "rms_before_updates_0_3": [p.first["residual_physical_rms"]] + [s["residual_physical_rms"] for s in cyc],
            "krylov_updates_0_3": [p.first["krylov_iterations"]] + [s["krylov_iterations"] for s in cyc],
            "linear_relative_residual": [s["linear_relative_residual"] for s in cyc],
            "kernel_times": kt, "geom": geom, "jv_fp64": jv}


def time_jv(dev, geom, launches=50):
    """FP64 Jv (dgh_dg_apply) on the protocol's last linearisation, HIP events."""
    import torch

    x = torch.randn(dev.size, dtype=torch.float64, device="cuda")
    y = dev.zeros()
    dev.dg_apply(x, y)
    dev.set_timing(True)
    dev.reset_kernel_times()
    for _ in range(launches):
        dev.dg_apply(x, y)
    ms, n = dev.kernel_time("jv")
    dev.set_timing(False)
    avg = ms / 1e3 / n
    return {"avg_launch_s": avg, "dof_per_s": geom["size"] / avg, "algorithmic_bytes": jv_bytes(geom),
            "achieved_GBs": jv_bytes(geom) / avg / 1e9, "frac_hbm": jv_bytes(geom) / avg / 1e9 / HBM_PEAK_GBS}


def run_precision_check(args, precision):
    """Untimed: updates 0 and 1 of the protocol also solve their Newton system
    through the FP64 path and report ||du_mixed - du_fp64||_M / ||du_fp64||_M."""
    from dg_hdg_euler_cpp_amd import DeviceContext

    with DeviceContext(solver_args(args.nx, args.k, args.init, precision, extra=["--check-update"])) as dev:
        p = Protocol(dev)
        second = p.run(1)[0]
    out = [{"update": i, "rms": s["residual_physical_rms"], "precision_error": s["precision_error"]}
           for i, s in enumerate([p.first, second])]
    log(f"[bench] {precision} vs fp64 update error: " + ", ".join(f"{o['precision_error']:.3e}" for o in out))
    return out


def run_1d():
    """The reference's 1D accuracy protocol at 128 cells (accuracy_test/1d/
    convergence.py:19-25: k=2, DIRK3, CFL 2, T 0.2, GMRES(20) + block Jacobi)
    on the device 1D path, DG and HDG, each one warm-up run then one timed run;
    beside them the reference's archived single-core times for the same runs
    (convergence_slurm_7622710.csv:6-7, a different CPU, not same-run figures)."""
    from dg_hdg_euler_cpp_amd import DeviceContext1D

    args = ["--cells", "128", "--k", "2", "--integrator", "ssp-dirk3", "--cfl", "2", "--T", "0.2",
            "--ntol", "1e-12", "--ltol", "1e-12", "--restart", "20", "--max-linear", "1500", "--max-newton", "30",
            "--safety", "1.1"]
    out = {"workload": "1D, 128 cells, k=2, DIRK3, CFL 2, T 0.2, GMRES(20) + block Jacobi"}
    published = {"dg": (0.45550448400000004, 0.0028767540316455686), "hdg": (0.116675362, 0.00073255362025316452)}
    for method in ("dg", "hdg"):
        with DeviceContext1D(args + ["--method", method]) as dev:
            dev.advance(dev.initial())
            u = dev.initial()
            st = dev.advance(u)
            err = dev.error_l2(u, st["final_time"])
        log(f"[bench] 1d {method} 128 cells k=2 DIRK3: {st['newton_iterations']} updates, "
            f"{st['krylov_iterations']} Krylov, {st['seconds']:.4f} s")
        out[method] = {"time_loop_s": st["seconds"], "newton_updates": st["newton_iterations"],
                       "krylov_iterations": st["krylov_iterations"],
                       "ms_per_update": 1e3 * st["seconds"] / st["newton_iterations"], "rho_L2": err[0],
                       "reference_time_loop_s_published": published[method][0],
                       "reference_mean_newton_s_published": published[method][1]}
    return out


# ------------------------------------------------------------ CPU baseline
def host_cores():
    """Cores this process may run on (the box's CPU share), at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_leg(lib, nx, k, init):
    """Child process: the oracle's DG BE update 0 of the speed protocol at
    nx x nx (residual, Jacobian refresh, every block LU, GMRES(20) with all its
    iterations, line search). Prints one JSON line."""
    if os.environ.get("DGH_CPU_PIN"):
        os.sched_setaffinity(0, {int(os.environ["DGH_CPU_PIN"])})
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as O

    O.select_library(lib)
    args = ["--nx", str(nx), "--ny", str(nx), "--k", str(k), "--testcase", init, "--integrator", "be", "--steps", "1",
            "--newton-steps", "1", "--cfl", "5", "--restart", "20", "--max-linear", "1500", "--ntol", "1e-12",
            "--ltol", "1e-12", "--safety", "1.1"]
    o = O.Oracle(args)
    u = o.initial()
    _, st = o.advance("dg", u)
    up = st["updates"][0]
    print(json.dumps({"iteration_s": up["iteration_s"], "krylov": up["krylov_iterations"],
                      "rms": up["residual_physical_rms"]}), flush=True)


def host_description():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": host_cores()}


def cpu_baseline(nx_full, k, init):
    """The oracle (FP64 CPU restatement of the reference's algorithm) timed on
    this host: one full Newton update (update 0 of the protocol) at a reduced
    mesh, scaled to nx_full^2 by the element count. Built here with -O3
    -march=native -fopenmp (oracle/Makefile `native`); falls back to the
    prebuilt x86-64-v3 builds if that compile fails. Two samples:
    1 thread pinned to one core at 32^2, and every available core at 64^2."""
    oracle = os.path.join(ROOT, "oracle")
    native = os.path.join(oracle, "libdgh_oracle_native.so")
    built = "native"
    t0 = time.perf_counter()
    try:
        subprocess.run(["make", "-s", "-B", "-j3", "-C", oracle, "native"], check=True, timeout=300,
                       stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)
        lib1 = lib_all = native
    except (subprocess.SubprocessError, OSError) as e:
        log(f"[cpu] native build failed ({e}); using the prebuilt x86-64-v3 oracle")
        built = "x86-64-v3"
        lib1, lib_all = os.path.join(oracle, "libdgh_oracle.so"), os.path.join(oracle, "libdgh_oracle_omp.so")
    log(f"[cpu] oracle build ({built}) {time.perf_counter() - t0:.1f} s")
    cores = host_cores()
    samples = {}
    for label, nx, threads, lib in (("1core", 32, 1, lib1), ("all_cores", 64, cores, lib_all)):
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="true")
        if threads == 1:
            env["DGH_CPU_PIN"] = str(sorted(os.sched_getaffinity(0))[0])
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-leg", lib, str(nx), str(k), init]
        t0 = time.perf_counter()
        out = subprocess.run(cmd, check=True, env=env, stdout=subprocess.PIPE, text=True, timeout=900).stdout
        r = json.loads(out.strip().splitlines()[-1])
        scale = (nx_full / nx) ** 2
        r.update(nx=nx, threads=threads, scale=scale, seconds_per_update_full=r["iteration_s"] * scale,
                 sample_wall_s=time.perf_counter() - t0)
        log(f"[cpu] {label}: {nx}x{nx} k={k} update 0 {r['iteration_s']:.2f} s, {r['krylov']} Krylov, "
            f"{threads} thread(s) -> {r['seconds_per_update_full']:.1f} s/update at {nx_full}^2")
        samples[label] = r
    return {"build": built, "host": host_description(), "samples": samples}

def probe(args):
    """Each hot kernel a few times on one 256^2 k=8 linearisation, beside a
    plain device copy of known size (the scale for hardware byte counters):
    the program that counter runs (rocprofv3 --pmc) profile."""
    import torch
    from dg_hdg_euler_cpp_amd import DeviceContext

    src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"[probe] copy 1 GiB read + 1 GiB write x3")
    del src, dst
    with DeviceContext(solver_args(args.nx, args.k, args.init, "mixed")) as dev:
        u = dev.initial()
        rate = dev.freeze(u)
        dev.set_step(5.0 / rate, u)
        dev.dg_refresh(u)
        x = torch.randn(dev.size, dtype=torch.float64, device="cuda")
        y = dev.zeros()
        for _ in range(3):
            dev.dg_apply(x, y)
        for _ in range(3):
            dev.dg_apply_transpose(x, y)
        log("[probe] jv x3, jtv x3")
        dev.bj_refresh()
        for _ in range(3):
            dev.bj_apply(x, y)
        log("[probe] bj_apply fp64 x3")
        dev.bj_refresh_mixed()
        x32 = x.to(torch.float32)
        y32 = torch.zeros_like(x32)
        for _ in range(3):
            dev.bj_apply_f32(x32, y32)
        torch.cuda.synchronize()
        log("[probe] bj_apply fp32 x3")
        dev.ilu_refresh()
        for _ in range(3):
            dev.ilu_apply(x, y)
        torch.cuda.synchronize()
        log("[probe] ilu_apply fp64 x3")
    with DeviceContext(solver_args(args.nx, args.k, args.init, "mixed-fp16")) as dev:
        u = dev.initial()
        rate = dev.freeze(u)
        dev.set_step(5.0 / rate, u)
        dev.dg_refresh(u)
        dev.bj_refresh_mixed()
        x32 = torch.randn(dev.size, dtype=torch.float32, device="cuda")
        y32 = torch.zeros_like(x32)
        for _ in range(3):
            dev.bj_apply_f32(x32, y32)
        torch.cuda.synchronize()
        log("[probe] bj_apply fp16 factors x3")
    with DeviceContext(solver_args(args.nx, args.k, args.init, "fp64", method="hdg")) as dev:
        u = dev.initial()
        rate = dev.freeze(u)
        dev.set_step(5.0 / rate, u)
        r = dev.dg_residual(u)
        dev.hdg_refresh_projected(u, r)
        t = dev.hdg_rhs()
        for _ in range(3):
            dev.hdg_apply(t)
        torch.cuda.synchronize()
        log("[probe] hdg condensation x1, schur apply x3")
    print(json.dumps({"probe": "done", "nx": args.nx, "k": args.k}), flush=True)


def summarize_counters(fetch_csv, write_csv, out_json):
    """HBM bytes per launch of every kernel from two rocprofv3 counter runs of
    `bench.py --probe` (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), scaled by
    what the same counters read for the probe's 1 GiB device-to-device copy.
    Prints a table and writes the per-kernel figures to out_json."""
    import collections
    import csv
    import re

    gib = float(1 << 30)

    def load(path, counter):
        # per (kernel, grid size): the block ILU(0) launches the block kernels
        # on one colour (half the grid) beside the full block-Jacobi launches.
        # The calibration copies are the probe's first three large copies (the
        # factor packing stages larger ones later).
        agg, copies = collections.defaultdict(list), []
        with open(path) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] != counter:
                    continue
                v = float(r["Counter_Value"]) * 1024.0
                agg[(r["Kernel_Name"], int(r["Grid_Size"]))].append(v)
                if r["Kernel_Name"] == "__amd_rocclr_copyBuffer" and v > 0.4 * gib:
                    copies.append((int(r["Dispatch_Id"]), v))
        return agg, [v for _, v in sorted(copies)[:3]]

    (fetch, copy_f), (write, copy_w) = load(fetch_csv, "FETCH_SIZE"), load(write_csv, "WRITE_SIZE")
    sf, sw = gib / (sum(copy_f) / len(copy_f)), gib / (sum(copy_w) / len(copy_w))
    cells, m = 256 * 256, 324
    d = cells * m
    alg = {"k_bj_apply<324, double": cells * m * m * 8 + cells * m * 4 + 2 * d * 8,
           "k_bj_apply<324, float": cells * m * m * 4 + cells * m * 4 + 2 * d * 4,
           "k_bj_apply_pairs<324": cells * m * m * 2 + cells * m * 4 + 2 * d * 4,
           "k_bj_apply_colpairs<324": cells * m * m * 2 + cells * m * 4 + 2 * d * 4}
    out = {"fetch_scale": sf, "write_scale": sw, "kernels": {}}
    print(f"# copy kernel (1 GiB read + 1 GiB write): FETCH_SIZE reads {1 / sf:.4f} GiB, WRITE_SIZE {1 / sw:.4f} GiB")
    print(f"# scale factors: fetch x{sf:.4f}, write x{sw:.4f}")
    print(f"# {'kernel [grid threads]':72s} {'fetch GB':>10s} {'write GB':>10s} {'total GB':>10s} "
          f"{'algorithmic GB':>15s} {'ratio':>6s}")
    full = {}
    for (name, grid) in fetch:
        if name.startswith(("void dgh::", "dgh::")):
            full[name] = max(full.get(name, 0), grid)
    for (name, grid) in fetch:
        if not name.startswith(("void dgh::", "dgh::")):
            continue
        f = sum(fetch[(name, grid)]) / len(fetch[(name, grid)]) * sf
        ws = write.get((name, grid), [0.0])
        w = sum(ws) / len(ws) * sw
        short = re.sub(r"^(void )?dgh::\(anonymous namespace\)::", "", name)
        short = short[:short.index("(")] if "(" in short else short
        a = next((v for k, v in alg.items() if short.startswith(k)), None) if grid == full[name] else None
        key = short if grid == full[name] else f"{short} [{'one colour' if 2 * grid == full[name] else 'batch'}]"
        out["kernels"][key] = {"grid_threads": grid, "fetch_bytes": f, "write_bytes": w,
                               "hbm_bytes_per_launch": f + w, "algorithmic_bytes": a}
        extra = "" if a is None else f"{a / 1e9:15.3f} {(f + w) / a:6.3f}"
        print(f"  {(key + f' [{grid}]')[:72]:72s} {f / 1e9:10.3f} {w / 1e9:10.3f} {(f + w) / 1e9:10.3f} {extra}")
    with open(out_json, "w") as f:
        json.dump(out, f, indent=1)


COMPUTE_COUNTERS = ["SQ_INSTS_VALU_FMA_F32", "SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_MFMA_MOPS_F32",
                    "SQ_INSTS_VALU_MFMA_MOPS_F64", "SQ_ACTIVE_INST_VALU", "SQ_VALU_MFMA_BUSY_CYCLES", "SQ_INSTS_LDS",
                    "SQ_INSTS_VALU", "GRBM_GUI_ACTIVE"]


def summarize_compute(counter_csv, trace_csv, kernels=("k_lu", "k_hdg_solve", "k_diag_inverses", "k_ilu_coupling",
                                                      "k_bj_apply", "k_assemble", "k_element_jv", "k_hdg_schur")):
    """Arithmetic of the dense kernels from one rocprofv3 --pmc run (the
    COMPUTE_COUNTERS above) with --kernel-trace durations of the same
    dispatches: vector FMA flops (FMA instructions x 64 lanes x 2; lanes masked
    off still count, so this is an upper bound), MFMA flops (MOPS x 512), the
    rate against the FP64 / FP32 peaks, VALU busy (rocprofv3's VALUBusy:
    SQ_ACTIVE_INST_VALU / CUs / GRBM_GUI_ACTIVE) and MFMA busy (MfmaUtil:
    SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE x SIMDs)). The counter file
    sums GRBM_GUI_ACTIVE over the 8 XCDs, so it is divided by 8 (the clock it
    then implies is printed as a check against the duration)."""
    import collections
    import csv
    import re

    cus, simds, xcds = 256, 1024, 8
    dur = {}
    with open(trace_csv) as f:
        for r in csv.DictReader(f):
            dur[int(r["Dispatch_Id"])] = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9
    per = collections.defaultdict(dict)
    names = {}
    with open(counter_csv) as f:
        for r in csv.DictReader(f):
            d = int(r["Dispatch_Id"])
            per[d][r["Counter_Name"]] = per[d].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
            names[d] = r["Kernel_Name"]
    agg = collections.defaultdict(lambda: collections.defaultdict(float))
    for d, c in per.items():
        short = re.sub(r"^(void )?dgh::\(anonymous namespace\)::", "", names[d])
        short = short[:short.index("(")] if "(" in short else short
        if not short.startswith(kernels) or d not in dur:
            continue
        a = agg[short]
        a["launches"] += 1
        a["seconds"] += dur[d]
        for k, v in c.items():
            a[k] += v
    print(f"# {'kernel':44s} {'calls':>5s} {'avg ms':>9s} {'FMA TFLOP/s':>11s} {'MFMA TFLOP/s':>12s} "
          f"{'of peak':>7s} {'VALU busy':>9s} {'MFMA busy':>9s} {'LDS/VALU':>8s} {'clock GHz':>9s}")
    for k, a in sorted(agg.items(), key=lambda kv: -kv[1]["seconds"]):
        t = a["seconds"]
        f64 = a["SQ_INSTS_VALU_FMA_F64"] * 128 + a["SQ_INSTS_VALU_MFMA_MOPS_F64"] * 512
        f32 = a["SQ_INSTS_VALU_FMA_F32"] * 128 + a["SQ_INSTS_VALU_MFMA_MOPS_F32"] * 512
        fma = (a["SQ_INSTS_VALU_FMA_F64"] + a["SQ_INSTS_VALU_FMA_F32"]) * 128 / t / 1e12
        mfma = (a["SQ_INSTS_VALU_MFMA_MOPS_F64"] + a["SQ_INSTS_VALU_MFMA_MOPS_F32"]) * 512 / t / 1e12
        peak = f64 / t / 1e12 / FP64_PEAK_TFLOPS if f64 >= f32 else f32 / t / 1e12 / FP32_PEAK_TFLOPS
        gui = a["GRBM_GUI_ACTIVE"] / xcds
        valu = 100 * a["SQ_ACTIVE_INST_VALU"] / cus / gui if gui else 0.0
        mf = 100 * a["SQ_VALU_MFMA_BUSY_CYCLES"] / (gui * simds) if gui else 0.0
        lds = a["SQ_INSTS_LDS"] / a["SQ_INSTS_VALU"] if a["SQ_INSTS_VALU"] else 0.0
        print(f"  {k[:44]:44s} {int(a['launches']):5d} {1e3 * t / a['launches']:9.3f} {fma:11.2f} {mfma:12.2f} "
              f"{peak:7.3f} {valu:8.1f}% {mf:8.1f}% {lds:8.3f} {gui / t / 1e9:9.2f}")


def summarize_kernel_stats(stats_csv, top=14):
    """Per-kernel time table from a rocprofv3 --stats kernel_stats.csv
    (template instances of the reduction kernels summed)."""
    import collections
    import csv
    import re

    if stats_csv.endswith(".db"):  # rocprofv3's default rocpd output: per-dispatch rows
        import sqlite3

        con = sqlite3.connect(stats_csv)
        per = collections.defaultdict(lambda: [0, 0.0])
        for name, dur in con.execute("select name, duration from kernels"):
            per[name][0] += 1
            per[name][1] += float(dur)
        rows = [{"Name": k, "Calls": c, "TotalDurationNs": t} for k, (c, t) in per.items()]
    else:
        with open(stats_csv) as f:
            rows = list(csv.DictReader(f))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    fam = collections.defaultdict(lambda: [0, 0.0])
    for r in rows:
        mt = re.match(r"(?:void )?(?:dgh::\(anonymous namespace\)::)?([A-Za-z_0-9:]+)(<[^(]*>)?", r["Name"])
        key = mt.group(1) if mt else r["Name"][:40]
        if mt and mt.group(2) and not key.startswith(("k_mdot", "k_mupdate")):
            key += mt.group(2)
        fam[key][0] += int(r["Calls"])
        fam[key][1] += float(r["TotalDurationNs"])
    print(f"# {'kernel (template instances of k_mdot / k_mupdate summed)':78s} {'calls':>6s} {'total ms':>10s} "
          f"{'avg us':>10s} {'share':>7s}")
    for k, (c, t) in sorted(fam.items(), key=lambda kv: -kv[1][1])[:top]:
        print(f"  {k[:78]:78s} {c:6d} {t / 1e6:10.2f} {t / 1e3 / c:10.1f} {100 * t / tot:6.2f}%")
    print(f"# total kernel time {tot / 1e6:.1f} ms")


def profile_figure(kernel_prefix, workload_tag):
    """Average duration (us) of a kernel in a committed rocprofv3 --stats summary
    (profiles/*<workload_tag>*.txt rows 'name calls total_ms avg_us share')."""
    import glob

    import re

    row = re.compile(r"^\s+(\S.*?)\s+(\d+)\s+([\d.]+)\s+([\d.]+)\s+([\d.]+)%\s*$")
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", f"{workload_tag}_kernel_stats.txt"))):
        with open(path) as f:
            for line in f:
                m = row.match(line)
                if m and m.group(1).startswith(kernel_prefix + "<"):
                    return {"avg_us": float(m.group(4)), "calls": int(m.group(2)), "source": os.path.relpath(path, ROOT)}
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--k", type=int, default=8)
    ap.add_argument("--init", default="diagonal", choices=["seeded", "diagonal"])
    ap.add_argument("--precision", default="both", choices=["both", "fp64", "mixed", "mixed-bf16", "mixed-fp16"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the HDG, ILU(0), 1D and precision-check runs")
    ap.add_argument("--no-graph", action="store_true", help="host-driven GMRES loop instead of captured hipGraphs")
    ap.add_argument("--only-hdg", action="store_true", help="run and report the condensed HDG form alone")
    ap.add_argument("--probe", action="store_true", help="run each hot kernel 3x for counter profiles")
    ap.add_argument("--counters", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "OUT_JSON"),
                    help="summarise two rocprofv3 counter runs of --probe into HBM bytes per launch")
    ap.add_argument("--kernel-stats", metavar="STATS_CSV", help="summarise a rocprofv3 kernel_stats.csv")
    ap.add_argument("--compute", nargs=2, metavar=("COUNTER_CSV", "TRACE_CSV"),
                    help="summarise a rocprofv3 --pmc run of the compute counters (with --kernel-trace)")
    ap.add_argument("--cpu-leg", nargs=4, metavar=("LIB", "NX", "K", "INIT"), help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_leg:
        lib, nx, k, init = args.cpu_leg
        cpu_leg(lib, int(nx), int(k), init)
        return
    if args.counters:
        summarize_counters(*args.counters)
        return
    if args.kernel_stats:
        summarize_kernel_stats(args.kernel_stats)
        return
    if args.compute:
        summarize_compute(*args.compute)
        return
    global GRAPH
    GRAPH = not args.no_graph
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X; multi-GPU runs are not part of this build "
              "(the path would shard by 2D element tiles with halo exchanges of side traces)", file=sys.stderr)
        sys.exit(2)
    if args.probe:
        probe(args)
        return
    if args.only_hdg:
        hdg = run_form(args, "fp64", args.steps, args.warmup, method="hdg",
                       groups=["hdg_condense", "schur", "face_bj_apply", "face_lu", "lu", "gmres_ortho"])
        hdg.pop("geom")
        print(json.dumps({
            "metric": "2D HDG k=8 256² Newton updates/s (FP64)", "value": hdg["updates_per_s"],
            "unit": "Newton updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": hdg["ms_per_update"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": f"synthetic ({args.init} state)",
            "config": {"workload": f"2D condensed HDG BE Newton update, k={args.k}, {args.nx}x{args.nx}, CFL 5, "
                                   f"Schur GMRES(20) + face block Jacobi, ntol=ltol=1e-12"},
            "hdg_fp64": hdg}), flush=True)
        return

    groups = ["bj_apply", "lu", "jv", "gmres_ortho", "assemble", "pack", "residual"]
    head = "mixed-fp16" if args.precision == "both" else args.precision
    forms = {head: run_form(args, head, args.steps, args.warmup, groups=groups)}
    if args.precision == "both":
        # the comparison forms run one protocol cycle each (updates 1-3)
        for p in ("fp64", "mixed"):
            forms[p] = run_form(args, p, CYCLE, 0, groups=groups)
    dev = forms[head]
    g = dev.pop("geom")
    for f in forms.values():
        f.pop("geom", None)
    extras = {}
    if not args.no_extras and args.precision == "both":
        extras["hdg_fp64"] = run_form(args, "fp64", CYCLE, 0, method="hdg",
                                      groups=["hdg_condense", "schur", "face_bj_apply", "face_lu", "gmres_ortho"])
        ilu = ["--precond", "block-ilu0", "--ordering", "redblack"]
        extras["ilu0_redblack_fp64"] = run_form(args, "fp64", CYCLE, 0, extra=ilu, groups=["ilu_apply", "lu",
                                                                                              "ilu_coupling"])
        extras["ilu0_redblack_mixed"] = run_form(args, "mixed", CYCLE, 0, extra=ilu, groups=["ilu_apply", "lu",
                                                                                                "ilu_coupling"])
        for e in extras.values():
            e.pop("geom", None)
        extras["update_error_vs_fp64"] = run_precision_check(args, head) if head != "fp64" else None
        extras["oned"] = run_1d()

    # equal residual: every form's RMS before updates 0-3 against the FP64
    # form's and the reference's archived 256^2 run (k8_dg_iterations.csv:2-5)
    equal = None
    if "fp64" in forms and len(forms) > 1:
        a = forms["fp64"]["rms_before_updates_0_3"]
        equal = {"rms_fp64": a}
        for p, f in forms.items():
            if p == "fp64":
                continue
            b = f["rms_before_updates_0_3"]
            equal[p] = {"rms": b, "max_rel_diff_updates_0_2": max(abs(x - y) / abs(x) for x, y in zip(a[:3], b[:3])),
                        "speedup_over_fp64": f["updates_per_s"] / forms["fp64"]["updates_per_s"]}
        gold = os.path.join(ROOT, "tests", "golden", "reference_runs.json")
        if args.nx == 256 and args.k == 8 and args.init == "diagonal" and os.path.exists(gold):
            with open(gold) as fh:
                ref = json.load(fh)["runs"]["k8_256_be"]["iterations"]
            equal["reference_rms_256_k8"] = [float(r["residual_physical_rms"]) for r in ref]
            equal["reference_krylov_256_k8"] = [int(r["krylov_iterations"]) for r in ref]

    # dominant kernel: the block-Jacobi apply of the head form, HIP events over
    # every launch of one protocol cycle (the launches the timed updates make)
    bj = dev["kernel_times"]["bj_apply"]
    avg_s = bj["avg_ms"] / 1e3
    bytes_bj = bj_bytes(g, head)
    achieved = bytes_bj / avg_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"bj_apply_traffic_{head}.json")
    if os.path.exists(tpath) and args.nx == 256 and args.k == 8:
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    prof = profile_figure("k_bj_apply_colpairs" if head in ("mixed-fp16", "mixed-bf16") else "k_bj_apply",
                          f"newton_update_{args.nx}_k{args.k}_{head.replace('-', '_')}")
    lu = dev["kernel_times"].get("lu")
    lu_rate = None
    if lu:
        lu_rate = {"avg_launch_s": lu["avg_ms"] / 1e3, "flops_per_refresh": lu_flops(g),
                   "GFLOPs": lu_flops(g) / (lu["avg_ms"] / 1e3) / 1e9,
                   "frac_peak": lu_flops(g) / (lu["avg_ms"] / 1e3) / 1e12 /
                   (FP64_PEAK_TFLOPS if head == "fp64" else FP32_PEAK_TFLOPS),
                   "factor_precision": "f64" if head == "fp64" else "f32"}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.nx, args.k, args.init)
    line = {
        "metric": "2D DG k=8 256² Newton updates/s: FP64 vs mixed-prec at equal residual",
        "value": dev["updates_per_s"],
        "unit": "Newton updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dev["ms_per_update"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f64+f32",
        "data": f"synthetic ({args.init} state{', seed 2025' if args.init == 'seeded' else ''})",
        "config": {"workload": f"2D DG BE Newton update, k={args.k}, {args.nx}x{args.nx} periodic quads, CFL 5, "
                               f"GMRES(20) + block-Jacobi LU, ntol=ltol=1e-12; steps cycle through updates 1-3 "
                               f"of the reference speed protocol",
                   "nx": args.nx, "ny": args.nx, "k": args.k, "init": args.init, "precision": head,
                   "factor_storage": {"fp64": "f64", "mixed": "f32", "mixed-bf16": "bf16",
                                      "mixed-fp16": "f16, U scaled by 2^-e per block"}[head],
                   "krylov_per_update": dev["krylov_per_update"]},
        "roofline": {"kernel": "bj_apply", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": bytes_bj, "avg_launch_s": avg_s, "launches_timed": bj["launches"],
                     "timing": "HIP events on the library stream over every apply of one protocol cycle",
                     "rocprof_avg_s": None if prof is None else prof["avg_us"] / 1e6,
                     "rocprof_frac": None if prof is None else bytes_bj / (prof["avg_us"] / 1e6) / 1e9 / HBM_PEAK_GBS,
                     "rocprof_source": None if prof is None else prof["source"]},
        "lu": lu_rate,
        "jv_fp64": forms.get("fp64", {}).get("jv_fp64"),
        "forms": forms,
        "equal_residual": equal,
        **extras,
    }
    if cpu is not None:
        one, allc = cpu["samples"]["1core"], cpu["samples"]["all_cores"]
        line["cpu_baseline"] = {
            "value": 1.0 / allc["seconds_per_update_full"], "unit": "Newton updates/s", "cores": allc["threads"],
            "kind": "port",
            "sample": (f"oracle FP64 ({cpu['build']} build, OpenMP) on {allc['threads']} host cores: one full BE "
                       f"Newton update (update 0, {allc['krylov']} GMRES(20) iterations) at {allc['nx']}x"
                       f"{allc['nx']} k={args.k} in {allc['iteration_s']:.1f} s, scaled by {allc['scale']:.0f}x "
                       f"elements to {args.nx}x{args.nx}"),
            "one_core": {"value": 1.0 / one["seconds_per_update_full"], "cores": 1,
                         "sample": (f"the same update at {one['nx']}x{one['nx']} on one pinned core in "
                                    f"{one['iteration_s']:.1f} s ({one['krylov']} iterations), scaled by "
                                    f"{one['scale']:.0f}x elements")},
            "host": cpu["host"],
            "reference_published_s_per_update": 2146.3353,
        }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
