"""Newton-update throughput of the 2D DG k=8 256x256 implicit Euler solve on one MI355X.

A step is one damped Newton update of the BE stage (residual, Jacobian refresh,
block-Jacobi assembly + batched 324x324 LU, right-preconditioned GMRES(20) to
ltol=1e-12, line search with positivity check) -- the reference's speed protocol
(speed_test/2d/benchmark.py:26-36: BE, CFL 5, 1 step, fixed Newton updates, the
first W are warm-up). Inputs are resident in HBM before the timed region.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--init seeded|diagonal]

Prints one JSON line on stdout.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
GRAPH = False  # --graph: GMRES Arnoldi steps replay captured hipGraphs


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def solver_args(nx, k, init, precision, newton_steps, warmup, device=True, method="dg"):
    args = ["--nx", str(nx), "--ny", str(nx), "--k", str(k), "--method", method, "--integrator", "be",
            "--steps", "1", "--newton-steps", str(newton_steps), "--warmup-updates", str(warmup), "--cfl", "5",
            "--ntol", "1e-12", "--ltol", "1e-12", "--restart", "20", "--max-linear", "1500", "--safety", "1.1",
            "--timing-only", "--testcase", init]
    if device:
        args += ["--seed", "2025", "--precision", precision]
        if GRAPH:
            args += ["--graph"]
    return args


def run_hdg(args, nx, k, init):
    """Condensed HDG (FP64) at the same workload: Newton updates/s and Schur
    Krylov iterations (the reference's second method, config 3)."""
    from dg_hdg_euler_cpp_amd import DeviceContext

    with DeviceContext(solver_args(nx, k, init, "fp64", args.warmup + args.steps, args.warmup,
                                   method="hdg")) as dev:
        u = dev.initial()
        stats = dev.advance(u)
    measured = stats["measured_newton_seconds"]
    upd = stats["updates"]
    log(f"[bench] hdg fp64 {init}: updates={len(upd)} mean_s={measured / max(1, stats['measured_updates']):.4f}")
    for x in upd:
        log(f"   update {x['iteration']}: rms={x['residual_physical_rms']:.6e} krylov={x['krylov_iterations']} "
            f"t={x['iteration_s']:.4f}s")
    return {"updates_per_s": stats["measured_updates"] / measured if measured > 0 else None,
            "krylov_per_update": sum(x["krylov_iterations"] for x in upd[args.warmup:]) / max(1, len(upd) - args.warmup),
            "rms": [x["residual_physical_rms"] for x in upd]}


def run_ilu(args, nx, k, init, precision="fp64"):
    """Block ILU(0) in red-black order at the same workload (FP64, or FP32
    factors inside the mixed-precision solve): Newton updates/s, Krylov
    iterations, and the preconditioner's refresh and apply times (HIP events on
    the library's stream)."""
    from dg_hdg_euler_cpp_amd import DeviceContext

    ilu = ["--precond", "block-ilu0", "--ordering", "redblack"]
    with DeviceContext(solver_args(nx, k, init, precision, args.warmup + args.steps, args.warmup) + ilu) as dev:
        u = dev.initial()
        dev.set_timing(True)
        dev.reset_kernel_times()
        stats = dev.advance(u)
        times = {name: dev.kernel_time(name) for name in ["assemble", "lu", "ilu_coupling", "ilu_update",
                                                          "ilu_apply", "jv"]}
        dev.set_timing(False)
        u2 = dev.initial()
        stats_untimed = dev.advance(u2)
    measured = stats_untimed["measured_newton_seconds"]
    upd = stats_untimed["updates"]
    log(f"[bench] ilu0 {precision} {init}: updates={len(upd)} "
        f"mean_s={measured / max(1, stats_untimed['measured_updates']):.4f}")
    for x in upd:
        log(f"   update {x['iteration']}: rms={x['residual_physical_rms']:.6e} krylov={x['krylov_iterations']} "
            f"t={x['iteration_s']:.4f}s")
    per = {name: {"ms_total": ms, "launches": n, "avg_ms": ms / max(1, n)} for name, (ms, n) in times.items()}
    return {"updates_per_s": stats_untimed["measured_updates"] / measured if measured > 0 else None,
            "ms_per_update": 1e3 * measured / max(1, stats_untimed["measured_updates"]),
            "krylov_per_update": sum(x["krylov_iterations"] for x in upd[args.warmup:]) / max(1, len(upd) - args.warmup),
            "rms": [x["residual_physical_rms"] for x in upd], "timed_regions": per}


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


def run_precision_check(args, nx, k, init, precision):
    """Untimed: every update of the mixed form also solves its Newton system
    through the FP64 path and reports ||du_mixed - du_fp64||_M / ||du_fp64||_M."""
    from dg_hdg_euler_cpp_amd import DeviceContext

    with DeviceContext(solver_args(nx, k, init, precision, 2, 0) + ["--check-update"]) as dev:
        u = dev.initial()
        stats = dev.advance(u)
    out = [{"update": x["iteration"], "rms": x["residual_physical_rms"], "precision_error": x["precision_error"]}
           for x in stats["updates"]]
    log(f"[bench] {precision} vs fp64 update error: " + ", ".join(f"{o['precision_error']:.3e}" for o in out))
    return out


def run_device(args, nx, k, init, precision):
    import torch
    from dg_hdg_euler_cpp_amd import DeviceContext

    with DeviceContext(solver_args(nx, k, init, precision, args.warmup + args.steps, args.warmup)) as dev:
        u = dev.initial()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stats = dev.advance(u)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        measured = stats["measured_newton_seconds"]
        ups = stats["measured_updates"] / measured if measured > 0 else float("nan")
        upd = stats["updates"]
        log(f"[bench] {precision} {init}: updates={len(upd)} measured={stats['measured_updates']} "
            f"mean_s={measured / max(1, stats['measured_updates']):.4f} wall={wall:.2f}s")
        for x in upd:
            log(f"   update {x['iteration']}: rms={x['residual_physical_rms']:.6e} krylov={x['krylov_iterations']} "
                f"lin={x['linear_relative_residual']:.3e} damping={x['damping']} t={x['iteration_s']:.4f}s")
        # Dominant kernel: block-Jacobi apply, timed with HIP events on the
        # library's stream over a few launches on the final linearisation.
        mixed = precision != "fp64"
        vdtype = torch.float32 if mixed else torch.float64
        r = torch.randn(dev.size, dtype=vdtype, device="cuda")
        z = torch.zeros_like(r)
        apply = dev.bj_apply_f32 if mixed else dev.bj_apply
        apply(r, z)
        dev.set_timing(True)
        dev.reset_kernel_times()
        for _ in range(10):
            apply(r, z)
        bj_ms, bj_n = dev.kernel_time("bj_apply")
        dev.reset_kernel_times()
        x64 = z.to(torch.float64)
        y = dev.zeros()
        for _ in range(10):
            dev.dg_apply(x64, y)
        jv_ms, jv_n = dev.kernel_time("jv")
        dev.set_timing(False)
        geom = dict(size=dev.size, cells=dev.cells, m=dev.m, n=dev.n, faces=dev.faces)
    return dict(ups=ups, stats=stats, wall=wall, bj_avg_s=bj_ms / 1e3 / max(1, bj_n),
                jv_avg_s=jv_ms / 1e3 / max(1, jv_n), geom=geom)


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


def cpu_baseline(nx_full, k, init, krylov_per_update):
    """The oracle (1 core, FP64) on a 64x64 sub-problem of the same k=8 workload
    (about 10 s of CPU work), scaled per element to 256^2:
    t_update = (C/4096) * (t_refresh + n_K * t_iter + 2 * t_residual), with n_K
    the device run's mean Krylov count per update."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle_py as O

    nx = 64
    o = O.Oracle(solver_args(nx, k, init, "fp64", 1, 0, device=False))
    u = o.seeded(2025) if init == "seeded" else o.initial()
    rate = o.freeze(u)
    o.set_step(5.0 / rate, u)
    t0 = time.perf_counter()
    r = o.dg_residual(u)
    t_res = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.dg_refresh(u)
    o.bj_refresh()
    t_ref = time.perf_counter() - t0
    iters = 20
    t0 = time.perf_counter()
    try:
        _, its, _, _ = o.dg_gmres(-r, iters, iters, 1e-300)
    except O.OracleError:
        its = iters
    t_g = time.perf_counter() - t0
    t_iter = t_g / (its + 1)  # + the end-of-cycle M^-1 V y and true-residual Jv
    scale = (nx_full * nx_full) / (nx * nx)
    t_update = scale * (t_ref + krylov_per_update * t_iter + 2 * t_res)
    sample_s = t_res + t_ref + t_g
    log(f"[cpu] {nx}x{nx} k={k}: residual {t_res:.3f}s refresh+LU {t_ref:.3f}s gmres/iter {t_iter:.4f}s "
        f"-> {t_update:.1f} s/update at {nx_full}^2 (sample {sample_s:.1f}s)")
    return dict(value=1.0 / t_update, unit="Newton updates/s", cores=1, kind="port",
                sample=(f"oracle FP64 on 1 host core: residual, Jacobian refresh + {nx * nx} block LUs and {its} "
                        f"GMRES(20) iterations on a {nx}x{nx} k={k} mesh ({sample_s:.1f}s), scaled by "
                        f"{scale:.0f}x elements and the device run's {krylov_per_update:.0f} Krylov "
                        f"iterations per update"),
                seconds_per_update=t_update)


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
    with DeviceContext(solver_args(args.nx, args.k, args.init, "mixed", 1, 0)) as dev:
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
    with DeviceContext(solver_args(args.nx, args.k, args.init, "mixed-fp16", 1, 0)) as dev:
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
           "k_bj_apply_pairs<324": cells * m * m * 2 + cells * m * 4 + 2 * d * 4}
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
        key = short if grid == full[name] else f"{short} [one colour]"
        out["kernels"][key] = {"grid_threads": grid, "fetch_bytes": f, "write_bytes": w,
                               "hbm_bytes_per_launch": f + w, "algorithmic_bytes": a}
        extra = "" if a is None else f"{a / 1e9:15.3f} {(f + w) / a:6.3f}"
        print(f"  {(key + f' [{grid}]')[:72]:72s} {f / 1e9:10.3f} {w / 1e9:10.3f} {(f + w) / 1e9:10.3f} {extra}")
    with open(out_json, "w") as f:
        json.dump(out, f, indent=1)


def summarize_kernel_stats(stats_csv, top=14):
    """Per-kernel time table from a rocprofv3 --stats kernel_stats.csv
    (template instances of the reduction kernels summed)."""
    import collections
    import csv
    import re

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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--k", type=int, default=8)
    ap.add_argument("--init", default="diagonal", choices=["seeded", "diagonal"])
    ap.add_argument("--precision", default="both", choices=["both", "fp64", "mixed", "mixed-bf16", "mixed-fp16"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-hdg", action="store_true")
    ap.add_argument("--no-check", action="store_true", help="skip the untimed mixed-vs-FP64 update comparison")
    ap.add_argument("--graph", action="store_true", help="GMRES Arnoldi steps as captured hipGraphs")
    ap.add_argument("--only-hdg", action="store_true", help="run and report the condensed HDG form alone")
    ap.add_argument("--no-ilu", action="store_true", help="skip the block ILU(0) form")
    ap.add_argument("--no-1d", action="store_true", help="skip the 1D DG protocol run")
    ap.add_argument("--probe", action="store_true", help="run each hot kernel 3x for counter profiles")
    ap.add_argument("--counters", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "OUT_JSON"),
                    help="summarise two rocprofv3 counter runs of --probe into HBM bytes per launch")
    ap.add_argument("--kernel-stats", metavar="STATS_CSV", help="summarise a rocprofv3 kernel_stats.csv")
    args = ap.parse_args()
    if args.counters:
        summarize_counters(*args.counters)
        return
    if args.kernel_stats:
        summarize_kernel_stats(args.kernel_stats)
        return
    global GRAPH
    GRAPH = args.graph
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X; multi-GPU runs are not part of this build "
              "(the path would shard by 2D element tiles with halo exchanges of side traces)", file=sys.stderr)
        sys.exit(2)

    if args.probe:
        probe(args)
        return
    if args.only_hdg:
        hdg = run_hdg(args, args.nx, args.k, args.init)
        print(json.dumps({
            "metric": "2D HDG k=8 256² Newton updates/s (FP64)", "value": hdg["updates_per_s"],
            "unit": "Newton updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": 1e3 / hdg["updates_per_s"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": f"synthetic ({args.init} state)",
            "config": {"workload": f"2D condensed HDG BE Newton update, k={args.k}, {args.nx}x{args.nx}, CFL 5, "
                                   f"Schur GMRES(20) + face block Jacobi, ntol=ltol=1e-12"},
            "hdg_fp64": hdg}), flush=True)
        return
    # with both: FP64, mixed precision with FP32-stored factors, and the
    # headline form, mixed precision with FP16-stored factors
    forms = ["fp64", "mixed", "mixed-fp16"] if args.precision == "both" else [args.precision]
    runs = {p: run_device(args, args.nx, args.k, args.init, p) for p in forms}
    head = forms[-1]
    dev = runs[head]
    g = dev["geom"]

    def summary(p):
        d = runs[p]
        upd = d["stats"]["updates"][args.warmup:]
        krylov = sum(x["krylov_iterations"] for x in upd) / max(1, len(upd))
        b = bj_bytes(d["geom"], p)
        ach = b / d["bj_avg_s"] / 1e9
        return {"updates_per_s": d["ups"], "ms_per_update": 1e3 / d["ups"], "krylov_per_update": krylov,
                "rms": [x["residual_physical_rms"] for x in d["stats"]["updates"]],
                "linear_relative_residual": [x["linear_relative_residual"] for x in d["stats"]["updates"]],
                "bj_apply": {"avg_launch_s": d["bj_avg_s"], "algorithmic_bytes": b, "achieved_GBs": ach,
                             "frac_hbm": ach / HBM_PEAK_GBS}}

    forms_out = {p: summary(p) for p in forms}
    krylov = forms_out[head]["krylov_per_update"]
    bytes_bj = bj_bytes(g, head)
    achieved = bytes_bj / dev["bj_avg_s"] / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"bj_apply_traffic_{head}.json")
    # the counter runs were taken at 256^2, k=8 (bench.py --probe)
    if os.path.exists(tpath) and args.nx == 256 and args.k == 8:
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    equal = None
    if "fp64" in forms and len(forms) > 1:
        a = forms_out["fp64"]["rms"]
        equal = {"rms_fp64": a}
        for p in forms[1:]:
            b = forms_out[p]["rms"]
            equal[p] = {"rms": b,
                        "max_rel_diff_first_two_updates": max(abs(x - y) / abs(x) for x, y in zip(a[:2], b[:2])),
                        "speedup_over_fp64": forms_out[p]["updates_per_s"] / forms_out["fp64"]["updates_per_s"]}
    hdg = None if args.no_hdg or args.precision != "both" else run_hdg(args, args.nx, args.k, args.init)
    oned = None if args.no_1d or args.precision != "both" else run_1d()
    ilu = None if args.no_ilu or args.precision != "both" else run_ilu(args, args.nx, args.k, args.init)
    ilu_mixed = None if args.no_ilu or args.precision != "both" else run_ilu(args, args.nx, args.k, args.init,
                                                                              "mixed")
    check = None if args.no_check or head == "fp64" else run_precision_check(args, args.nx, args.k, args.init, head)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.nx, args.k, args.init, forms_out.get("fp64", forms_out[head])["krylov_per_update"])
    line = {
        "metric": "2D DG k=8 256² Newton updates/s: FP64 vs mixed-prec at equal residual",
        "value": dev["ups"],
        "unit": "Newton updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 / dev["ups"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f64+f32",
        "data": f"synthetic ({args.init} state{', seed 2025' if args.init == 'seeded' else ''})",
        "config": {"workload": f"2D DG BE Newton update, k={args.k}, {args.nx}x{args.nx} periodic quads, "
                               f"CFL 5, GMRES(20) + block-Jacobi LU, ntol=ltol=1e-12",
                   "nx": args.nx, "ny": args.nx, "k": args.k, "init": args.init, "precision": head,
                   "factor_storage": {"fp64": "f64", "mixed": "f32", "mixed-bf16": "bf16",
                                      "mixed-fp16": "f16, U scaled by 2^-e per block"}[head],
                   "krylov_per_update": krylov},
        "roofline": {"kernel": "bj_apply", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": bytes_bj, "avg_launch_s": dev["bj_avg_s"]},
        "forms": forms_out,
        "equal_residual": equal,
        "update_error_vs_fp64": check,
        "hdg_fp64": hdg,
        "ilu0_redblack_fp64": ilu,
        "ilu0_redblack_mixed": ilu_mixed,
        "oned": oned,
        "jv": {"dof_per_s": g["size"] / dev["jv_avg_s"], "achieved_GBs": jv_bytes(g) / dev["jv_avg_s"] / 1e9,
               "avg_launch_s": dev["jv_avg_s"]},
        "cpu_baseline": None if cpu is None else {k: v for k, v in cpu.items() if k != "seconds_per_update"},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
