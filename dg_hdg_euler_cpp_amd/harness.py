"""Experiment drivers for the device executable, with the reference's protocols.

    python -m dg_hdg_euler_cpp_amd.harness check    --output DIR
    python -m dg_hdg_euler_cpp_amd.harness accuracy --output DIR [--grids 32 64 128 256]
    python -m dg_hdg_euler_cpp_amd.harness speed    --output DIR [--cells 256] [--degrees 8 7 ...]
    python -m dg_hdg_euler_cpp_amd.harness accuracy1d --output DIR [--grids 32 64 128 256 512]

Each runs `bin/dg_hdg_2d --device gpu` with the command lines of the
reference's drivers and applies their acceptance rules:
* check: code_test/2d/test_solver.py:13-23 -- 4x4, k=2, DIRK3 / BE / entropy
  wave, DG and HDG with --check-update; DG/HDG max nodal difference <= 1e-9,
  conservation drift <= 1e-10;
* accuracy: accuracy_test/2d/convergence.py:20-47 -- Q2, DIRK3, CFL 2, T 0.2,
  paired DG/HDG per grid; observed orders per component and the DG/HDG
  difference columns, written to convergence.csv;
* speed: speed_test/2d/benchmark.py:21-43 -- BE, CFL 5, four Newton updates
  with the first as warm-up, DG then HDG per degree; timings.csv and
  speedup.csv, with the checks of scripts/postprocess.py:190-195 (four
  updates, warm-up flags 0,1,1,1, the summary mean equals the mean of the
  three measured iteration times).
* accuracy1d: accuracy_test/1d/convergence.py:19-47 with `bin/dg_hdg_1d` --
  k=2, DIRK3, CFL 2, T 0.2, GMRES(20) + block Jacobi, DG and HDG paired per
  grid, log2 orders per component and the DG/HDG difference columns, written
  to convergence.csv.
A SUCCESS file is written when every run passed; any failure raises.
"""
from __future__ import annotations

import argparse
import csv
import math
import os
import subprocess
import sys
import time
from pathlib import Path

EXE = Path(__file__).resolve().parent / "bin" / "dg_hdg_2d"
EXE1D = Path(__file__).resolve().parent / "bin" / "dg_hdg_1d"


def rows(path):
    with Path(path).open(newline="") as stream:
        return list(csv.DictReader(stream))


def write_csv(path, data):
    if not data:
        return
    with Path(path).open("w", newline="") as stream:
        writer = csv.DictWriter(stream, fieldnames=list(data[0]))
        writer.writeheader()
        writer.writerows(data)


def run(command, prefix):
    """Run one solver invocation, log to PREFIX.log, return wall seconds."""
    command = [str(part) for part in command]
    start = time.perf_counter()
    with Path(f"{prefix}.log").open("w") as log:
        proc = subprocess.run(command, stdout=log, stderr=subprocess.STDOUT)
    if proc.returncode:
        raise RuntimeError(f"Command failed: {' '.join(command)}; see {prefix}.log")
    return {"wall_s": time.perf_counter() - start}


def require(ok, message):
    if not ok:
        raise RuntimeError(message)


def check(output: Path, exe: Path):
    for name, extra in [("dirk3", ["--T", .02]), ("be", ["--integrator", "be", "--T", .02]),
                        ("entropy", ["--testcase", "wave", "--T", .02])]:
        prefix = output / name
        run([exe, "--device", "gpu", "--nx", 4, "--ny", 4, "--k", 2, "--cfl", 2, "--method", "both",
             "--check-update", "--ntol", "1e-12", "--ltol", "1e-12", "--no-save-solution", *extra,
             "--prefix", prefix], prefix)
        diff = rows(f"{prefix}_difference.csv")[0]
        require(float(diff["max_nodal"]) <= 1e-9, f"{name}: DG/HDG discrepancy")
        for summary in rows(f"{prefix}_summary.csv"):
            require(float(summary["max_conservation_drift"]) <= 1e-10, f"{name}: conservation")
        print(f"{name} passed", flush=True)
    (output / "SUCCESS").write_text("All production-path 2D checks passed.\n")


def accuracy(output: Path, exe: Path, grids):
    require(all(n >= 3 for n in grids) and grids == sorted(set(grids)), "Grids must be distinct and increasing, N>=3")
    records, previous = [], {}
    for n in grids:
        prefix = output / f"n{n}"
        print(f"Start N={n}, Q2, DIRK3, CFL=2, T=0.2", flush=True)
        run([exe, "--device", "gpu", "--nx", n, "--ny", n, "--k", 2, "--integrator", "ssp-dirk3", "--cfl", 2,
             "--T", .2, "--method", "both", "--precond", "block-jacobi", "--ordering", "natural", "--ntol", "1e-12",
             "--ltol", "1e-12", "--restart", 20, "--max-linear", 1500, "--max-newton", 30, "--safety", 1.1,
             "--no-save-solution", "--prefix", prefix], prefix)
        difference = rows(f"{prefix}_difference.csv")[0]
        summaries = rows(f"{prefix}_summary.csv")
        require([r["method"] for r in summaries] == ["dg", "hdg"], "Missing paired results")
        for data in summaries:
            require(float(data["final_time"]) == .2 and data["integrator"] == "ssp-dirk3"
                    and int(data["fixed_newton_steps"]) == 0, "Incorrect convergence protocol")
            method = data["method"]
            for component in ["rho", "momentum_x", "momentum_y", "energy"]:
                error = float(data[f"{component}_L2"])
                require(math.isfinite(error) and error > 0, "Invalid error")
                key = (method, component)
                data[f"{component}_order"] = (math.log(previous[key][1] / error) / math.log(n / previous[key][0])
                                              if key in previous else "")
                previous[key] = (n, error)
                data[f"difference_{component}_L2"] = difference[f"{component}_L2"]
            data["difference_joint_L2"] = difference["joint_L2"]
            data["difference_max_nodal"] = difference["max_nodal"]
            records.append(data)
        write_csv(output / "convergence.csv", records)
        print(f"N={n} completed", flush=True)
    (output / "SUCCESS").write_text("All requested paired 2D grids completed.\n")
    return records


def accuracy1d(output: Path, exe: Path, grids):
    require(all(n >= 3 for n in grids) and grids == sorted(set(grids)), "Grids must be distinct and increasing, N>=3")
    records, previous = [], {}
    for cells in grids:
        prefix = output / f"n{cells}"
        print(f"Start N={cells}, k=2, DIRK3, CFL=2, T=0.2", flush=True)
        run([exe, "--cells", cells, "--k", 2, "--integrator", "ssp-dirk3", "--cfl", 2, "--T", .2, "--method", "both",
             "--precond", "block-jacobi", "--ordering", "natural", "--ntol", "1e-12", "--ltol", "1e-12",
             "--restart", 20, "--max-linear", 1500, "--max-newton", 30, "--safety", 1.1, "--prefix", prefix], prefix)
        difference = rows(f"{prefix}_difference.csv")[0]
        summaries = rows(f"{prefix}_summary.csv")
        require([r["method"] for r in summaries] == ["dg", "hdg"], "Missing paired solver results")
        for data in summaries:
            method = data["method"]
            require(float(data["final_time"]) == .2 and data["integrator"] == "ssp-dirk3"
                    and int(data["fixed_newton_steps"]) == 0, "Unexpected convergence protocol")
            for component in ("rho", "momentum", "energy"):
                error = float(data[f"{component}_L2"])
                require(math.isfinite(error) and error > 0, "Nonfinite/zero convergence error")
                key = method, component
                data[f"{component}_order"] = math.log2(previous[key] / error) if key in previous else ""
                previous[key] = error
                data[f"difference_{component}_L2"] = difference[f"{component}_L2"]
            data["difference_joint_L2"] = difference["joint_L2"]
            data["difference_max_nodal"] = difference["max_nodal"]
            records.append(data)
        write_csv(output / "convergence.csv", records)
        print(f"N={cells} completed", flush=True)
    (output / "SUCCESS").write_text("All requested paired 1D grids completed.\n")
    return records


def speed(output: Path, exe: Path, cells, degrees):
    require(cells >= 3 and len(set(degrees)) == len(degrees) and all(k in range(9) for k in degrees),
            "Invalid mesh or degree list")
    records, ratios = [], []
    for k in degrees:
        pair = {}
        for method in ["dg", "hdg"]:
            prefix = output / f"k{k}_{method}"
            print(f"Start {method}: N={cells}x{cells}, k={k}, exact block Jacobi", flush=True)
            resources = run([exe, "--device", "gpu", "--nx", cells, "--ny", cells, "--k", k, "--method", method,
                             "--integrator", "be", "--steps", 1, "--newton-steps", 4, "--warmup-updates", 1,
                             "--cfl", 5, "--precond", "block-jacobi", "--ordering", "natural", "--ntol", "1e-12",
                             "--ltol", "1e-12", "--restart", 20, "--max-linear", 1500, "--safety", 1.1,
                             "--timing-only", "--prefix", prefix], prefix)
            summary = rows(f"{prefix}_summary.csv")[0]
            updates = rows(f"{prefix}_iterations.csv")
            require(len(updates) == 4 and [r["measured"] for r in updates] == ["0", "1", "1", "1"]
                    and int(summary["measured_updates"]) == 3, "Expected exactly four updates, first warmup")
            require(int(summary["steps"]) == 1 and int(summary["newton_iterations"]) == 4
                    and int(summary["fixed_newton_steps"]) == 4 and summary["integrator"] == "be",
                    "Expected four BE updates")
            mean = sum(float(r["iteration_s"]) for r in updates[1:]) / 3
            require(math.isfinite(mean) and mean > 0 and math.isclose(mean, float(summary["mean_newton_s"]),
                                                                          rel_tol=1e-13), "Incorrect mean Newton time")
            summary.update(resources)
            summary["warmup_s"] = updates[0]["iteration_s"]
            records.append(summary)
            pair[method] = mean
            write_csv(output / "timings.csv", records)
            print(f"Completed {method} k={k}: mean={mean:.8g} s", flush=True)
        ratios.append(dict(k=k, nx=cells, ny=cells, preconditioner="block-jacobi", dg_jacobi_local="dense",
                           hdg_jacobi_local="face-dense", integrator="be", cfl=5, updates=4, warmup=1,
                           measured_updates=3, dg_mean_s=pair["dg"], hdg_mean_s=pair["hdg"],
                           dg_over_hdg=pair["dg"] / pair["hdg"]))
        write_csv(output / "speedup.csv", ratios)
    (output / "SUCCESS").write_text("All requested sequential 2D timing pairs completed.\n")
    return records, ratios


def main(argv=None):
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument("mode", choices=["check", "accuracy", "speed", "accuracy1d"])
    parser.add_argument("--output", required=True)
    parser.add_argument("--exe", default=None)
    parser.add_argument("--grids", type=int, nargs="+", default=None)
    parser.add_argument("--cells", type=int, default=256)
    parser.add_argument("--degrees", type=int, nargs="+", default=list(range(8, -1, -1)))
    args = parser.parse_args(argv)
    exe = Path(args.exe or (EXE1D if args.mode == "accuracy1d" else EXE))
    grids = args.grids or ([32, 64, 128, 256, 512] if args.mode == "accuracy1d" else [32, 64, 128, 256])
    if not exe.is_file():
        raise SystemExit(f"{exe} is missing: build it with `python -m dg_hdg_euler_cpp_amd.build`")
    output = Path(args.output)
    output.mkdir(parents=True, exist_ok=True)
    if args.mode == "check":
        check(output, exe)
    elif args.mode == "accuracy":
        accuracy(output, exe, grids)
    elif args.mode == "accuracy1d":
        accuracy1d(output, exe, grids)
    else:
        speed(output, exe, args.cells, args.degrees)
    return 0


if __name__ == "__main__":
    sys.exit(main())
