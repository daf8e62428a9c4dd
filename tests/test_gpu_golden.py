"""Whole device Newton runs against the reference's archived runs
(tests/golden/reference_runs.json, copied from the reference's result CSVs).

The reference reproduces itself (rebuilt, older vs current version) to ~4e-16
on the initial RMS, ~1e-14 on the first Newton update and ~2e-8 on the second
(its own archived CSVs), so the tolerances grow with the update index.
"""
import json
import math
import os

import pytest

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_runs.json")))
RMS_TOL = [1e-13, 1e-10, 1e-6, 1e-2]


def run_device(args, method):
    from dg_hdg_euler_cpp_amd import DeviceContext

    print(f"\n[golden] {method} {' '.join(args)}", flush=True)
    with DeviceContext(args + ["--method", method]) as dev:
        u = dev.initial()
        stats = dev.advance(u)
        err = dev.error_l2(u, stats["final_time"])
    return stats, err


@pytest.mark.parametrize("key", ["be_4x4_k2", "dirk3_4x4_k2", "entropy_4x4_k2"])
def test_small_runs_match_reference(key, torch_cuda):
    run = GOLD["runs"][key]
    stats, err = run_device(run["args"], "dg")
    ref = [r for r in run["iterations"] if r["method"] == "dg"]
    assert len(stats["updates"]) == len(ref)
    for got, want in zip(stats["updates"], ref):
        it = int(want["iteration"])
        w = float(want["residual_physical_rms"])
        first = float([r for r in ref if r["stage"] == want["stage"] and r["iteration"] == "0"][0]["residual_physical_rms"])
        assert abs(got["residual_physical_rms"] - w) <= max(RMS_TOL[it] * w, 1e-12 * first), (key, it, got["residual_physical_rms"], w)
        assert abs(got["krylov_iterations"] - int(want["krylov_iterations"])) <= 1
        assert abs(got["cfl"] - float(want["cfl"])) <= 1e-13 * float(want["cfl"])
        assert got["linear_relative_residual"] <= 1e-12
    summ = [s for s in run["summary"] if s["method"] == "dg"][0]
    for v, k in enumerate(["rho_L2", "momentum_x_L2", "momentum_y_L2", "energy_L2"]):
        assert abs(err[v] - float(summ[k])) <= 1e-9 * float(summ[k]), (k, err[v], summ[k])
    assert stats["conservation_drift"] <= 2e-11


def test_k8_256_initial_residual_and_dt(torch_cuda):
    """RMS0 and dt of the 256^2 k=8 `diagonal` BE CFL=5 run (k8_dg_iterations.csv:2)."""
    from dg_hdg_euler_cpp_amd import DeviceContext

    run = GOLD["runs"]["k8_256_be"]
    want_rms = float(run["iterations"][0]["residual_physical_rms"])
    want_dt = float(run["iterations"][0]["dt"])
    with DeviceContext(["--nx", "256", "--ny", "256", "--k", "8", "--cfl", "5", "--integrator", "be"]) as dev:
        u = dev.initial()
        rate = dev.freeze(u)
        dt = 5.0 / rate
        dev.set_step(dt, u)
        rms = dev.residual_norm(dev.dg_residual(u))
    assert abs(dt - want_dt) <= 4e-16 * want_dt
    assert abs(rms - want_rms) <= 1e-13 * want_rms, (rms, want_rms)


def test_k8_256_first_newton_updates(torch_cuda):
    """Two BE Newton updates at 256^2, k=8, GMRES(20): RMS sequence and Krylov
    count of the reference's speed run (k8_dg_iterations.csv:2-3)."""
    run = GOLD["runs"]["k8_256_be"]
    args = list(run["args"])
    i = args.index("--newton-steps")
    args[i + 1] = "2"
    args[args.index("--warmup-updates") + 1] = "0"
    stats, _ = run_device([a for a in args if a not in ("--method", "dg")], "dg")
    ref = run["iterations"]
    for it, got in enumerate(stats["updates"]):
        w = float(ref[it]["residual_physical_rms"])
        assert abs(got["residual_physical_rms"] - w) <= RMS_TOL[it] * w, (it, got["residual_physical_rms"], w)
    assert abs(stats["updates"][0]["krylov_iterations"] - int(ref[0]["krylov_iterations"])) <= 3


CONV = GOLD["convergence"]


@pytest.mark.parametrize("method,nx", [("dg", 32), ("dg", 64), ("dg", 128), ("dg", 256), ("hdg", 32), ("hdg", 64)])
def test_convergence_study_matches_reference(method, nx, torch_cuda):
    """The reference's Q2 accuracy study (k=2, DIRK3, CFL 2, T 0.2, rotated
    simple wave; accuracy_test/.../convergence.csv:2-9): step and Newton
    counts, L2 errors to 1e-9 and the observed order between meshes."""
    row = [r for r in CONV["rows"] if r["method"] == method and int(r["nx"]) == nx][0]
    stats, err = run_device(["--nx", str(nx), "--ny", str(nx)] + [a for a in CONV["args"] if a != "both"
                                                                    and a != "--method"], method)
    assert int(stats["steps"]) == int(row["steps"])
    # a stage whose last Newton residual lands within rounding of the 1e-12
    # threshold may take one update more or less than the reference
    assert abs(int(stats["newton_iterations"]) - int(row["newton_iterations"])) <= 1
    for v, k in enumerate(["rho_L2", "momentum_x_L2", "momentum_y_L2", "energy_L2"]):
        assert abs(err[v] - float(row[k])) <= 1e-9 * float(row[k]), (k, err[v], row[k])
    assert stats["conservation_drift"] <= 1e-12
    if row["rho_order"]:
        coarse = [r for r in CONV["rows"] if r["method"] == method and int(r["nx"]) == nx // 2][0]
        order = math.log(float(coarse["rho_L2"]) / err[0]) / math.log(2.0)
        print(f"{method} {nx}: rho L2 {err[0]:.6e}, order {order:.4f} (reference {row['rho_order']})")
        assert abs(order - float(row["rho_order"])) <= 1e-8
        assert order > 2.6   # Q2 elements: third order in the limit
