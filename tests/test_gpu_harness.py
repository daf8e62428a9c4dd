"""The reference's experiment protocols through the device executable
(dg_hdg_euler_cpp_amd/harness.py): the production checks, a two-grid
accuracy sweep and a small speed pair, with the reference drivers'
acceptance rules and CSV schemas."""
import csv

import pytest

pytestmark = pytest.mark.gpu


def test_check_protocol(tmp_path, torch_cuda):
    from dg_hdg_euler_cpp_amd import harness

    harness.check(tmp_path, harness.EXE)
    assert (tmp_path / "SUCCESS").exists()


def test_accuracy_protocol(tmp_path, torch_cuda):
    from dg_hdg_euler_cpp_amd import harness

    records = harness.accuracy(tmp_path, harness.EXE, [32, 64])
    rows = list(csv.DictReader((tmp_path / "convergence.csv").open()))
    assert [r["method"] for r in rows] == ["dg", "hdg", "dg", "hdg"]
    # reference accuracy_test/.../convergence.csv:4-5 (N=64 rho order)
    for r in rows[2:]:
        assert abs(float(r["rho_order"]) - 2.6176117079127) <= 1e-8
        assert float(r["difference_max_nodal"]) <= 1e-9
    assert len(records) == 4


def test_speed_protocol(tmp_path, torch_cuda):
    from dg_hdg_euler_cpp_amd import harness

    records, ratios = harness.speed(tmp_path, harness.EXE, 32, [4, 2])
    assert [int(r["k"]) for r in ratios] == [4, 2]
    assert all(float(r["dg_mean_s"]) > 0 and float(r["hdg_mean_s"]) > 0 for r in ratios)
    assert (tmp_path / "speedup.csv").exists() and (tmp_path / "SUCCESS").exists()


def test_accuracy_1d_protocol(tmp_path, torch_cuda):
    """accuracy_test/1d/convergence.py:19-47 through bin/dg_hdg_1d: the paired
    DG and HDG rows of the reference's 1D table (convergence_slurm_7622710.csv:
    2-7) at 32, 64 and 128 cells, errors, observed orders and the DG/HDG
    difference columns."""
    import json
    import os

    from dg_hdg_euler_cpp_amd import harness

    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_runs.json")))
    ref = {int(r["cells"]): r for r in gold["convergence_1d"]["rows"] if r["method"] == "dg"}
    refh = {int(r["cells"]): r for r in gold["convergence_1d"]["rows"] if r["method"] == "hdg"}
    records = harness.accuracy1d(tmp_path, harness.EXE1D, [32, 64, 128])
    rows = list(csv.DictReader((tmp_path / "convergence.csv").open()))
    assert [(int(r["cells"]), r["method"]) for r in rows] == [(c, m) for c in (32, 64, 128) for m in ("dg", "hdg")]
    assert len(records) == 6
    for r in rows:
        want = (ref if r["method"] == "dg" else refh)[int(r["cells"])]
        assert int(r["newton_iterations"]) == int(want["newton_iterations"])
        for key in ["rho_L2", "momentum_L2", "energy_L2"]:
            assert abs(float(r[key]) - float(want[key])) <= 2e-9 * float(want[key])
        assert float(r["difference_max_nodal"]) <= 1e-13
    for r in rows[2:]:
        want = (ref if r["method"] == "dg" else refh)[int(r["cells"])]
        assert abs(float(r["rho_order"]) - float(want["rho_order"])) <= 1e-6
    iters = list(csv.DictReader((tmp_path / "n128_iterations.csv").open()))
    assert len(iters) == 2 * int(ref[128]["newton_iterations"])
    sol = [s for s in csv.DictReader((tmp_path / "n128_solution.csv").open()) if s["method"] == "dg"]
    assert len(sol) == 128 * 3
    # the nodal solution is within the L2 error scale of the exact wave
    assert max(abs(float(s["rho"]) - float(s["rho_exact"])) for s in sol) < 1e-4
