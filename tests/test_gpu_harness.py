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
