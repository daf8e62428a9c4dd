"""Host logic of the experiment protocols (dg_hdg_euler_cpp_amd/harness.py) on
the CPU: a stand-in solver script writes the reference's CSV files from its
flags, so the case tables, acceptance rules and reports are exercised without
a device. The device runs of the same protocols are in test_gpu_harness.py."""
import csv
import math
import os
import stat
import sys
import textwrap

import pytest

from dg_hdg_euler_cpp_amd import harness

FAKE = textwrap.dedent('''\
    import csv, os, sys
    a = sys.argv[1:]
    flag = lambda name, default=None: a[a.index(name) + 1] if name in a else default
    prefix = flag("--prefix")
    n = int(flag("--nx", flag("--cells", "4")))
    k = int(flag("--k", "2"))
    bad = os.environ.get("FAKE_BREAK", "")
    methods = ["dg", "hdg"] if flag("--method", "dg") == "both" else [flag("--method")]
    fixed = int(flag("--newton-steps", "0"))
    comps = ["rho", "momentum", "energy"] if "--cells" in a else ["rho", "momentum_x", "momentum_y", "energy"]
    def write(path, rows):
        with open(path, "w", newline="") as f:
            w = csv.DictWriter(f, fieldnames=list(rows[0]))
            w.writeheader(); w.writerows(rows)
    rows = []
    for m in methods:
        r = dict(method=m, integrator=flag("--integrator", "ssp-dirk3"), final_time=flag("--T", "0.2"),
                 fixed_newton_steps=fixed, steps=1, newton_iterations=fixed or 7, measured_updates=3 if fixed else 0,
                 mean_newton_s=0.2, max_conservation_drift=1e-12 if bad != "drift" else 1e-6, k=k)
        for c in comps:
            r[c + "_L2"] = 0.5 * (2.0 / n) ** 3
        rows.append(r)
    write(prefix + "_summary.csv", rows)
    its = [dict(update=i, measured=int(i > 0), iteration_s=[0.9, 0.1, 0.2, 0.3][i]) for i in range(4)]
    write(prefix + "_iterations.csv", its)
    d = dict(max_nodal=1e-14 if bad != "nodal" else 1e-3, joint_L2=1e-15)
    for c in comps:
        d[c + "_L2"] = 1e-15
    write(prefix + "_difference.csv", [d])
    print("fake solver", " ".join(a))
    ''')


@pytest.fixture
def fake_exe(tmp_path):
    p = tmp_path / "fake_solver.py"
    p.write_text("#!" + sys.executable + "\n" + FAKE)
    p.chmod(p.stat().st_mode | stat.S_IEXEC)
    return p


def _read(path):
    return list(csv.DictReader(path.open()))


def test_check_protocol_passes_and_fails(tmp_path, fake_exe, monkeypatch):
    harness.check(tmp_path / "ok", fake_exe)
    assert (tmp_path / "ok" / "SUCCESS").exists()
    for name in ("dirk3", "be", "entropy"):
        assert (tmp_path / "ok" / f"{name}.log").exists()
    monkeypatch.setenv("FAKE_BREAK", "nodal")
    with pytest.raises(harness.ProtocolFailure, match="agree to 1e-9"):
        harness.check(tmp_path / "bad", fake_exe)
    assert not (tmp_path / "bad" / "SUCCESS").exists()
    monkeypatch.setenv("FAKE_BREAK", "drift")
    with pytest.raises(harness.ProtocolFailure, match="conservation"):
        harness.check(tmp_path / "bad2", fake_exe)


def test_accuracy_orders_and_schema(tmp_path, fake_exe):
    rows = harness.accuracy(tmp_path, fake_exe, [8, 16, 32])
    disk = _read(tmp_path / "convergence.csv")
    assert [r["method"] for r in disk] == ["dg", "hdg"] * 3
    assert len(rows) == 6
    # the stand-in's errors fall as h^3: every order after the first grid is 3
    for r in disk[2:]:
        for c in ("rho", "momentum_x", "momentum_y", "energy"):
            assert abs(float(r[c + "_order"]) - 3.0) < 1e-12
    assert disk[0]["rho_order"] == ""
    for col in ("difference_rho_L2", "difference_joint_L2", "difference_max_nodal"):
        assert col in disk[0]
    with pytest.raises(harness.ProtocolFailure, match="increasing"):
        harness.accuracy(tmp_path / "g", fake_exe, [16, 8])


def test_accuracy1d_uses_1d_components(tmp_path, fake_exe):
    harness.accuracy1d(tmp_path, fake_exe, [16, 32])
    disk = _read(tmp_path / "convergence.csv")
    assert "momentum_order" in disk[0] and "momentum_x_order" not in disk[0]
    assert abs(float(disk[2]["energy_order"]) - 3.0) < 1e-12


def test_speed_tables(tmp_path, fake_exe):
    timings, speedup = harness.speed(tmp_path, fake_exe, cells=8, degrees=[3, 1])
    assert [(int(r["k"]), r["method"]) for r in timings] == [(3, "dg"), (3, "hdg"), (1, "dg"), (1, "hdg")]
    assert [int(r["k"]) for r in speedup] == [3, 1]
    for r in _read(tmp_path / "timings.csv"):
        assert float(r["warmup_s"]) == 0.9 and float(r["wall_s"]) > 0 and "peak_rss_kib" in r
    s = _read(tmp_path / "speedup.csv")[0]
    assert math.isclose(float(s["dg_over_hdg"]), 1.0) and int(s["measured_updates"]) == 3
    with pytest.raises(harness.ProtocolFailure, match="degrees"):
        harness.speed(tmp_path / "x", fake_exe, cells=8, degrees=[9])


def test_failed_solver_names_the_case(tmp_path):
    p = tmp_path / "failing.py"
    p.write_text("#!" + sys.executable + "\nimport sys\nprint('ERROR: Invalid initial solution')\nsys.exit(1)\n")
    p.chmod(p.stat().st_mode | stat.S_IEXEC)
    with pytest.raises(harness.ProtocolFailure, match="dirk3: solver exited with status 1"):
        harness.check(tmp_path / "out", p)
    assert "Invalid initial solution" in (tmp_path / "out" / "dirk3.log").read_text()
