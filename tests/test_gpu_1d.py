"""The 1D DG device path (include/dgh/dgh1d_device.h) against the oracle's 1D
restatement (dg_hdg_1d/src), which tests/test_oracle_1d.py pins against the
reference's archived convergence table:

* element operations on the same inputs: projected initial state, frozen
  speeds and rate, residual, Jacobian apply, element blocks, block-Jacobi
  apply, admissibility, integrals and L2 errors;
* GMRES: the reference's true-residual acceptance and the oracle's iteration
  count;
* whole BE / DIRK3 runs update by update, and the convergence protocol
  (accuracy_test/1d/convergence.py:19-25) against the reference's table.
"""
import json
import os

import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_runs.json")))["convergence_1d"]

CASES = [("k2-32", ["--cells", "32", "--k", "2"]), ("k4-q7-5", ["--cells", "5", "--k", "4", "--quad", "7"]),
         ("k0-6", ["--cells", "6", "--k", "0"]), ("k8-4", ["--cells", "4", "--k", "8"]),
         ("k3-600", ["--cells", "600", "--k", "3"])]


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("name,args", CASES, ids=[c[0] for c in CASES])
def test_1d_operators_match_oracle(name, args, torch_cuda):
    from dg_hdg_euler_cpp_amd import DeviceContext1D

    torch = torch_cuda
    print(f"\n[case] 1d operators {args}", flush=True)
    orc = O.Oracle1D(args)
    with DeviceContext1D(args) as dev:
        assert dev.size == orc.size
        u_o = orc.initial()
        u_d = dev.initial()
        assert rel(u_d.cpu().numpy(), u_o) <= 1e-13
        u = torch.tensor(u_o, device="cuda")
        rate_o = orc.freeze(u_o)
        rate_d = dev.freeze(u)
        assert abs(rate_d - rate_o) <= 1e-14 * rate_o
        assert rel(np.array(dev.speeds).reshape(-1, 2), orc.speeds) <= 1e-14
        tau = 2.0 / rate_o
        old = u_o * (1 + 1e-3 * np.sin(np.arange(u_o.size)))
        orc.set_step(tau, old)
        dev.set_step(tau, torch.tensor(old, device="cuda"))
        r_o = orc.residual(u_o)
        r_d = dev.residual(u).cpu().numpy()
        # On fine meshes the flux terms (size |u|) cancel to a residual a few
        # hundred times smaller, so rounding differences of the summation
        # order show relative to |r|; they stay a few ulps of the terms.
        scale = 64 * np.finfo(np.float64).eps * np.abs(u_o).max()
        print(f"residual rel {rel(r_d, r_o):.3e}, max diff {np.abs(r_d - r_o).max():.3e} (term ulps {scale:.1e})")
        assert rel(r_d, r_o) <= 1e-11 and np.abs(r_d - r_o).max() <= scale
        assert dev.admissible(u) and orc.admissible(u_o)
        bad = u_o.copy()
        bad[1] = -1.0
        assert not dev.admissible(torch.tensor(bad, device="cuda"))
        assert np.allclose(dev.integral(u), orc.integral(u_o), rtol=1e-13, atol=1e-15)
        assert np.allclose(dev.error_l2(u, 0.01), orc.error_l2(u_o, 0.01), rtol=1e-10, atol=1e-15)
        orc.refresh(u_o)
        dev.refresh(u)
        x = np.random.default_rng(11).standard_normal(u_o.size)
        y_o = orc.apply(x)
        y_d = dev.apply(torch.tensor(x, device="cuda")).cpu().numpy()
        print(f"Jv rel {rel(y_d, y_o):.3e}")
        assert rel(y_d, y_o) <= 1e-12
        for row in [0, orc.cells // 2, orc.cells - 1]:
            assert rel(dev.block(row), orc.block(row, row)) <= 1e-13, row
        orc.bj_refresh()
        dev.bj_refresh()
        z_o = orc.bj_apply(x)
        z_d = dev.bj_apply(torch.tensor(x, device="cuda")).cpu().numpy()
        print(f"block Jacobi rel {rel(z_d, z_o):.3e}")
        assert rel(z_d, z_o) <= 1e-11
        rhs = -r_o
        x_o, its_o, rel_o = orc.gmres(rhs, 30, 1500, 1e-12)
        x_d, its_d, rel_d = dev.gmres(torch.tensor(rhs, device="cuda"), 1500, 1e-12)
        print(f"GMRES its device {its_d}, oracle {its_o}")
        assert rel_d <= 1e-12 and abs(its_d - its_o) <= max(1, its_o // 20)
        res = orc.apply(x_d.cpu().numpy()) - rhs
        w = 1.0 / np.tile(0.5 * orc.dx * O.gauss(orc.n)[1], 3 * orc.cells)
        assert np.sqrt(np.sum(res ** 2 * w) / np.sum(rhs ** 2 * w)) <= 2e-12


RUNS = [("be-k2-16", ["--cells", "16", "--k", "2", "--integrator", "be", "--T", "0.05"]),
        ("dirk3-k3-q5-12", ["--cells", "12", "--k", "3", "--quad", "5", "--T", "0.05"]),
        ("dirk3-k5-16-steps", ["--cells", "16", "--k", "5", "--steps", "2"]),
        ("density-dirk3-k2-32", ["--cells", "32", "--k", "2", "--T", "0.2", "--testcase", "density"])]


@pytest.mark.parametrize("name,args", RUNS, ids=[r[0] for r in RUNS])
def test_1d_runs_match_oracle(name, args, torch_cuda):
    from dg_hdg_euler_cpp_amd import DeviceContext1D

    full = args + ["--cfl", "2", "--ntol", "1e-12", "--ltol", "1e-12", "--restart", "20"]
    print(f"\n[case] 1d run {full}", flush=True)
    orc = O.Oracle1D(full)
    uN_o, st_o = orc.advance(orc.initial())
    with DeviceContext1D(full) as dev:
        u = dev.initial()
        st = dev.advance(u)
        uN = u.cpu().numpy()
    assert len(st["updates"]) == len(st_o["updates"])
    diffs = []
    for got, want in zip(st["updates"], st_o["updates"]):
        w = want["residual_physical_rms"]
        assert abs(got["residual_physical_rms"] - w) <= max(1e-8 * w, 1e-13), (got, want)
        diffs.append(int(got["krylov_iterations"] - want["krylov_iterations"]))
    # the reductions sum in another order than the oracle's, so a solve that
    # ends near the ltol threshold can take one or two steps more or fewer
    print(f"Krylov differences per update {diffs}; final state rel {rel(uN, uN_o):.3e}")
    assert max(abs(d) for d in diffs) <= 2 and abs(sum(diffs)) <= max(2, st_o["krylov_iterations"] // 100)
    assert rel(uN, uN_o) <= 1e-11
    assert st["conservation_drift"] <= 1e-14
    # with --steps the final time sums CFL / rate of each step's state
    assert st["steps"] == st_o["steps"] and abs(st["final_time"] - st_o["final_time"]) <= 1e-14 * st_o["final_time"]


@pytest.mark.parametrize("cells", [32, 64, 128])
def test_1d_convergence_protocol_matches_reference(cells, torch_cuda):
    """accuracy_test/1d/convergence.py:19-25 (k=2, DIRK3, CFL 2, T 0.2,
    GMRES(20), block Jacobi) against the reference's archived DG rows."""
    from dg_hdg_euler_cpp_amd import DeviceContext1D

    ref = [r for r in GOLD["rows"] if r["method"] == "dg" and int(r["cells"]) == cells][0]
    args = ["--cells", str(cells)] + GOLD["args"]
    print(f"\n[case] 1d convergence {args}", flush=True)
    with DeviceContext1D(args) as dev:
        u = dev.initial()
        st = dev.advance(u)
        err = dev.error_l2(u, st["final_time"])
    print(f"steps {st['steps']} Newton {st['newton_iterations']} Krylov {st['krylov_iterations']} "
          f"(reference {ref['steps']}, {ref['newton_iterations']}, {ref['krylov_iterations']}), "
          f"{st['seconds']:.3f} s", flush=True)
    assert st["steps"] == int(ref["steps"]) and st["final_time"] == 0.2
    assert st["newton_iterations"] == int(ref["newton_iterations"])
    assert abs(st["krylov_iterations"] - int(ref["krylov_iterations"])) <= 2
    assert st["max_linear_residual"] <= 1e-12
    assert st["conservation_drift"] <= 1e-14
    for e, key in zip(err, ["rho_L2", "momentum_L2", "energy_L2"]):
        assert abs(e - float(ref[key])) <= 2e-9 * float(ref[key]), (key, e, ref[key])


HDG_CASES = [("k2-32", ["--cells", "32", "--k", "2"]), ("k3-q5-7", ["--cells", "7", "--k", "3", "--quad", "5"]),
             ("k0-6", ["--cells", "6", "--k", "0"]), ("k8-5", ["--cells", "5", "--k", "8"])]


@pytest.mark.parametrize("name,args", HDG_CASES, ids=[c[0] for c in HDG_CASES])
def test_1d_hdg_operators_match_oracle(name, args, torch_cuda):
    """hybrid.cpp (1D) on the same state and DG residual: Schur right-hand
    side, apply, face blocks, face block Jacobi, recovery, recovered residual
    and the Schur GMRES."""
    from dg_hdg_euler_cpp_amd import DeviceContext1D

    torch = torch_cuda
    print(f"\n[case] 1d hdg operators {args}", flush=True)
    orc = O.Oracle1D(args)
    with DeviceContext1D(args) as dev:
        u_o = orc.initial()
        u = torch.tensor(u_o, device="cuda")
        rate = orc.freeze(u_o)
        dev.freeze(u)
        tau = 2.0 / rate
        old = u_o * (1 + 1e-3 * np.sin(np.arange(u_o.size)))
        orc.set_step(tau, old)
        dev.set_step(tau, torch.tensor(old, device="cuda"))
        r_o = orc.residual(u_o)
        r = torch.tensor(r_o, device="cuda")
        orc.refresh(u_o)
        orc.hdg_refresh(u_o, r_o)
        dev.hdg_refresh(u, r)
        tol = 1e-12 if name != "k8-5" else 1e-10  # k=8 element blocks: kappa ~1e4
        g_o = orc.hdg_rhs()
        print(f"rhs rel {rel(dev.hdg_rhs().cpu().numpy(), g_o):.3e}")
        assert rel(dev.hdg_rhs().cpu().numpy(), g_o) <= tol
        dl = np.random.default_rng(4).standard_normal(3 * orc.cells)
        dlt = torch.tensor(dl, device="cuda")
        assert rel(dev.hdg_apply(dlt).cpu().numpy(), orc.hdg_apply(dl)) <= tol
        for f in [0, orc.cells // 2, orc.cells - 1]:
            assert rel(dev.hdg_block(f), orc.hdg_block(f, f)) <= tol, f
        assert rel(dev.hdg_bj_apply(dlt).cpu().numpy(), orc.hdg_bj_apply(dl)) <= 100 * tol
        assert rel(dev.hdg_recover(dlt).cpu().numpy(), orc.hdg_recover(dl)) <= tol
        assert rel(dev.hdg_recovered_residual(dlt).cpu().numpy(), orc.hdg_recovered_residual(dl)) <= tol
        x_d, its, rel_d = dev.hdg_gmres(torch.tensor(g_o, device="cuda"), 1500, 1e-12)
        assert rel_d <= 1e-12 and its > 0
        res = orc.hdg_apply(x_d.cpu().numpy()) - g_o
        assert np.linalg.norm(res) <= 2e-12 * np.linalg.norm(g_o)


HDG_RUNS = [("be-k2-16", ["--cells", "16", "--k", "2", "--integrator", "be", "--T", "0.05"]),
            ("dirk3-k3-q5-12", ["--cells", "12", "--k", "3", "--quad", "5", "--T", "0.05"]),
            ("density-dirk3-k2-32", ["--cells", "32", "--k", "2", "--T", "0.2", "--testcase", "density"])]


@pytest.mark.parametrize("name,args", HDG_RUNS, ids=[r[0] for r in HDG_RUNS])
def test_1d_hdg_runs_match_oracle(name, args, torch_cuda):
    from dg_hdg_euler_cpp_amd import DeviceContext1D

    full = args + ["--cfl", "2", "--ntol", "1e-12", "--ltol", "1e-12", "--restart", "20", "--method", "hdg"]
    print(f"\n[case] 1d hdg run {full}", flush=True)
    orc = O.Oracle1D(full)
    uN_o, st_o = orc.advance(orc.initial(), method="hdg")
    with DeviceContext1D(full) as dev:
        u = dev.initial()
        st = dev.advance(u)
        uN = u.cpu().numpy()
    assert len(st["updates"]) == len(st_o["updates"])
    diffs = []
    for got, want in zip(st["updates"], st_o["updates"]):
        w = want["residual_physical_rms"]
        assert abs(got["residual_physical_rms"] - w) <= max(1e-8 * w, 1e-13), (got, want)
        diffs.append(int(got["krylov_iterations"] - want["krylov_iterations"]))
    print(f"Krylov differences per update {diffs}; final state rel {rel(uN, uN_o):.3e}")
    assert max(abs(d) for d in diffs) <= 2
    assert rel(uN, uN_o) <= 1e-11
    assert st["conservation_drift"] <= 1e-14


@pytest.mark.parametrize("cells", [32, 64, 128])
def test_1d_hdg_convergence_protocol_matches_reference(cells, torch_cuda):
    """The reference table's hdg rows (convergence_slurm_7622710.csv)."""
    from dg_hdg_euler_cpp_amd import DeviceContext1D

    ref = [r for r in GOLD["rows"] if r["method"] == "hdg" and int(r["cells"]) == cells][0]
    args = ["--cells", str(cells)] + GOLD["args"] + ["--method", "hdg"]
    with DeviceContext1D(args) as dev:
        u = dev.initial()
        st = dev.advance(u)
        err = dev.error_l2(u, st["final_time"])
    print(f"\n[case] 1d hdg convergence {cells}: steps {st['steps']} Newton {st['newton_iterations']} "
          f"Krylov {st['krylov_iterations']} (reference {ref['steps']}, {ref['newton_iterations']}, "
          f"{ref['krylov_iterations']}), {st['seconds']:.3f} s", flush=True)
    assert st["steps"] == int(ref["steps"]) and st["final_time"] == 0.2
    assert st["newton_iterations"] == int(ref["newton_iterations"])
    assert abs(st["krylov_iterations"] - int(ref["krylov_iterations"])) <= 2
    assert st["conservation_drift"] <= 1e-14
    for e, key in zip(err, ["rho_L2", "momentum_L2", "energy_L2"]):
        assert abs(e - float(ref[key])) <= 2e-9 * float(ref[key]), (key, e, ref[key])


def test_1d_density_protocol(torch_cuda):
    """The stated 1D initial condition rho = 1 + 0.2 sin(pi x) (`--testcase
    density`, pinned on the oracle by tests/test_oracle_1d_density.py) through
    the 1D protocol at 64 cells (k=2, DIRK3, CFL 2, T 0.2): DG and condensed
    HDG on the device against the oracle, and DG = HDG."""
    from dg_hdg_euler_cpp_amd import DeviceContext1D

    args = ["--cells", "64", "--k", "2", "--integrator", "ssp-dirk3", "--cfl", "2", "--T", "0.2", "--ntol", "1e-12",
            "--ltol", "1e-12", "--restart", "20", "--max-linear", "1500", "--max-newton", "30", "--safety", "1.1",
            "--testcase", "density"]
    finals = {}
    for method in ("dg", "hdg"):
        print(f"\n[case] 1d density {method}", flush=True)
        orc = O.Oracle1D(args + ["--method", method])
        u0_o = orc.initial()
        uN_o, st_o = orc.advance(u0_o, method=method)
        with DeviceContext1D(args + ["--method", method]) as dev:
            u = dev.initial()
            assert rel(u.cpu().numpy(), u0_o) <= 1e-13
            st = dev.advance(u)
            uN = u.cpu().numpy()
            err = dev.error_l2(u, st["final_time"])
        assert st["steps"] == st_o["steps"] and st["newton_iterations"] == st_o["newton_iterations"]
        assert rel(uN, uN_o) <= 1e-11
        assert np.allclose(err, orc.error_l2(uN_o, st_o["final_time"]), rtol=1e-9, atol=0)
        assert st["conservation_drift"] <= 1e-13
        finals[method] = uN
    assert np.abs(finals["dg"] - finals["hdg"]).max() <= 1e-9
