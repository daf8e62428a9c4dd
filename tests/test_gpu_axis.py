"""The `axis` test case (dg_hdg_2d/src/example.cpp:18-37: the simple wave along
x on the 2 x 2 periodic square) through whole device runs: BE and SSP-DIRK3,
DG and condensed HDG, update by update against the oracle's advance; the DG and
HDG final states agree (check_2d.cpp:87-97, 2e-10); the device solution stays
independent of y with zero y-momentum. The operator-level parity of the case
(residual, Jv, J^T w, blocks, block Jacobi, GMRES) is in test_gpu_dg_parity.py;
the oracle's axis path is pinned against the 1D oracle in test_oracle_axis.py.
"""
import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu

RUNS = [("be-k2-6x3", ["--nx", "6", "--ny", "3", "--k", "2", "--integrator", "be"]),
        ("dirk3-k3-8x4", ["--nx", "8", "--ny", "4", "--k", "3", "--integrator", "ssp-dirk3"]),
        ("dirk3-k8-4x3", ["--nx", "4", "--ny", "3", "--k", "8", "--integrator", "ssp-dirk3"])]


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def _full(args, method):
    return args + ["--testcase", "axis", "--cfl", "2", "--T", "0.02", "--ntol", "1e-12", "--ltol", "1e-12",
                   "--restart", "20", "--method", method]


@pytest.mark.parametrize("name,args", RUNS, ids=[r[0] for r in RUNS])
def test_axis_runs_match_oracle_and_dg_equals_hdg(name, args, torch_cuda):
    from dg_hdg_euler_cpp_amd import DeviceContext

    finals = {}
    for method in ("dg", "hdg"):
        full = _full(args, method)
        print(f"\n[case] axis run {full}", flush=True)
        orc = O.Oracle(full)
        uN_o, st_o = orc.advance(method, orc.initial())
        with DeviceContext(full) as dev:
            u = dev.initial()
            st = dev.advance(u)
            uN = u.cpu().numpy()
            err = dev.error_l2(u, st["final_time"])
        assert len(st["updates"]) == len(st_o["updates"])
        for got, want in zip(st["updates"], st_o["updates"]):
            w = want["residual_physical_rms"]
            assert abs(got["residual_physical_rms"] - w) <= max(1e-8 * w, 1e-13), (method, got, want)
            assert abs(got["krylov_iterations"] - want["krylov_iterations"]) <= 1
        assert rel(uN, uN_o) <= 1e-11, (method, rel(uN, uN_o))
        assert st["conservation_drift"] <= 2e-11
        assert np.allclose(err, orc.error_l2(uN_o, st_o["final_time"]), rtol=1e-9, atol=1e-15)
        nx, ny, n = int(args[1]), int(args[3]), int(args[5]) + 1
        W = uN.reshape(ny, nx, 4, n, n)
        assert np.abs(W - W[:1]).max() < 1e-12 and np.abs(W[:, :, 2]).max() < 1e-13
        finals[method] = uN
    # DG and condensed HDG solve the same discrete system (check_2d.cpp:87-97)
    assert np.abs(finals["dg"] - finals["hdg"]).max() <= 2e-10
