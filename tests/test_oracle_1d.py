"""The oracle's 1D DG restatement (dg_hdg_1d/src) pinned without retyping it:

* whole DIRK3 runs against the reference's archived 1D convergence table
  (accuracy_test/results/1d_slurm_7622710/convergence_slurm_7622710.csv,
  copied into tests/golden/reference_runs.json): step, Newton and Krylov
  counts, minimum step, L2 errors and conservation;
* the Jacobian against central differences of the residual (frozen speeds
  make the residual smooth in u), collocated, over-integrated and k = 0;
* the element blocks against the matrix-free Jacobian on unit vectors, and
  the block-Jacobi apply against a dense solve of each diagonal block;
* GMRES's true residual against its tolerance (gmres.cpp:78-83);
* the condensed HDG operator (hybrid.cpp, 1D): the reference's HDG rows of the
  same table (Krylov counts included), the DG/HDG agreement of the final
  states, the recovered-residual identity J recover(dl) + r =
  recovered_residual(dl) against the pinned DG Jacobian, and the face blocks
  against the Schur apply on unit vectors.
"""
import json
import os

import numpy as np
import pytest

import oracle_py as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_runs.json")))["convergence_1d"]


def protocol(cells):
    return ["--cells", str(cells), "--k", "2", "--integrator", "ssp-dirk3", "--cfl", "2", "--T", "0.2", "--ntol",
            "1e-12", "--ltol", "1e-12", "--restart", "20", "--max-linear", "1500", "--max-newton", "30",
            "--safety", "1.1"]


@pytest.mark.parametrize("cells", [32, 64, 128, 256])
def test_convergence_table(cells):
    ref = [r for r in GOLD["rows"] if r["method"] == "dg" and int(r["cells"]) == cells][0]
    o = O.Oracle1D(protocol(cells))
    uN, st = o.advance(o.initial())
    assert int(st["steps"]) == int(ref["steps"])
    assert int(st["newton_iterations"]) == int(ref["newton_iterations"])
    assert abs(int(st["krylov_iterations"]) - int(ref["krylov_iterations"])) <= 2
    assert st["final_time"] == 0.2
    assert abs(st["minimum_step"] - float(ref["min_step"])) <= 1e-12 * float(ref["min_step"])
    assert st["max_linear_residual"] <= 1e-12
    assert st["conservation_drift"] <= 1e-14
    err = o.error_l2(uN, st["final_time"])
    for e, key in zip(err, ["rho_L2", "momentum_L2", "energy_L2"]):
        assert abs(e - float(ref[key])) <= 2e-9 * float(ref[key]), (key, e, ref[key])


@pytest.mark.parametrize("cells", [32, 64, 128])
def test_hdg_convergence_table(cells):
    ref = [r for r in GOLD["rows"] if r["method"] == "hdg" and int(r["cells"]) == cells][0]
    o = O.Oracle1D(protocol(cells))
    uN, st = o.advance(o.initial(), method="hdg")
    assert int(st["steps"]) == int(ref["steps"])
    assert int(st["newton_iterations"]) == int(ref["newton_iterations"])
    assert abs(int(st["krylov_iterations"]) - int(ref["krylov_iterations"])) <= 2
    assert st["conservation_drift"] <= 1e-14
    err = o.error_l2(uN, st["final_time"])
    for e, key in zip(err, ["rho_L2", "momentum_L2", "energy_L2"]):
        assert abs(e - float(ref[key])) <= 2e-9 * float(ref[key]), (key, e, ref[key])
    uD, _ = O.Oracle1D(protocol(cells)).advance(o.initial())
    assert np.abs(uD - uN).max() <= 1e-13 and float(ref["difference_max_nodal"]) <= 1e-13


@pytest.mark.parametrize("args", [["--cells", "6", "--k", "2"], ["--cells", "5", "--k", "3", "--quad", "5"],
                                  ["--cells", "4", "--k", "0"]], ids=["k2", "k3-q5", "k0"])
def test_hdg_identities(args):
    o, u = setup(args)
    r = o.residual(u)
    o.refresh(u)
    o.hdg_refresh(u, r)
    dl = np.random.default_rng(1).standard_normal(3 * o.cells)
    rr = o.hdg_recovered_residual(dl)
    assert np.linalg.norm(o.apply(o.hdg_recover(dl)) + r - rr) <= 1e-13 * np.linalg.norm(rr)
    C = o.cells
    blocks = np.zeros((3 * C, 3 * C))
    for a in range(C):
        for b in range(C):
            blocks[3 * a:3 * a + 3, 3 * b:3 * b + 3] = o.hdg_block(a, b)
    E = np.eye(3 * C)
    applied = np.stack([o.hdg_apply(E[:, j]) for j in range(3 * C)], axis=1)
    assert np.abs(blocks - applied).max() <= 1e-14 * np.abs(applied).max()


def setup(args, cfl=2.0):
    o = O.Oracle1D(args)
    u = o.initial()
    rate = o.freeze(u)
    o.set_step(cfl / rate, u * (1 + 1e-3 * np.sin(np.arange(u.size))))
    return o, u


@pytest.mark.parametrize("args", [["--cells", "7", "--k", "2"], ["--cells", "5", "--k", "4", "--quad", "7"],
                                  ["--cells", "6", "--k", "0"], ["--cells", "4", "--k", "8"]],
                         ids=["k2", "k4-q7", "k0", "k8"])
def test_jacobian_matches_central_differences(args):
    o, u = setup(args)
    o.refresh(u)
    x = np.random.default_rng(3).standard_normal(u.size)
    eps = 1e-6 * np.linalg.norm(u) / np.linalg.norm(x)
    fd = (o.residual(u + eps * x) - o.residual(u - eps * x)) / (2 * eps)
    jv = o.apply(x)
    assert np.linalg.norm(jv - fd) <= 1e-7 * np.linalg.norm(jv)


@pytest.mark.parametrize("args", [["--cells", "5", "--k", "2"], ["--cells", "4", "--k", "3", "--quad", "6"],
                                  ["--cells", "3", "--k", "0"]], ids=["k2", "k3-q6", "k0"])
def test_blocks_match_matrix_free(args):
    o, u = setup(args)
    o.refresh(u)
    n, m, C = o.size, o.m, o.cells
    dense = np.zeros((n, n))
    for r in range(C):
        for c in range(C):
            dense[r * m:(r + 1) * m, c * m:(c + 1) * m] = o.block(r, c)
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1
        assert np.abs(o.apply(e) - dense[:, j]).max() <= 1e-12 * np.abs(dense).max()
    o.bj_refresh()
    r = np.random.default_rng(4).standard_normal(n)
    z = o.bj_apply(r)
    for c in range(C):
        B = dense[c * m:(c + 1) * m, c * m:(c + 1) * m]
        assert np.allclose(B @ z[c * m:(c + 1) * m], r[c * m:(c + 1) * m], rtol=0, atol=1e-12 * np.abs(r).max())


def test_gmres_true_residual():
    o, u = setup(["--cells", "32", "--k", "2"], cfl=2.0)
    o.refresh(u)
    o.bj_refresh()
    rhs = -o.residual(u)
    x, its, rel = o.gmres(rhs, 20, 1500, 1e-12)
    w = 1.0 / np.array([0.5 * o.dx * wq for wq in O.gauss(o.n)[1]] * (3 * o.cells))
    res = o.apply(x) - rhs
    assert rel <= 1e-12 and its > 0
    assert np.sqrt(np.sum(res ** 2 * w) / np.sum(rhs ** 2 * w)) <= 1e-12


def test_initial_state_is_the_projected_wave():
    o = O.Oracle1D(["--cells", "16", "--k", "3"])
    u = o.initial()
    assert o.admissible(u)
    bad = u.copy()
    bad[0] = -1.0
    assert not o.admissible(bad)
    # the projection is exact to the 20-point rule's accuracy: tiny L2 error at t = 0
    assert np.all(o.error_l2(u, 0.0) < 1e-5)
    # total mass of the simple wave over [-1, 1]
    assert o.integral(u)[0] > 0
