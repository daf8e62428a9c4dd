"""Pins the CPU oracle before the tests compare the HIP path with it.

Each check is one the reference itself runs (check_2d.cpp / check_1d.cpp) or a
closed form / archived golden vector of the reference (tests/golden/
reference_runs.json, copied from its result CSVs). A plausible restatement
mistake (dropped term, wrong sign or index, transposed operand) fails at least
one of them.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle_py as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_runs.json")))


def direction(size):
    """check_2d.cpp:9-11 deterministic test direction."""
    i = np.arange(size, dtype=np.float64)
    return 0.2 * np.sin(0.31 * i) + 0.1 * np.cos(0.79 * i)


# ---------------------------------------------------------------- basis
@pytest.mark.parametrize("n", [1, 2, 3, 5, 9, 20, 21])
def test_gauss_rule_exactness(n):
    """sum w = 2 and exact integration of x^p, p <= 2n-1 (check_1d.cpp:67-83)."""
    x, w = O.gauss(n)
    assert abs(w.sum() - 2) < 1e-14
    assert np.all(np.diff(x) > 0)
    assert np.allclose(x, -x[::-1], atol=0, rtol=0)
    for p in range(2 * n):
        exact = 0.0 if p % 2 else 2.0 / (p + 1)
        assert abs(np.dot(w, x ** p) - exact) < 1e-13, p


@pytest.mark.parametrize("k", [1, 2, 4, 8])
def test_derivative_matrix_on_polynomials(k):
    """weighted_derivatives = diag(w) D reproduces polynomial derivatives (check_1d.cpp:75-82)."""
    o = O.Oracle(["--nx", "3", "--ny", "3", "--k", str(k)])
    b = o.basis()
    x, w, wd = b["x"], b["w"], b["wd"]
    D = wd / w[:, None]  # wd(q, i) = w_q l'_i(x_q)
    for p in range(k + 1):
        f = x ** p
        df = p * x ** (p - 1) if p else np.zeros_like(x)
        assert np.allclose(D @ f, df, atol=2e-12)
    assert np.allclose(b["left"].sum(), 1.0, atol=1e-14) and np.allclose(b["right"].sum(), 1.0, atol=1e-14)
    assert np.allclose(b["right"] @ x ** k, 1.0, atol=1e-13)
    assert np.allclose(b["left"] @ x ** k, (-1.0) ** k, atol=1e-13)


# -------------------------------------------------------------- physics
@pytest.mark.parametrize("rho", [0.3, 0.7, 1.3])
@pytest.mark.parametrize("axis", [0, 1])
def test_physics_jacobians(rho, axis):
    """check_2d.cpp:26-36: cached == analytic to 1e-13, central FD of the flux to 1e-8."""
    u = np.array([rho, 0.7 * rho, -0.2 * rho, 1.7 + 0.265 * rho])
    _, analytic, cached = O.physics(u, axis)
    assert np.linalg.norm(cached - analytic) < 1e-13
    d = direction(4)
    eps = 2e-7
    fp, _, _ = O.physics(u + eps * d, axis)
    fm, _, _ = O.physics(u - eps * d, axis)
    assert np.linalg.norm((fp - fm) / (2 * eps) - cached @ d) < 1e-8
    # Euler flux is homogeneous of degree one: F(u) = A(u) u
    f, _, _ = O.physics(u, axis)
    assert np.allclose(analytic @ u, f, atol=1e-13)


# ------------------------------------------------------ sum factorisation
@pytest.mark.parametrize("k", [0, 1, 2, 4, 8])
@pytest.mark.parametrize("over", [False, True])
def test_sum_factorisation_brute_force(k, over):
    """check_2d.cpp:37-51: subtract_weak vs an independent quadruple sum, 3e-12."""
    n = k + 1
    Q = n + (2 if over else 0)
    hx, hy = 0.3, 0.7
    values, wd, qw = O.basis_tables(k, Q)
    q = np.arange(Q * Q)
    fx = np.stack([np.sin(0.23 * q + 0.4 * v) for v in range(4)], axis=1)
    fy = np.stack([np.cos(0.12 * q - 0.7 * v) for v in range(4)], axis=1)
    out = O.subtract_weak(k, Q, hx, hy, fx.ravel(order="F"), fy.ravel(order="F"), np.zeros(4 * n * n))
    ref = np.zeros((n * n, 4))
    if n > 1:
        for v in range(4):
            for j in range(n):
                for i in range(n):
                    s = 0.0
                    for qy in range(Q):
                        for qx in range(Q):
                            s += (0.5 * hy * wd[qx, i] * qw[qy] * values[qy, j] * fx[qx + Q * qy, v]
                                  + 0.5 * hx * qw[qx] * values[qx, i] * wd[qy, j] * fy[qx + Q * qy, v])
                    ref[i + n * j, v] = -s
    assert np.linalg.norm(out - ref.ravel(order="F")) < 3e-12


# ------------------------------------------------------- block LU / pool
@pytest.mark.parametrize("width", [4, 12, 36, 324])
def test_lu_forced_permutation(width):
    """check_2d.cpp:12-25: cyclic permutation matrix forces pivoting, exact solve to 2e-12."""
    a = np.zeros((width, width))
    for i in range(width):
        a[i, (i + 1) % width] = 2.0 + 0.01 * i
        a[i, i] += 0.1
    exact = direction(width)
    x = O.lu_solve(a, a @ exact)
    assert np.linalg.norm(x - exact) < 2e-12


# ---------------------------------------------------------- operators
@pytest.mark.parametrize("k", [0, 1, 2, 4, 8])
@pytest.mark.parametrize("over", [False, True])
def test_operators(k, over):
    """check_2d.cpp:52-86 on a 3x4 mesh with dt = 0.003."""
    args = ["--nx", "3", "--ny", "4", "--k", str(k), "--quad", str(k + 1 + (2 if over else 0))]
    o = O.Oracle(args)
    u = o.initial()
    o.freeze(u)
    o.set_step(0.003)  # old = u from freeze
    o.dg_refresh(u)
    d = direction(u.size)
    jv = o.dg_apply(d)
    r = o.dg_residual(u)
    eps = 2e-7
    fd = (o.dg_residual(u + eps * d) - o.dg_residual(u - eps * d)) / (2 * eps)
    assert np.linalg.norm(jv - fd) / max(1.0, np.linalg.norm(jv)) < 2e-8, "DG derivative finite difference"
    m = o.m
    assembled = np.zeros_like(u)
    for row in range(o.cells):
        cols = sorted({row, *[nb for nb in neighbors(row, o.nx, o.ny)]})
        for col in cols:
            assembled[row * m:(row + 1) * m] += o.dg_block(row, col) @ d[col * m:(col + 1) * m]
    assert np.linalg.norm(assembled - jv) / max(1.0, np.linalg.norm(jv)) < 2e-12, "DG matrix-free/block identity"
    lam = o.trace(u)
    hr, g = o.hybrid_residual(u, lam)
    assert np.linalg.norm(hr - r) < 1e-12, "Projected DG/HDG residual identity"
    assert np.linalg.norm(g) < 1e-12, "Projected skeleton equation"
    o.hdg_refresh_projected(u, r)
    dl = direction(o.trace_size)
    sdl = o.hdg_apply(dl)
    du = o.hdg_recover(dl)
    check = o.hdg_recovered_residual(sdl - o.hdg_rhs())
    ref = o.dg_apply(du) + r
    assert np.linalg.norm(check - ref) / max(1.0, np.linalg.norm(ref)) < 3e-11, "Recovered DG residual identity"
    fw = 4 * o.n
    sblock = np.zeros(o.trace_size)
    for row in range(o.faces):
        for col in trace_neighbors(row, o):
            sblock[row * fw:(row + 1) * fw] += o.hdg_block(row, col) @ dl[col * fw:(col + 1) * fw]
    assert np.linalg.norm(sblock - sdl) / max(1.0, np.linalg.norm(sdl)) < 2e-12, "Schur matrix-free/block identity"


def neighbors(cell, nx, ny):
    x, y = cell % nx, cell // nx
    return [(x + nx - 1) % nx + nx * y, (x + 1) % nx + nx * y, x + nx * ((y + ny - 1) % ny), x + nx * ((y + 1) % ny)]


def cell_faces(cell, nx, ny):
    C = nx * ny
    nb = neighbors(cell, nx, ny)
    return [nb[0], cell, C + nb[2], C + cell]


def trace_neighbors(face, o):
    cols = set()
    for cell in range(o.cells):
        fs = cell_faces(cell, o.nx, o.ny)
        if face in fs:
            cols.update(fs)
    return sorted(cols)


# ------------------------------------------------------------- adjoint
def test_adjoint_oracle_matches_dense_transpose():
    """The oracle's J^T w (block-row transpose) equals the dense Jacobian's
    transpose at 8x8, k=2 (2304 unknowns) and the FD duality identity."""
    o = O.Oracle(["--nx", "8", "--ny", "8", "--k", "2", "--cfl", "5"])
    u = o.seeded(3)
    rate = o.freeze(u)
    o.set_step(5.0 / rate)
    o.dg_refresh(u)
    J = o.dense_jacobian()
    rng = np.random.default_rng(1)
    w = rng.standard_normal(u.size)
    v = rng.standard_normal(u.size)
    z = o.dg_apply_transpose(w)
    assert np.linalg.norm(z - J.T @ w) / np.linalg.norm(z) < 1e-13
    assert np.linalg.norm(o.dg_apply(v) - J @ v) / np.linalg.norm(J @ v) < 1e-13
    h = 1e-7
    fd = (o.dg_residual(u + h * v) - o.dg_residual(u - h * v)) / (2 * h)
    assert abs(np.dot(w, fd) - np.dot(z, v)) <= 1e-6 * abs(np.dot(z, v))


# ------------------------------------------------------------ seeded state
def test_seeded_generator_known_answers():
    """splitmix64 finaliser: the stream of seed 0 starts at the published
    0xe220a8397b1dcdaf (state 0x9e3779b97f4a7c15)."""
    u = O.uniform(0, 3)
    assert u[0] == (0xE220A8397B1DCDAF >> 11) * 2.0 ** -53
    assert np.all((u >= 0) & (u < 1))


def test_seeded_state_admissible_and_mean_preserving():
    o = O.Oracle(["--nx", "8", "--ny", "8", "--k", "8", "--testcase", "seeded"])
    u = o.seeded(2025)
    assert o.admissible(u)
    assert np.array_equal(u, o.seeded(2025))
    assert not np.array_equal(u, o.seeded(2026))
    # positivity scaling keeps the cell mean of the raw draw
    n, P, m = o.n, o.n * o.n, o.m
    w = o.basis()["w"]
    wij = 0.25 * np.outer(w, w).T.ravel()  # node i + n j -> w_i w_j / 4
    raw = np.zeros(4 * P)
    cell = 5
    for k in range(P):
        base = 4 * (cell * P + k)
        rho = 0.5 + O.uniform(2025, base + 1)[base]
        vx = -0.5 + O.uniform(2025, base + 2)[base + 1]
        vy = -0.5 + O.uniform(2025, base + 3)[base + 2]
        p = 0.5 + O.uniform(2025, base + 4)[base + 3]
        raw[k], raw[P + k], raw[2 * P + k] = rho, rho * vx, rho * vy
        raw[3 * P + k] = p / 2 + 0.5 * rho * (vx * vx + vy * vy)
    got = u[cell * m:(cell + 1) * m]
    for v in range(4):
        assert abs(np.dot(wij, got[v * P:(v + 1) * P]) - np.dot(wij, raw[v * P:(v + 1) * P])) < 1e-13


# ------------------------------------------------------- golden vectors
@pytest.mark.parametrize("key", ["be_4x4_k2", "dirk3_4x4_k2", "entropy_4x4_k2"])
@pytest.mark.parametrize("method", ["dg", "hdg"])
def test_reference_runs_4x4(key, method):
    """Per-update RMS, Krylov counts, CFL, L2 errors and conservation of the
    reference's archived 4x4, k=2 runs (job_7687730/checks/*.csv)."""
    run = GOLD["runs"][key]
    o = O.Oracle(run["args"])
    u1, st = o.advance(method, o.initial())
    ref = [r for r in run["iterations"] if r["method"] == method]
    assert len(st["updates"]) == len(ref)
    tol = [1e-13, 1e-10, 1e-6]
    for got, want in zip(st["updates"], ref):
        it = int(want["iteration"])
        w = float(want["residual_physical_rms"])
        # relative band per update index; near round-off the band is absolute
        # against the stage's first RMS (Newton stops at 1e-12 of it)
        first = float([r for r in ref if r["stage"] == want["stage"] and r["iteration"] == "0"][0]["residual_physical_rms"])
        assert abs(got["residual_physical_rms"] - w) <= max(tol[it] * w, 1e-12 * first), (it, got["residual_physical_rms"], w)
        assert abs(got["krylov_iterations"] - int(want["krylov_iterations"])) <= 1
        assert got["cfl"] == pytest.approx(float(want["cfl"]), rel=1e-14)
    summ = [s for s in run["summary"] if s["method"] == method][0]
    err = o.error_l2(u1, st["final_time"])
    for v, k in enumerate(["rho_L2", "momentum_x_L2", "momentum_y_L2", "energy_L2"]):
        assert err[v] == pytest.approx(float(summ[k]), rel=1e-9)
    assert st["conservation_drift"] <= 2e-11


def test_dg_equals_hdg_stepping():
    """check_2d.cpp:87-97: DG and HDG final states agree to 2e-10 (BE and DIRK3)."""
    for integ in ["be", "ssp-dirk3"]:
        o = O.Oracle(["--nx", "4", "--ny", "4", "--k", "2", "--T", "0.01", "--cfl", "2", "--integrator", integ,
                      "--check-update"])
        u0 = o.initial()
        ud, sd = o.advance("dg", u0)
        uh, sh = o.advance("hdg", u0)
        assert sd["conservation_drift"] <= 2e-11 and sh["conservation_drift"] <= 2e-11
        assert np.abs(ud - uh).max() <= 2e-10


@pytest.mark.slow
def test_k8_256_initial_rms_and_dt():
    """RMS0 = 0.40279582096926586, dt = 0.013547128620806707 at 256^2, k=8, CFL 5
    (speed_test/.../job_7635952/data/k8_dg_iterations.csv:2)."""
    want = GOLD["runs"]["k8_256_be"]["iterations"][0]
    o = O.Oracle(["--nx", "256", "--ny", "256", "--k", "8", "--cfl", "5", "--integrator", "be", "--steps", "1"])
    u = o.initial()
    rate = o.freeze(u)
    dt = 5.0 / rate
    o.set_step(dt, u)
    rms = o.residual_norm(o.dg_residual(u))
    assert dt == pytest.approx(float(want["dt"]), rel=4e-16)
    assert rms == pytest.approx(float(want["residual_physical_rms"]), rel=1e-13)


@pytest.mark.slow
def test_convergence_run_32():
    """32^2, k=2, DIRK3, CFL 2, T 0.2: L2 errors of the reference's accuracy run
    (accuracy_test/.../convergence.csv:2)."""
    conv = GOLD["convergence"]
    row = [r for r in conv["rows"] if r["method"] == "dg" and r["nx"] == "32"][0]
    o = O.Oracle(["--nx", "32", "--ny", "32"] + conv["args"])
    u, st = o.advance("dg", o.initial())
    err = o.error_l2(u, st["final_time"])
    assert int(st["steps"]) == int(row["steps"])
    assert int(st["newton_iterations"]) == int(row["newton_iterations"])
    for v, k in enumerate(["rho_L2", "momentum_x_L2", "momentum_y_L2", "energy_L2"]):
        assert err[v] == pytest.approx(float(row[k]), rel=1e-9)
