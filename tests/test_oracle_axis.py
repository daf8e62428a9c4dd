"""The `axis` test case (dg_hdg_2d/src/example.cpp:18-37, the simple wave along
x on the 2 x 2 periodic square) on the oracle, pinned by the mathematics of a
y-independent state rather than by retyping the formulas:

  * a 2D state that is constant along y with zero y-momentum is a 1D state
    extended: the y fluxes are constant along y, so their weak derivative and
    their (jump-free) HLL face fluxes cancel exactly, and the 2D residual and
    Jacobian action equal the 1D ones (dg_hdg_1d/src/discretization.cpp,
    volume_jacobian.cpp, dg_operator.cpp) times the y quadrature weight
    h_y w_j / 2. The 1D oracle is pinned separately against the reference's
    archived 1D convergence table (tests/test_oracle_1d.py), so this checks
    the 2D x-direction physics, flux and lifting at real degree;
  * a DIRK3 run stays y-independent with zero y-momentum to round-off;
  * its L2 error against the exact simple wave converges at the Q2 rate the
    reference observes for its rotated wave (2.6 at N = 64,
    accuracy_test/.../convergence.csv).
"""
import numpy as np
import pytest

import oracle_py as O

MAP = [(0, 0), (1, 1), (3, 2)]  # 2D variable <- 1D variable (rho, m_x, E); m_y = 0


def extend(u1, nx, ny, n):
    """1D vector [cell][v][i] -> 2D vector [y][x][v][j][i], constant in y."""
    U1 = u1.reshape(nx, 3, n)
    U2 = np.zeros((ny, nx, 4, n, n))
    for v2, v1 in MAP:
        U2[:, :, v2, :, :] = U1[None, :, v1, None, :]
    return U2.ravel()


def compare(vec2, vec1, nx, ny, n, w, hy):
    """max |2D - (hy w_j / 2) 1D| over rho, m_x, E; max |2D m_y|; scale."""
    V2 = vec2.reshape(ny, nx, 4, n, n)
    V1 = vec1.reshape(nx, 3, n)
    scale = 0.5 * hy * np.asarray(w)
    err, big = 0.0, 0.0
    for v2, v1 in MAP:
        pred = V1[None, :, v1, None, :] * scale[None, None, :, None]
        err = max(err, float(np.abs(V2[:, :, v2] - pred).max()))
        big = max(big, float(np.abs(pred).max()))
    return err, float(np.abs(V2[:, :, 2]).max()), big


@pytest.mark.parametrize("nx,ny,k", [(6, 3, 3), (5, 4, 8), (7, 3, 0)])
def test_axis_operators_reduce_to_1d(nx, ny, k):
    n = k + 1
    o1 = O.Oracle1D(["--cells", str(nx), "--k", str(k), "--cfl", "2"])
    o2 = O.Oracle(["--nx", str(nx), "--ny", str(ny), "--k", str(k), "--testcase", "axis", "--cfl", "2",
                   "--integrator", "be"])
    w, hy = o2.basis()["w"], o2.hy
    assert abs(o1.dx - o2.hx) < 1e-15
    u1 = o1.initial()
    u2 = extend(u1, nx, ny, n)
    o1.freeze(u1)
    o2.freeze(u2)
    tau = 0.05
    old1 = u1 * (1 + 1e-3 * np.cos(np.arange(u1.size)))
    o1.set_step(tau, old1)
    o2.set_step(tau, extend(old1, nx, ny, n))
    err, my, big = compare(o2.dg_residual(u2), o1.residual(u1), nx, ny, n, w, hy)
    assert err <= 1e-13 * big and my <= 1e-13 * big, (err, my, big)
    o1.refresh(u1)
    o2.dg_refresh(u2)
    x1 = np.sin(0.37 * np.arange(u1.size)) + 0.2
    err, my, big = compare(o2.dg_apply(extend(x1, nx, ny, n)), o1.apply(x1), nx, ny, n, w, hy)
    assert err <= 1e-13 * big and my <= 1e-13 * big, (err, my, big)


def _axis_run(nx, ny=3, k=2):
    o = O.Oracle(["--nx", str(nx), "--ny", str(ny), "--k", str(k), "--testcase", "axis", "--integrator",
                  "ssp-dirk3", "--cfl", "2", "--T", "0.2", "--restart", "20", "--ntol", "1e-12", "--ltol", "1e-12"])
    uN, st = o.advance("dg", o.initial())
    return o, uN, st


def test_axis_run_stays_one_dimensional_and_converges():
    errors = {}
    for nx in (32, 64):
        o, uN, st = _axis_run(nx)
        assert abs(st["final_time"] - 0.2) < 1e-14
        n = o.n
        W = uN.reshape(3, nx, 4, n, n)  # [y][x][v][j][i]
        assert np.abs(W - W[:1]).max() < 1e-12  # the same in every cell row
        assert np.abs(W - W[:, :, :, :1, :]).max() < 1e-12  # and along y inside a cell
        assert np.abs(W[:, :, 2]).max() < 1e-13  # no y-momentum
        errors[nx] = o.error_l2(uN, st["final_time"])
    for v in (0, 1, 3):
        order = np.log(errors[32][v] / errors[64][v]) / np.log(2.0)
        assert 2.3 < order < 3.3, (v, order)
    assert errors[64][2] < 1e-14
