"""The 1D protocol's stated initial condition, rho = 1 + 0.2 sin(pi x) (no
reference counterpart: the reference runs its own simple wave,
dg_hdg_1d/src/simple_wave.cpp:3-25), on the oracle. It is the same gamma = 3
simple wave family with the invariant u + c = 2 and p = rho^3 / 3, so that
c = rho and u = 2 - rho at t = 0 (`--testcase density`). The reference holds
no numbers for it, so it is pinned by what the mathematics fixes:

* the initial state: nodal density, velocity and pressure of the projected
  state against the closed forms;
* the solution at T = 0.2 against an exact solution computed here
  independently (the characteristic foot xi of x = xi - 0.4 t sin(pi xi) by
  bisection; rho = 1 + 0.2 sin(pi xi));
* the observed order of the L2 errors (k = 2, DIRK3, CFL 2; the reference's own
  wave shows 2.8-3.0 at the same grids);
* conservation, and DG = condensed HDG on the final state.
"""
import numpy as np
import pytest

import oracle_py as O


def protocol(cells, k=2):
    return ["--cells", str(cells), "--k", str(k), "--integrator", "ssp-dirk3", "--cfl", "2", "--T", "0.2", "--ntol",
            "1e-12", "--ltol", "1e-12", "--restart", "20", "--max-linear", "1500", "--max-newton", "30",
            "--safety", "1.1", "--testcase", "density"]


def node_positions(cells, k):
    x, _ = O.gauss(k + 1)
    h = 2.0 / cells
    return np.array([[-1 + h * (e + 0.5) + 0.5 * h * xi for xi in x] for e in range(cells)])  # [cell][i]


def primitives(u, cells, n):
    U = u.reshape(cells, 3, n)
    rho, m, E = U[:, 0], U[:, 1], U[:, 2]
    v = m / rho
    return rho, v, 2.0 * (E - 0.5 * rho * v * v)  # p = (gamma - 1)(E - rho v^2 / 2), gamma = 3


def exact_rho(x, t):
    """rho(x, t) = 1 + 0.2 sin(pi xi), x = xi - 0.4 t sin(pi xi) (bisection)."""
    lo, hi = x - 0.4 * t - 1e-12, x + 0.4 * t + 1e-12
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        f = mid - 0.4 * t * np.sin(np.pi * mid) - x
        lo, hi = np.where(f > 0, lo, mid), np.where(f > 0, mid, hi)
    return 1 + 0.2 * np.sin(np.pi * 0.5 * (lo + hi))


def test_initial_state_matches_the_closed_forms():
    cells, k = 64, 3
    o = O.Oracle1D(protocol(cells, k))
    u = o.initial()
    rho, v, p = primitives(u, cells, k + 1)
    x = node_positions(cells, k)
    assert np.abs(rho - (1 + 0.2 * np.sin(np.pi * x))).max() < 1e-7
    assert np.abs(v - (2 - rho)).max() < 1e-7
    assert np.abs(p - rho ** 3 / 3).max() < 1e-7
    assert rho.min() > 0.79 and v.min() > 0.79 and p.min() > 0.17
    assert np.all(o.error_l2(u, 0.0) < 1e-8)


def test_solution_at_T_matches_an_independent_exact_solution():
    cells, k = 128, 2
    o = O.Oracle1D(protocol(cells, k))
    uN, st = o.advance(o.initial())
    assert abs(st["final_time"] - 0.2) < 1e-14
    rho, v, p = primitives(uN, cells, k + 1)
    ref = exact_rho(node_positions(cells, k), 0.2)
    assert np.abs(rho - ref).max() < 2e-5
    assert np.abs(v - (2 - ref)).max() < 2e-5
    # the oracle's own error routine agrees with this exact solution's scale
    assert o.error_l2(uN, 0.2)[0] < 1e-6


def test_order_conservation_and_dg_equals_hdg():
    errors = {}
    for cells in (32, 64, 128):
        o = O.Oracle1D(protocol(cells))
        u0 = o.initial()
        mass0 = o.integral(u0)
        uN, st = o.advance(u0)
        errors[cells] = o.error_l2(uN, st["final_time"])
        assert np.abs(o.integral(uN) - mass0).max() <= 1e-12
        assert st["conservation_drift"] <= 1e-12
        if cells == 64:
            uH, _ = O.Oracle1D(protocol(cells)).advance(u0, method="hdg")
            assert np.abs(uH - uN).max() <= 1e-9
    for v in range(3):
        for a, b in ((32, 64), (64, 128)):
            order = np.log2(errors[a][v] / errors[b][v])
            assert 2.6 < order < 3.4, (v, a, order)


def test_unknown_testcase_is_refused():
    with pytest.raises(O.OracleError, match="simple or density"):
        O.Oracle1D(["--cells", "8", "--testcase", "bogus"])
