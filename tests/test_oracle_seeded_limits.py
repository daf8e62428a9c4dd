"""Why the throughput workload is the reference's smooth `diagonal` state and
not the i.i.d. seeded nodal state, shown on the oracle (CPU).

The seeded state (rho, p ~ U[.5, 1.5], u, v ~ U[-.5, .5] per node, then the
reference's mean-preserving positivity scaling, initial_and_error.cpp:19-51) is
white noise at the node level. With the reference's own solver settings
(time_integrator.cpp:48-213, gmres.cpp:17-86: GMRES(20), max-linear 1500,
ntol = ltol = 1e-12, HLL safety 1.1):
  * one BE Newton update at CFL 5 cannot be solved: GMRES(20) + block Jacobi
    stops at the 1500-iteration limit, for DG (k=8) and for the condensed HDG
    Schur system (k=8);
  * SSP-DIRK(4,3) cannot take ten steps even at CFL 0.02 (k=4) or three at
    CFL 0.02 (k=8): the damped Newton line search exhausts its 30 halvings.
The same settings on smooth states converge (the controls below), so
the failures belong to the state, not to the solver. The device reaches the
same limits through the same code path (dgh_advance), which is why the
seeded state is used at CFL <= 0.1 for operator parity and the Newton
throughput is measured on `diagonal`.
"""
import pytest

import oracle_py as O

SOLVER = ["--restart", "20", "--max-linear", "1500", "--ntol", "1e-12", "--ltol", "1e-12", "--safety", "1.1"]


def _advance(args, method, state):
    o = O.Oracle(args + SOLVER)
    u = o.seeded(2025) if state == "seeded" else o.initial()
    return o.advance(method, u)


def _be(nx, k, cfl, state):
    return ["--nx", str(nx), "--ny", str(nx), "--k", str(k), "--testcase", state, "--integrator", "be",
            "--cfl", str(cfl), "--steps", "1", "--newton-steps", "1"]


def _dirk3(nx, k, cfl, steps, state):
    return ["--nx", str(nx), "--ny", str(nx), "--k", str(k), "--testcase", state, "--integrator", "ssp-dirk3",
            "--cfl", str(cfl), "--steps", str(steps)]


@pytest.mark.slow
def test_seeded_be_cfl5_dg_hits_the_gmres_limit():
    """Config 2's update (seeded, k=8, BE, CFL 5, GMRES(20)) on 8x8 cells."""
    with pytest.raises(O.OracleError, match="GMRES iteration limit reached"):
        _advance(_be(8, 8, 5, "seeded"), "dg", "seeded")


def test_seeded_be_cfl5_hdg_hits_the_gmres_limit():
    """Config 3's update (seeded, k=8, condensed HDG, BE, CFL 5) on 6x6 cells:
    the Schur GMRES(20) with face block Jacobi stops at 1500 iterations."""
    with pytest.raises(O.OracleError, match="GMRES iteration limit reached"):
        _advance(_be(6, 8, 5, "seeded"), "hdg", "seeded")


def test_smooth_be_cfl5_converges_with_the_same_settings():
    """Control: a smooth state (the reference's entropy wave; the `diagonal`
    wave would steepen into its shock within one CFL 5 step on so coarse a
    mesh) at the same k, CFL, mesh and solver settings converges within the
    limit, DG and HDG, undamped."""
    for method in ("dg", "hdg"):
        _, st = _advance(_be(6, 8, 5, "wave"), method, "wave")
        upd = st["updates"]
        assert len(upd) == 1 and upd[0]["damping"] == 1.0
        assert 0 < upd[0]["krylov_iterations"] < 1500


def test_seeded_dirk3_line_search_fails_k4():
    """Config 5's trajectory (seeded, k=4, 10 DIRK3 steps) on 8x8 cells:
    fails at CFL 1 and still at CFL 0.02."""
    for cfl in (1, 0.02):
        with pytest.raises(O.OracleError, match="Newton line search failed"):
            _advance(_dirk3(8, 4, cfl, 10, "seeded"), "dg", "seeded")


@pytest.mark.slow
def test_seeded_dirk3_line_search_fails_k8():
    """Config 5 at k=8: three DIRK3 steps at CFL 0.02 on 8x8 cells."""
    with pytest.raises(O.OracleError, match="Newton line search failed"):
        _advance(_dirk3(8, 8, 0.02, 3, "seeded"), "dg", "seeded")


def test_smooth_dirk3_ten_steps_converge():
    """Control: ten DIRK3 steps of the entropy wave at CFL 1, k=4, 8x8."""
    _, st = _advance(_dirk3(8, 4, 1, 10, "wave"), "dg", "wave")
    assert st["steps"] == 10
    assert all(x["damping"] > 0 for x in st["updates"])
