"""CPU pins of the discrete-adjoint recursion (dense restatement on the oracle).

The reverse sweep is checked against central finite differences of the same
frozen-schedule map and against conservation: with periodic boundaries the
density integral over the whole domain is invariant, so its gradient with
respect to u_0 is exactly its gradient with respect to u_N.
"""
import numpy as np
import pytest

import adjoint_ref as AR
import oracle_py as O


def _case(integrator, k=2, nx=4, cfl="0.5"):
    args = ["--nx", str(nx), "--ny", str(nx), "--k", str(k), "--testcase", "seeded", "--integrator", integrator,
            "--cfl", cfl, "--restart", "20"]
    return O.Oracle(args)


@pytest.mark.parametrize("integrator,stages,steps", [("be", 1, 2), ("ssp-dirk3", 4, 2)])
def test_adjoint_matches_central_differences(integrator, stages, steps):
    orc = _case(integrator)
    u0 = orc.seeded(7)
    g = AR.quadrant_weights(orc)
    uN, record = AR.run(orc, u0, stages, steps, 0.5)
    grad = AR.adjoint(orc, record, g, stages)
    rng = np.random.default_rng(3)
    for v in (rng.standard_normal(orc.size), grad / np.linalg.norm(grad)):
        eps = 1e-6
        up, _ = AR.run(orc, u0 + eps * v, stages, steps, 0.5, schedule=record)
        um, _ = AR.run(orc, u0 - eps * v, stages, steps, 0.5, schedule=record)
        fd = (g @ up - g @ um) / (2 * eps)
        an = grad @ v
        assert abs(fd - an) <= 1e-7 * abs(an), (fd, an)
    # a dropped mass term or a wrong tableau entry moves the gradient far
    # from the identity dJ/du_0 = dJ/du_N of the trivial (zero-step) map
    assert np.linalg.norm(grad - g) > 1e-3 * np.linalg.norm(g)


@pytest.mark.parametrize("integrator,stages", [("be", 1), ("ssp-dirk3", 4)])
def test_adjoint_conserves_total_density(integrator, stages):
    orc = _case(integrator)
    u0 = orc.seeded(7)
    g = AR.quadrant_weights(orc, full=True)
    uN, record = AR.run(orc, u0, stages, 2, 0.5)
    assert abs(g @ uN - g @ u0) <= 1e-13 * abs(g @ u0)
    grad = AR.adjoint(orc, record, g, stages)
    assert np.linalg.norm(grad - g) <= 1e-11 * np.linalg.norm(g)


def test_replay_reproduces_recorded_run():
    orc = _case("ssp-dirk3")
    u0 = orc.seeded(7)
    uN, record = AR.run(orc, u0, 4, 2, 0.5)
    uR, _ = AR.run(orc, u0, 4, 2, 0.5, schedule=record)
    assert np.array_equal(uN, uR)
