"""Dense CPU restatement of BE / DIRK3 on a frozen schedule and of its
discrete adjoint, built on the oracle's residual, freeze and dense Jacobian.

The stage structure follows time_integrator.cpp:48-213 (tableau :52-56, stage
base :81-83, freeze before stages > 0 :84, step size a_ss h :85, slopes :195,
step combination :197-201). The schedule (step sizes and the states each stage
froze the HLL speeds at) is recorded on the first run and reused by replays,
so the map u_0 -> u_N is smooth and its gradient is what the adjoint returns.
Stage equations are solved by Newton with dense LU (small meshes only).
Test infrastructure: a reference for the device adjoint, not a product path.
"""
import numpy as np

from oracle_py import gauss

ALPHA = .24169426078821
BETA = ALPHA / 4.
ETA = .12915286960590
DIRK_A = np.array([[ALPHA, 0, 0, 0], [-ALPHA, ALPHA, 0, 0], [0, 1 - ALPHA, ALPHA, 0],
                   [BETA, ETA, .5 - BETA - ETA - ALPHA, ALPHA]])
DIRK_B = np.array([0, 1. / 6, 1. / 6, 2. / 3])


def tableau(stages):
    if stages == 1:
        return np.ones((1, 1)), np.ones(1)
    return DIRK_A, DIRK_B


def node_mass(orc):
    """Diagonal DG mass per state entry: .25 hx hy w_i w_j at node i + n j."""
    _, w = gauss(orc.n)
    cell = (.25 * orc.hx * orc.hy * np.outer(w, w)).ravel(order="F")
    return np.tile(cell, 4 * orc.cells)


def quadrant_weights(orc, full=False):
    """Gradient of the density integral over the cells with centre x > 0,
    y > 0 (all cells if full)."""
    mass = node_mass(orc).reshape(orc.cells, 4, orc.n * orc.n)
    g = np.zeros_like(mass)
    for cell in range(orc.cells):
        cx, cy = cell % orc.nx, cell // orc.nx
        if full or (2 * cx + 1 > orc.nx and 2 * cy + 1 > orc.ny):
            g[cell, 0] = mass[cell, 0]
    return g.ravel()


def newton(orc, U, iterations=12):
    for _ in range(iterations):
        r = orc.dg_residual(U)
        orc.dg_refresh(U)
        dU = np.linalg.solve(orc.dense_jacobian(), -r)
        U = U + dU
        if np.linalg.norm(dU) <= 1e-15 * np.linalg.norm(U):
            break
    return U


def run(orc, u0, stages, steps, cfl, schedule=None):
    """Primal run; returns (u_N, record). record[n] = (h, freeze_states, U_s)."""
    A, b = tableau(stages)
    u = np.array(u0, dtype=np.float64)
    record = []
    for n in range(steps):
        if schedule is None:
            h = cfl / orc.freeze(u)
            frozen = [u.copy()]
        else:
            h, frozen, _ = schedule[n]
            orc.freeze(frozen[0])
        k, Us, U = [], [], u.copy()
        for s in range(stages):
            base = u + h * sum(A[s, j] * k[j] for j in range(s)) if s else u.copy()
            if s > 0:
                if schedule is None:
                    frozen.append(U.copy())
                orc.freeze(frozen[s])
            tau = A[s, s] * h
            orc.set_step(tau, base)
            U = newton(orc, U)
            k.append((U - base) / tau)
            Us.append(U.copy())
        u = U.copy() if stages == 1 else u + h * sum(b[s] * k[s] for s in range(stages))
        record.append((h, frozen, Us))
    return u, record


def adjoint(orc, record, final_gradient, stages):
    """dJ/du_0 from dJ/du_N by the reverse sweep over the recorded stages."""
    A, b = tableau(stages)
    mass = node_mass(orc)
    ubar = np.array(final_gradient, dtype=np.float64)
    for h, frozen, Us in reversed(record):
        kbar = [h * b[s] * ubar for s in range(stages)]
        for s in reversed(range(stages)):
            tau = A[s, s] * h
            orc.freeze(frozen[s])
            orc.set_step(tau, Us[s])
            orc.dg_refresh(Us[s])
            lam = np.linalg.solve(orc.dense_jacobian().T, kbar[s] / tau)
            basebar = (mass * lam - kbar[s]) / tau
            ubar = ubar + basebar
            for j in range(s):
                kbar[j] = kbar[j] + h * A[s, j] * basebar
    return ubar
