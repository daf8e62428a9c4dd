"""Adjoint J^T w on the device and the discrete adjoint through BE / DIRK3.

* J^T w against the transpose of the oracle's dense Jacobian (assembled from
  the block provider, as the reference's --linear direct path does), entry by
  entry at 8x8, k=2;
* J^T w as the backward of a torch.autograd.Function whose forward is the DG
  residual, against central differences (h = 1e-7) of the residual at 128^2,
  k = 4 and 8 (the frozen HLL speeds are constants of the residual, as in the
  reference's Jacobians);
* the transposed block Jacobi and transposed GMRES against dense solves;
* the reverse sweep over a recorded trajectory against the dense CPU
  restatement (tests/adjoint_ref.py) at 4x4, k=2, and at 128^2 against
  central differences of replays on the recorded schedule and against the
  conservation identity of the total density.
"""
import numpy as np
import pytest

import adjoint_ref as AR
import oracle_py as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _dev(args):
    from dg_hdg_euler_cpp_amd import DeviceContext

    print(f"\n[case] {args}", flush=True)
    return DeviceContext(args)


def test_jtv_matches_dense_jacobian_transpose(torch_cuda):
    torch = torch_cuda
    args = ["--nx", "8", "--ny", "8", "--k", "2", "--testcase", "seeded", "--cfl", "0.5"]
    orc = O.Oracle(args)
    dev = _dev(args)
    u_o = orc.seeded(11)
    u = torch.tensor(u_o, device="cuda")
    rate = orc.freeze(u_o)
    dev.freeze(u)
    tau = 0.5 / rate
    orc.set_step(tau, u_o)
    dev.set_step(tau, u)
    orc.dg_refresh(u_o)
    dev.dg_refresh(u)
    J = orc.dense_jacobian()
    N = J.shape[0]
    rows = np.zeros((N, N))
    e = torch.zeros(N, dtype=torch.float64, device="cuda")
    out = torch.zeros(N, dtype=torch.float64, device="cuda")
    for i in range(N):
        e.zero_()
        e[i] = 1.0
        dev.dg_apply_transpose(e, out=out)
        rows[i] = out.cpu().numpy()   # J^T e_i = row i of J
    # entrywise, relative to the largest entry (the mass / tau diagonal)
    err = np.abs(rows - J).max() / np.abs(J).max()
    print(f"max |J^T_dev - J_dense^T| / max|J| = {err:.3e}")
    assert err <= 1e-13


class DGResidual:
    """torch.autograd.Function: forward = DG residual (frozen speeds), backward
    = the device adjoint J^T g at the forward state."""

    @staticmethod
    def make(dev):
        import torch

        class Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, u):
                ctx.save_for_backward(u)
                return dev.dg_residual(u.detach().contiguous())

            @staticmethod
            def backward(ctx, g):
                (u,) = ctx.saved_tensors
                dev.dg_refresh(u.detach().contiguous())
                return dev.dg_apply_transpose(g.contiguous())

        return Fn


@pytest.mark.parametrize("k", [4, 8])
def test_jtv_autograd_matches_central_differences(k, torch_cuda):
    torch = torch_cuda
    args = ["--nx", "128", "--ny", "128", "--k", str(k), "--testcase", "seeded", "--cfl", "0.5"]
    dev = _dev(args)
    u0 = dev.seeded(2025)
    rate = dev.freeze(u0)
    dev.set_step(0.5 / rate, u0 * 0.999)
    g = torch.Generator(device="cuda")
    g.manual_seed(17)
    v = torch.randn(u0.numel(), dtype=torch.float64, device="cuda", generator=g)
    w = torch.randn(u0.numel(), dtype=torch.float64, device="cuda", generator=g)
    fn = DGResidual.make(dev)
    u = u0.clone().requires_grad_(True)
    (fn.apply(u) * w).sum().backward()
    analytic = float(torch.dot(u.grad, v))
    h = 1e-7
    rp = dev.dg_residual(u0 + h * v)
    rm = dev.dg_residual(u0 - h * v)
    fd = float(torch.dot(w, rp - rm)) / (2 * h)
    err = abs(fd - analytic) / abs(analytic)
    print(f"k={k}: <J^T w, v> = {analytic:.12e}, central FD = {fd:.12e}, rel {err:.3e}")
    assert err <= 1e-6


def test_transposed_block_jacobi_and_gmres(torch_cuda):
    torch = torch_cuda
    args = ["--nx", "8", "--ny", "8", "--k", "2", "--testcase", "seeded", "--cfl", "0.5", "--restart", "20"]
    orc = O.Oracle(args)
    dev = _dev(args)
    u_o = orc.seeded(11)
    u = torch.tensor(u_o, device="cuda")
    tau = 0.5 / dev.freeze(u)
    dev.set_step(tau, u)
    dev.dg_refresh(u)
    dev.bj_refresh_transposed()
    r = torch.tensor(np.random.default_rng(4).standard_normal(u.numel()), device="cuda")
    z = dev.bj_apply_transposed(r).cpu().numpy()
    m = dev.m
    for row in (0, 17, dev.cells - 1):
        B = dev.dg_block(row)
        want = np.linalg.solve(B.T, r.cpu().numpy()[row * m:(row + 1) * m])
        # forward error of a partial-pivot LU solve: kappa * eps scale
        assert rel(z[row * m:(row + 1) * m], want) <= 1e-13 * np.linalg.cond(B)
    # the primal apply must refuse the transposed factors
    with pytest.raises(Exception):
        dev.bj_apply(r)
    x, its, relres, res = dev.dg_gmres_transposed(r, 1500, 1e-12)
    check = r - dev.dg_apply_transpose(x)
    print(f"transposed GMRES: {its} iterations, rel {relres:.3e}")
    assert relres <= 1e-12
    assert dev.dual_norm(check) <= 2e-12 * dev.dual_norm(r)
    assert rel(res.cpu().numpy(), check.cpu().numpy()) <= 1e-6


@pytest.mark.parametrize("integrator,stages", [("be", 1), ("ssp-dirk3", 4)])
def test_adjoint_matches_dense_restatement(integrator, stages, torch_cuda):
    torch = torch_cuda
    args = ["--nx", "4", "--ny", "4", "--k", "2", "--testcase", "seeded", "--integrator", integrator,
            "--cfl", "0.5", "--steps", "2", "--restart", "20"]
    orc = O.Oracle(args)
    u0 = orc.seeded(7)
    g = AR.quadrant_weights(orc)
    uN_o, record = AR.run(orc, u0, stages, 2, 0.5)
    grad_o = AR.adjoint(orc, record, g, stages)
    dev = _dev(args)
    u = torch.tensor(u0, device="cuda")
    dev.trajectory_record(u)
    # stage states converge to ntol = 1e-12 on the device, to roundoff here
    assert rel(u.cpu().numpy(), uN_o) <= 1e-11
    value, g_d = dev.quadrant_density(u)
    assert rel(g_d.cpu().numpy(), g) <= 1e-15
    assert abs(value - g @ uN_o) <= 1e-11 * abs(value)
    grad, st = dev.adjoint_backward(g_d)
    err = rel(grad.cpu().numpy(), grad_o)
    print(f"{integrator}: device adjoint vs dense restatement rel {err:.3e}, {st}")
    assert st["linear_solves"] == 2 * stages
    assert err <= 1e-9


# DIRK(4,3) on the per-node random seeded state is not usable: the base of
# stage 2, u_n + (1 - alpha) h k_1, leaves the admissible set and the stage's
# Newton line search stalls (the oracle's advance fails the same way, from 8^2
# to 128^2 and down to CFL 0.02); BE stalls the same way on seeded k=8 states
# (oracle, 8^2, CFL 0.02-0.1) and on seeded k=4 states at 128^2. The
# trajectory gradients therefore run on the smooth rotated simple wave; seeded
# states are covered by the J^T w checks above and the 4x4 dense restatement.
TRAJECTORIES = [
    ("dirk3-diagonal-k4", "ssp-dirk3", "diagonal", 4, "1", 2),
    ("dirk3-diagonal-k8", "ssp-dirk3", "diagonal", 8, "1", 1),
    ("be-diagonal-k8", "be", "diagonal", 8, "1", 1),
]


@pytest.mark.parametrize("name,integrator,testcase,k,cfl,directions", TRAJECTORIES, ids=[t[0] for t in TRAJECTORIES])
def test_adjoint_trajectory_central_differences_128(name, integrator, testcase, k, cfl, directions, torch_cuda):
    """Gradient of the quadrant density after 10 steps on 128^2."""
    torch = torch_cuda
    args = ["--nx", "128", "--ny", "128", "--k", str(k), "--testcase", testcase, "--integrator", integrator,
            "--cfl", cfl, "--steps", "10", "--restart", "20"]
    dev = _dev(args)
    u0 = dev.seeded(2025) if testcase == "seeded" else dev.initial()
    u = u0.clone()
    st = dev.trajectory_record(u)
    print(f"record: {st['newton_iterations']} Newton, {st['krylov_iterations']} Krylov", flush=True)
    J0, g = dev.quadrant_density(u)
    grad, ast = dev.adjoint_backward(g)
    print(f"adjoint: {ast}", flush=True)
    # conservation: the total density is invariant, so its gradient is itself
    x, w = O.gauss(k + 1)
    side = 2 / np.cos(np.pi / 4) if testcase == "diagonal" else 2.0   # example.hpp:12-13 domain
    cell_mass = (.25 * (side / 128) ** 2 * np.outer(w, w)).ravel(order="F")
    full = np.zeros((128 * 128, 4, (k + 1) ** 2))
    full[:, 0, :] = cell_mass
    full = torch.tensor(full.ravel(), device="cuda")
    grad_full, _ = dev.adjoint_backward(full)
    cons = rel(grad_full.cpu().numpy(), full.cpu().numpy())
    print(f"conservation identity rel {cons:.3e}", flush=True)
    assert cons <= 1e-10
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    dirs = [grad / torch.linalg.norm(grad),
            torch.randn(u0.numel(), dtype=torch.float64, device="cuda", generator=gen)][:directions]
    for v in dirs:
        h = 1e-6 * float(torch.linalg.norm(u0)) / float(torch.linalg.norm(v))
        up = u0 + h * v
        um = u0 - h * v
        dev.trajectory_replay(up)
        dev.trajectory_replay(um)
        fd = (dev.quadrant_density(up, with_gradient=False) - dev.quadrant_density(um, with_gradient=False)) / (2 * h)
        an = float(torch.dot(grad, v))
        err = abs(fd - an) / abs(an)
        print(f"k={k}: <grad, v> = {an:.12e}, central FD = {fd:.12e}, rel {err:.3e}", flush=True)
        assert err <= 1e-6
