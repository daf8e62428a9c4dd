"""Mixed-precision device path (no reference counterpart): FP32, FP16 (with a
power-of-two scale of U per block) or BF16-stored block-Jacobi factors, FP32
Jv, FP32 inner GMRES inside FP64 flexible GMRES.

The FP64 oracle pins it: the FP32 pieces against round-off bounds of their
storage precision, the FGMRES solution against the oracle's own linear system
at the reference's ltol (1e-12), and whole Newton runs against the reference's
archived RMS sequences ("equal residual").
"""
import json
import os

import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_runs.json")))


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def mass_vector(orc):
    w = orc.basis()["w"]
    n = orc.n
    node = np.array([0.25 * orc.hx * orc.hy * w[k % n] * w[k // n] for k in range(n * n)])
    return np.tile(node, 4 * orc.cells)


@pytest.mark.parametrize("precision,bound", [("mixed", 1e-5), ("mixed-bf16", 6e-2), ("mixed-fp16", 1e-2)])
@pytest.mark.parametrize("k,nx,state,cfl", [(2, 8, "diagonal", 5.0), (8, 12, "diagonal", 5.0),
                                            (8, 8, "seeded", 0.1), (4, 10, "seeded", 0.5)])
def test_fp32_pieces_and_fgmres(precision, bound, k, nx, state, cfl, torch_cuda):
    """FP32 Jv, the FP32 / 2-byte factor apply (backward error of every sampled
    block solve at the storage precision) and FP64 FGMRES over FP32 inner
    solves, on the smooth projected state and on seeded (node-level random)
    states, whose blocks carry pivots over a wide range of magnitudes."""
    from dg_hdg_euler_cpp_amd import DeviceContext

    args = ["--nx", str(nx), "--ny", str(nx), "--k", str(k), "--cfl", str(cfl), "--integrator", "be",
            "--testcase", state]
    print(f"\n[case] {precision} {args}", flush=True)
    orc = O.Oracle(args)
    dev = DeviceContext(args + ["--precision", precision, "--restart", "20"])
    u_o = orc.initial() if state != "seeded" else orc.seeded(2025)
    u = torch_cuda.tensor(u_o, device="cuda")
    rate = orc.freeze(u_o)
    dev.freeze(u)
    tau = cfl / rate
    orc.set_step(tau, u_o)
    dev.set_step(tau, u)
    orc.dg_refresh(u_o)
    dev.dg_refresh(u)
    dev.bj_refresh_mixed()
    rng = np.random.default_rng(2)
    x = rng.standard_normal(u_o.size)
    # FP32 Jv (FP32 in/out, FP64 arithmetic on the FP64 snapshot)
    y32 = dev.dg_apply_f32(torch_cuda.tensor(x, dtype=torch_cuda.float32, device="cuda")).cpu().numpy()
    assert rel(y32, orc.dg_apply(x.astype(np.float32).astype(np.float64))) < 1e-6
    # FP32 / BF16 factor apply: backward error per block at the storage precision
    r = rng.standard_normal(u_o.size)
    z32 = dev.bj_apply_f32(torch_cuda.tensor(r, dtype=torch_cuda.float32, device="cuda")).cpu().numpy()
    m = orc.m
    worst = 0.0
    for row in range(0, orc.cells, max(1, orc.cells // 8)):
        B = orc.dg_block(row, row)
        zb = z32[row * m:(row + 1) * m].astype(np.float64)
        res = B @ zb - r[row * m:(row + 1) * m].astype(np.float32).astype(np.float64)
        worst = max(worst, np.linalg.norm(res) / (np.linalg.norm(B) * np.linalg.norm(zb)))
    print(f"worst block backward error {worst:.3e} (bound {bound:.0e})", flush=True)
    assert np.all(np.isfinite(z32))
    assert worst < bound, worst
    # FP64 FGMRES with FP32 inner solves reaches the reference's ltol on the
    # oracle's own FP64 system
    rhs = -orc.dg_residual(u_o)
    xd, inner, outer, reld = dev.dg_fgmres(torch_cuda.tensor(rhs, device="cuda"), 1500, 1e-12)
    xd = xd.cpu().numpy()
    assert reld <= 1e-12 and inner > 0 and outer > 0
    w = 1.0 / mass_vector(orc)
    res = orc.dg_apply(xd) - rhs
    assert np.sqrt(np.sum(res ** 2 * w) / np.sum(rhs ** 2 * w)) <= 2e-12
    dev.close()


@pytest.mark.parametrize("precision", ["mixed", "mixed-bf16", "mixed-fp16"])
def test_mixed_newton_runs_match_reference(precision, torch_cuda):
    from dg_hdg_euler_cpp_amd import DeviceContext

    run = GOLD["runs"]["be_4x4_k2"]
    with DeviceContext(run["args"] + ["--method", "dg", "--precision", precision]) as dev:
        u = dev.initial()
        st = dev.advance(u)
        err = dev.error_l2(u, st["final_time"])
    ref = [r for r in run["iterations"] if r["method"] == "dg"]
    first = float(ref[0]["residual_physical_rms"])
    tol = [1e-13, 1e-10, 1e-6]
    for got, want in zip(st["updates"], ref):
        it = int(want["iteration"])
        w = float(want["residual_physical_rms"])
        assert abs(got["residual_physical_rms"] - w) <= max(tol[it] * w, 1e-12 * first)
        assert got["linear_relative_residual"] <= 1e-12
        assert got["outer_iterations"] > 0
    summ = [s for s in run["summary"] if s["method"] == "dg"][0]
    for v, k in enumerate(["rho_L2", "momentum_x_L2", "momentum_y_L2", "energy_L2"]):
        assert abs(err[v] - float(summ[k])) <= 1e-9 * float(summ[k])


@pytest.mark.parametrize("precision", ["mixed", "mixed-fp16", "mixed-bf16"])
def test_mixed_update_error_vs_fp64_256(precision, torch_cuda):
    """256^2, k=8: two BE Newton updates with FP32 factors (stored in FP32, in
    FP16 with a per-block scale of U, or in BF16) + FP32 inner GMRES; each update's
    direction is compared with the FP64 path's solution of the same system
    (precision_error) and the RMS sequence with the reference's."""
    from dg_hdg_euler_cpp_amd import DeviceContext

    run = GOLD["runs"]["k8_256_be"]
    args = list(run["args"])
    args[args.index("--newton-steps") + 1] = "2"
    args[args.index("--warmup-updates") + 1] = "0"
    args = [a for a in args if a not in ("--method", "dg")]
    with DeviceContext(args + ["--method", "dg", "--precision", precision, "--check-update"]) as dev:
        u = dev.initial()
        st = dev.advance(u)
    for it, got in enumerate(st["updates"]):
        w = float(run["iterations"][it]["residual_physical_rms"])
        print(f"{precision} update {it}: inner Krylov {got['krylov_iterations']}, "
              f"error vs FP64 {got['precision_error']:.3e}", flush=True)
        assert abs(got["residual_physical_rms"] - w) <= [1e-13, 1e-10][it] * w
        assert 0 <= got["precision_error"] <= 1e-9, got["precision_error"]


@pytest.mark.parametrize("precision", ["mixed", "mixed-fp16"])
def test_mixed_512_config(precision, torch_cuda):
    """512^2, k=8 (the mixed-precision configuration): two BE Newton updates with
    FP32 factors, stored in FP32 or in FP16 with a per-block scale, and the FP32
    inner GMRES(30) inside FP64 FGMRES(20); the RMS sequence of the reference's
    512^2 run (k8_512_be) and, per update, the direction's error against the
    FP64 path (the 2-byte and FP32 factors are released before the FP64 factors
    are built: 111 GB, then 221 GB)."""
    from dg_hdg_euler_cpp_amd import DeviceContext

    run = GOLD["runs"]["k8_512_be"]
    # the reference's speed-run command line (speed_test/2d/benchmark.py:26-36) at 512^2
    args = ["--nx", "512", "--ny", "512", "--k", "8", "--method", "dg", "--integrator", "be", "--steps", "1",
            "--newton-steps", "2", "--cfl", "5", "--ntol", "1e-12", "--ltol", "1e-12", "--restart", "20",
            "--max-linear", "1500", "--safety", "1.1", "--timing-only"]
    print(f"\n[case] 512 {precision} {args}", flush=True)
    with DeviceContext(args + ["--precision", precision, "--check-update"]) as dev:
        u = dev.initial()
        st = dev.advance(u)
    for it, got in enumerate(st["updates"]):
        w = float(run["iterations"][it]["residual_physical_rms"])
        print(f"update {it}: rms {got['residual_physical_rms']:.12e} (reference {w:.12e}), inner Krylov "
              f"{got['krylov_iterations']}, error vs FP64 {got['precision_error']:.3e}", flush=True)
        # RMS_0 sums 85M squared residual entries: the reference's vectorised
        # sequential sum carries about sqrt(N / 8) eps ~ 3e-13 of rounding
        assert abs(got["residual_physical_rms"] - w) <= [5e-13, 1e-10][it] * w
        assert 0 <= got["precision_error"] <= 1e-9, got["precision_error"]


@pytest.mark.parametrize("precision", ["mixed-fp16", "mixed-bf16"])
@pytest.mark.parametrize("k,nx,state", [(1, 6, "seeded"), (2, 8, "seeded"), (4, 10, "diagonal"), (8, 12, "seeded")])
def test_two_byte_layouts_agree(precision, k, nx, state, torch_cuda, monkeypatch):
    """The column-pair layout of the 2-byte factors (the default) and the
    row-pair layout (DGH_BJ_ROWPAIRS=1, read when the factors are built) apply
    every row's updates in the same order, so the block-Jacobi solves are equal
    to the last bit."""
    from dg_hdg_euler_cpp_amd import DeviceContext

    torch = torch_cuda
    args = ["--nx", str(nx), "--ny", str(nx), "--k", str(k), "--testcase", state, "--precision", precision,
            "--integrator", "be", "--cfl", "0.1"]
    out = {}
    for layout in ("rows", "columns"):
        if layout == "rows":
            monkeypatch.setenv("DGH_BJ_ROWPAIRS", "1")
        else:
            monkeypatch.delenv("DGH_BJ_ROWPAIRS", raising=False)
        print(f"\n[case] {precision} k={k} {nx}^2 {state} {layout}", flush=True)
        with DeviceContext(args) as dev:
            u = dev.seeded(5) if state == "seeded" else dev.initial()
            rate = dev.freeze(u)
            dev.set_step(0.1 / rate, u)
            dev.dg_refresh(u)
            dev.bj_refresh_mixed()
            g = torch.Generator(device="cuda")
            g.manual_seed(11)
            x = torch.randn(dev.size, dtype=torch.float32, device="cuda", generator=g)
            out[layout] = dev.bj_apply_f32(x).clone()
    assert torch.equal(out["rows"], out["columns"])
    assert bool(torch.isfinite(out["columns"]).all())
