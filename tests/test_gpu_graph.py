"""GMRES with graph-captured Arnoldi steps (--graph) against the plain host-
driven loop: the same Krylov counts and the same solutions to rounding (the
Givens rotations run on the device instead of the host), across several
restart cycles (so every step graph is replayed), for the DG and HDG systems
and through whole Newton updates."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _solve(args, graph, torch):
    from dg_hdg_euler_cpp_amd import DeviceContext

    print(f"\n[case] {args} graph={graph}", flush=True)
    with DeviceContext(args + (["--graph"] if graph else ["--no-graph"])) as dev:
        u = dev.initial()
        rate = dev.freeze(u)
        dev.set_step(5.0 / rate, u)
        r = dev.dg_residual(u)
        dev.dg_refresh(u)
        dev.bj_refresh()
        x, its, relres, _ = dev.dg_gmres(-r, 1500, 1e-12)
        out = {"x": x.cpu().numpy(), "its": its, "rel": relres}
        dev.hdg_refresh_projected(u, r)
        dev.hdg_bj_refresh()
        xl, its_h, rel_h, _ = dev.hdg_gmres(dev.hdg_rhs(), 1500, 1e-12)
        out.update(xh=xl.cpu().numpy(), its_h=its_h, rel_h=rel_h)
    return out


@pytest.mark.parametrize("args", [["--nx", "16", "--ny", "16", "--k", "8"], ["--nx", "12", "--ny", "8", "--k", "4"]],
                         ids=["k8-16x16", "k4-12x8"])
def test_graph_gmres_matches_host_loop(args, torch_cuda):
    base = args + ["--cfl", "5", "--integrator", "be", "--restart", "20"]
    plain = _solve(base, False, torch_cuda)
    graph = _solve(base, True, torch_cuda)
    print(f"DG its {plain['its']} / {graph['its']}, HDG its {plain['its_h']} / {graph['its_h']}")
    assert plain["its"] > 20 and plain["its_h"] > 20   # several restart cycles
    assert abs(graph["its"] - plain["its"]) <= 1 and abs(graph["its_h"] - plain["its_h"]) <= 1
    assert graph["rel"] <= 1e-12 and graph["rel_h"] <= 1e-12
    # both solutions meet the same 1e-12 residual; they differ by rounding of
    # the rotations, amplified by the conditioning of the Newton operator
    assert rel(graph["x"], plain["x"]) <= 1e-9
    assert rel(graph["xh"], plain["xh"]) <= 1e-9


def test_graph_newton_updates_match(torch_cuda):
    from dg_hdg_euler_cpp_amd import DeviceContext

    args = ["--nx", "32", "--ny", "32", "--k", "8", "--method", "dg", "--integrator", "be", "--cfl", "5",
            "--steps", "1", "--newton-steps", "3", "--restart", "20"]
    runs = []
    for graph in (False, True):
        with DeviceContext(args + (["--graph"] if graph else ["--no-graph"])) as dev:
            u = dev.initial()
            runs.append(dev.advance(u))
    a = [x["residual_physical_rms"] for x in runs[0]["updates"]]
    b = [x["residual_physical_rms"] for x in runs[1]["updates"]]
    print("rms", a, b)
    for k, (x, y) in enumerate(zip(a, b)):
        assert abs(x - y) <= [1e-13, 1e-10, 1e-6][k] * x + 1e-12 * a[0]


def test_graph_newton_updates_with_ilu_match(torch_cuda):
    """The captured Arnoldi step with the block ILU(0) apply (two one-colour
    block solves, restrictions, the off-diagonal and red back-substitution
    kernels) against the host-driven loop, whole Newton updates."""
    from dg_hdg_euler_cpp_amd import DeviceContext

    args = ["--nx", "32", "--ny", "32", "--k", "8", "--method", "dg", "--integrator", "be", "--cfl", "5",
            "--steps", "1", "--newton-steps", "3", "--restart", "20", "--precond", "block-ilu0",
            "--ordering", "redblack"]
    runs = []
    for graph in (False, True):
        print(f"\n[case] ilu graph={graph}", flush=True)
        with DeviceContext(args + (["--graph"] if graph else ["--no-graph"])) as dev:
            u = dev.initial()
            runs.append(dev.advance(u))
    a = [x["residual_physical_rms"] for x in runs[0]["updates"]]
    b = [x["residual_physical_rms"] for x in runs[1]["updates"]]
    ka = [x["krylov_iterations"] for x in runs[0]["updates"]]
    kb = [x["krylov_iterations"] for x in runs[1]["updates"]]
    print("rms", a, b, "krylov", ka, kb)
    for k, (x, y) in enumerate(zip(a, b)):
        assert abs(x - y) <= [1e-13, 1e-10, 1e-6][k] * x + 1e-12 * a[0]
    assert all(abs(x - y) <= 1 for x, y in zip(ka, kb))


@pytest.mark.parametrize("precision", ["mixed", "mixed-fp16"])
def test_graph_mixed_newton_updates_match(precision, torch_cuda):
    """The inner FP32 GMRES of the mixed forms with captured Arnoldi steps
    (block-Jacobi apply in FP32 or from FP16 factors, FP32 Jv, one Gram-Schmidt
    pass, the Givens update on the device) against its host-driven loop: the
    same Newton residual sequence to rounding and the same outer and inner
    Krylov counts to one iteration, over updates that replay the graphs."""
    from dg_hdg_euler_cpp_amd import DeviceContext

    args = ["--nx", "32", "--ny", "32", "--k", "8", "--method", "dg", "--integrator", "be", "--cfl", "5",
            "--steps", "1", "--newton-steps", "3", "--restart", "20", "--precision", precision]
    runs = []
    for graph in (False, True):
        print(f"\n[case] {precision} graph={graph}", flush=True)
        with DeviceContext(args + (["--graph"] if graph else ["--no-graph"])) as dev:
            u = dev.initial()
            runs.append(dev.advance(u))
    a = [x["residual_physical_rms"] for x in runs[0]["updates"]]
    b = [x["residual_physical_rms"] for x in runs[1]["updates"]]
    ka = [(x["krylov_iterations"], x["outer_iterations"]) for x in runs[0]["updates"]]
    kb = [(x["krylov_iterations"], x["outer_iterations"]) for x in runs[1]["updates"]]
    print("rms", a, b, "krylov (inner, outer)", ka, kb)
    for k, (x, y) in enumerate(zip(a, b)):
        assert abs(x - y) <= [1e-13, 1e-8, 1e-4][k] * x + 1e-12 * a[0]
    assert all(abs(x[0] - y[0]) <= 1 and x[1] == y[1] for x, y in zip(ka, kb))


def test_graphs_follow_the_step_size(torch_cuda):
    """Captured graphs are kept across the linear solves of a stage and
    captured again when the step size changes: two SSP-DIRK3 steps (each with
    its own dt, four stages) of the mixed and FP64 forms agree with the
    host-driven loop to the solver tolerance."""
    from dg_hdg_euler_cpp_amd import DeviceContext

    for precision in ("fp64", "mixed"):
        args = ["--nx", "12", "--ny", "10", "--k", "4", "--testcase", "wave", "--integrator", "ssp-dirk3",
                "--steps", "2", "--cfl", "3", "--restart", "20", "--precision", precision]
        out = []
        for graph in (False, True):
            print(f"\n[case] dirk3 {precision} graph={graph}", flush=True)
            with DeviceContext(args + (["--graph"] if graph else ["--no-graph"])) as dev:
                u = dev.initial()
                st = dev.advance(u)
                out.append((u.cpu().numpy(), st))
        (ua, sa), (ub, sb) = out
        assert sa["steps"] == sb["steps"] == 2
        assert abs(sa["newton_iterations"] - sb["newton_iterations"]) <= 1
        assert rel(ub, ua) <= 1e-10, precision
