"""dgh_newton_update: one damped Newton update of the current stage through the
C ABI. A sequence of calls from the frozen state reproduces the updates of
dgh_advance (the same loop body, time_integrator.cpp:92-179) bit for bit: the
same residual RMS before each update, Krylov counts, damping and final state.
This is the step the throughput benchmark repeats."""
import pytest

pytestmark = pytest.mark.gpu

FORMS = [("dg", "fp64"), ("dg", "mixed"), ("dg", "mixed-fp16"), ("hdg", "fp64")]


@pytest.mark.parametrize("method,precision", FORMS, ids=[f"{m}-{p}" for m, p in FORMS])
def test_newton_updates_reproduce_advance(method, precision, torch_cuda):
    from dg_hdg_euler_cpp_amd import DeviceContext

    updates = 3
    args = ["--nx", "12", "--ny", "10", "--k", "4", "--testcase", "wave", "--integrator", "be", "--steps", "1",
            "--newton-steps", str(updates), "--cfl", "5", "--restart", "20", "--method", method,
            "--precision", precision]
    print(f"\n[case] newton_update {args}", flush=True)
    with DeviceContext(args) as dev:
        ua = dev.initial()
        st = dev.advance(ua)
        u0 = dev.initial()
        rate = dev.freeze(u0)
        dev.set_step(5.0 / rate, u0)  # BE: tau = CFL / rate, base state u0
        u = u0.clone()
        for k in range(updates):
            s = dev.newton_update(u)
            want = st["updates"][k]
            assert s["residual_physical_rms"] == want["residual_physical_rms"], (k, s, want)
            assert s["krylov_iterations"] == want["krylov_iterations"]
            assert s["damping"] == want["damping"]
            assert s["iteration_s"] > 0
        assert torch_cuda.equal(u, ua)
        # the call refuses a stage without a step size
    with DeviceContext(args) as dev:
        from dg_hdg_euler_cpp_amd.device import DeviceError

        u = dev.initial()
        dev.freeze(u)
        with pytest.raises(DeviceError, match="Set the time step"):
            dev.newton_update(u)
