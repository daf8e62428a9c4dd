"""dgh_hdg_apply_noalloc, the Schur apply of dgh_hdg_apply (hybrid.cpp:73-82)
with its intermediate vectors in the library's fixed per-GPU scratch instead of
context buffers: called through the C ABI, it returns dgh_hdg_apply's result
bit for bit (the same kernels), on two calls in a row whose second mesh is
smaller than the first (the scratch keeps the first call's contents), and it
refuses a mesh whose intermediates exceed the fixed maximum."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def condensed(args, torch):
    from dg_hdg_euler_cpp_amd import DeviceContext

    dev = DeviceContext(args + ["--cfl", "5", "--integrator", "be"])
    u = dev.initial()
    rate = dev.freeze(u)
    dev.set_step(5.0 / rate, u)
    dev.hdg_refresh_projected(u, dev.dg_residual(u))
    return dev


def noalloc(dev, x):
    y = dev.tzeros()
    dev._call("dgh_hdg_apply_noalloc", dev._tv(x), dev._tv(y))
    return y


def test_noalloc_apply_matches_hdg_apply(torch_cuda):
    # first call on the larger mesh, then a smaller one with the scratch
    # still holding the first call's vectors
    for args, seed in ((["--nx", "12", "--ny", "9", "--k", "8"], 1), (["--nx", "5", "--ny", "4", "--k", "3"], 2)):
        print(f"\n[case] hdg_apply_noalloc {args}", flush=True)
        with condensed(args, torch_cuda) as dev:
            rng = np.random.default_rng(seed)
            x = torch_cuda.tensor(rng.standard_normal(dev.trace_size), device="cuda")
            want = dev.hdg_apply(x)
            got = noalloc(dev, x)
            assert torch_cuda.equal(got, want)
            assert float(want.abs().max()) > 0
            # and again on the same context: nothing is left over from the first call
            assert torch_cuda.equal(noalloc(dev, 2 * x), dev.hdg_apply(2 * x))


def test_noalloc_apply_refuses_meshes_beyond_the_scratch(torch_cuda):
    from dg_hdg_euler_cpp_amd import DeviceContext
    from dg_hdg_euler_cpp_amd.device import DeviceError

    # 200 x 200 at k = 8: 2*40000*36 + 40000*144 = 8.64M doubles > 8388608
    print("\n[case] hdg_apply_noalloc oversize 200x200 k8", flush=True)
    with DeviceContext(["--nx", "200", "--ny", "200", "--k", "8"]) as dev:
        x = dev.tzeros()
        with pytest.raises(DeviceError, match="beyond the fixed maximum of 8388608"):
            noalloc(dev, x)


def test_noalloc_apply_needs_a_condensation(torch_cuda):
    from dg_hdg_euler_cpp_amd import DeviceContext
    from dg_hdg_euler_cpp_amd.device import DeviceError

    print("\n[case] hdg_apply_noalloc before refresh", flush=True)
    with DeviceContext(["--nx", "4", "--ny", "4", "--k", "2"]) as dev:
        with pytest.raises(DeviceError, match="Refresh the HDG condensation"):
            noalloc(dev, dev.tzeros())
