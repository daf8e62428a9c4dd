"""The two LU kernels of the k = 8 element blocks against each other
(block_jacobi.hpp:12-23; Eigen PartialPivLU semantics in both).

The default k_lu (18-column panels factorised in LDS, 2 / 4 trailing columns
per wavefront pass in FP64 / FP32) and k_lu_mfma (trailing update on MFMA,
selected with DGH_LU_MFMA=1, read at each launch) apply the same products in
the same order to every entry, so the block-Jacobi apply through either set of
factors returns the same vector entry for entry. The apply reads every factor
entry and the row permutation, so equal outputs on random right-hand sides pin
both; each kernel is pinned against the oracle by the block-Jacobi tests."""
import os

import pytest

pytestmark = pytest.mark.gpu

CASES = [(8, "diagonal", 5.0), (8, "seeded", 0.1), (6, "wave", 5.0)]


def _refresh_apply(dev, r, precision, mfma):
    if mfma:
        os.environ["DGH_LU_MFMA"] = "1"
    try:
        if precision == "fp64":
            dev.bj_refresh()
        else:
            dev.bj_refresh_mixed()
    finally:
        os.environ.pop("DGH_LU_MFMA", None)
    return dev.bj_apply(r) if precision == "fp64" else dev.bj_apply_f32(r)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("nx,state,cfl", CASES)
def test_lu_kernels_give_equal_factors(precision, nx, state, cfl, torch_cuda):
    from dg_hdg_euler_cpp_amd import DeviceContext

    torch = torch_cuda
    args = ["--nx", str(nx), "--ny", str(nx), "--k", "8", "--cfl", str(cfl), "--integrator", "be",
            "--testcase", state, "--precision", precision]
    print(f"\n[case] {args}", flush=True)
    dtype = torch.float64 if precision == "fp64" else torch.float32
    with DeviceContext(args) as dev:
        u = dev.seeded(2025) if state == "seeded" else dev.initial()
        rate = dev.freeze(u)
        dev.set_step(cfl / rate, u)
        dev.dg_refresh(u)
        gen = torch.Generator(device="cuda").manual_seed(11)
        for trial in range(2):
            r = torch.randn(dev.size, dtype=dtype, device="cuda", generator=gen)
            z_default = _refresh_apply(dev, r, precision, False).clone()
            z_mfma = _refresh_apply(dev, r, precision, True).clone()
            assert torch.isfinite(z_default).all()
            diff = (z_default != z_mfma).sum().item()
            print(f"trial {trial}: entries that differ {diff} of {z_default.numel()}", flush=True)
            assert torch.equal(z_default, z_mfma)
