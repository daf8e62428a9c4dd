"""CPU-side checks of the C ABI: the library loads without a GPU, exports every
symbol include/dgh/dgh_device.h declares, and parses the reference CLI with the
reference's defaults and error messages (config.cpp:6-122)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "dgh", "dgh_device.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dgh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from dg_hdg_euler_cpp_amd import load_library

    L = load_library()
    syms = header_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_library_exports_every_1d_symbol():
    from dg_hdg_euler_cpp_amd import load_library

    text = open(os.path.join(ROOT, "include", "dgh", "dgh1d_device.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    syms = sorted(set(re.findall(r"\b(dgh1_[a-z0-9_]+)\s*\(", text)))
    assert len(syms) >= 18
    L = load_library()
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


@pytest.mark.parametrize("args,message", [
    (["--cells", "2"], "Invalid mesh or degree"),
    (["--k", "9"], "Invalid mesh or degree"),
    (["--k", "3", "--quad", "3"], "Quadrature must be >= degree+1"),
    (["--T", "0.1", "--steps", "3"], "Choose either --T or --steps"),
    (["--integrator", "rk4"], "Invalid integrator"),
    (["--method", "both2"], "Method must be dg, hdg, or both"),
    (["--ltol", "1"], "Invalid solver tolerance or wave-speed safety"),
    (["--bogus"], "Unknown option: --bogus"),
    (["--precond", "block-ilu0"], "The 1D device path preconditions with block Jacobi"),
])
def test_1d_config_errors_use_reference_messages(args, message):
    """dg_hdg_1d/src/config.cpp:6-116 messages and defaults (config.hpp:7-18)."""
    from dg_hdg_euler_cpp_amd.device import DeviceError, parse_config_1d

    c = parse_config_1d([])
    assert (c.cells, c.degree, c.restart, c.integrator, c.cfl, c.final_time) == (32, 3, 30, 1, 2.0, 0.2)
    with pytest.raises(DeviceError, match=re.escape(message)):
        parse_config_1d(args)


def parse(args):
    from dg_hdg_euler_cpp_amd.device import parse_config

    return parse_config(args)


class Cfg(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("degree", C.c_int), ("quadrature", C.c_int),
                ("restart", C.c_int), ("max_linear", C.c_int), ("max_newton", C.c_int), ("steps", C.c_int),
                ("newton_steps", C.c_int), ("warmup_updates", C.c_int), ("cfl", C.c_double),
                ("final_time", C.c_double), ("safety", C.c_double), ("nonlinear_tolerance", C.c_double),
                ("linear_tolerance", C.c_double), ("direct", C.c_int), ("colamd", C.c_int),
                ("timing_only", C.c_int), ("check_update", C.c_int), ("save_solution", C.c_int),
                ("integrator", C.c_int), ("method", C.c_int), ("testcase", C.c_int), ("precision", C.c_int),
                ("inner_restart", C.c_int), ("inner_max", C.c_int), ("inner_tolerance", C.c_double),
                ("seed", C.c_ulonglong), ("graph", C.c_int), ("precond", C.c_int), ("ordering", C.c_int),
                ("prefix", C.c_char * 512)]


def test_config_struct_layout_and_defaults():
    from dg_hdg_euler_cpp_amd import load_library

    L = load_library()
    assert L.dgh_config_size() == C.sizeof(Cfg)
    c = Cfg.from_buffer_copy(parse([]))
    # config.hpp:7-18
    assert (c.nx, c.ny, c.degree, c.quadrature, c.restart, c.max_linear, c.max_newton) == (8, 8, 3, 0, 30, 1500, 30)
    assert (c.cfl, c.final_time, c.safety, c.nonlinear_tolerance, c.linear_tolerance) == (2.0, 0.2, 1.1, 1e-12, 1e-12)
    assert c.integrator == 1 and c.method == 2 and c.testcase == 0 and c.prefix == b"output/run_2d"
    c = Cfg.from_buffer_copy(parse(["--nx", "256", "--k", "8", "--integrator", "be", "--steps", "1",
                                    "--precision", "mixed", "--device", "gpu", "--restart", "20"]))
    assert (c.nx, c.degree, c.integrator, c.steps, c.precision, c.restart) == (256, 8, 0, 1, 1, 20)
    assert c.graph == 1 and Cfg.from_buffer_copy(parse(["--no-graph"])).graph == 0
    assert Cfg.from_buffer_copy(parse(["--no-graph", "--graph"])).graph == 1
    # 2-byte factor storage: BF16, or FP16 with a per-block scale of U
    assert Cfg.from_buffer_copy(parse(["--precision", "mixed-bf16"])).precision == 2
    assert Cfg.from_buffer_copy(parse(["--precision", "mixed-fp16"])).precision == 3
    # device block ILU(0): DG, FP64, red-black order on an even grid
    c = Cfg.from_buffer_copy(parse(["--precond", "block-ilu0", "--ordering", "redblack", "--method", "dg"]))
    assert (c.precond, c.ordering, c.colamd) == (1, 2, 0)
    c = Cfg.from_buffer_copy(parse(["--ordering", "colamd"]))
    assert (c.precond, c.ordering, c.colamd) == (0, 1, 1)


@pytest.mark.parametrize("args,message", [
    (["--k", "9"], "Invalid mesh or degree"),
    (["--T", "0.1", "--steps", "2"], "Choose either --T or --steps"),
    (["--precond", "ilu"], "Invalid block preconditioner"),
    (["--precond", "block-ilu0", "--method", "dg"],
     "Block ILU(0) on the device runs in red-black order: use --ordering redblack"),
    (["--precond", "block-ilu0", "--ordering", "redblack", "--method", "dg", "--nx", "5"],
     "Block ILU(0) on the device runs in red-black order: nx and ny must be even and >= 4"),
    (["--precond", "block-ilu0", "--ordering", "redblack"], "Block ILU(0) on the device is built for --method dg"),
    (["--precond", "block-ilu0", "--ordering", "redblack", "--method", "dg", "--precision", "mixed-bf16"],
     "Block ILU(0) on the device is built for --precision fp64 and mixed"),
    (["--precond", "block-ilu0", "--ordering", "redblack", "--method", "dg", "--precision", "mixed-fp16"],
     "Block ILU(0) on the device is built for --precision fp64 and mixed"),
    (["--ordering", "redblack"], "Red-black ordering applies to --precond block-ilu0"),
    (["--precision", "fp8"], "Invalid precision"),
    (["--linear", "lu"], "Invalid linear solver"),
    (["--integrator", "rk4"], "Invalid integrator"),
    (["--method", "fem"], "Method must be dg, hdg, or both"),
    (["--ltol", "2"], "Invalid solver tolerance or wave-speed safety"),
    (["--bogus"], "Unknown option: --bogus"),
    (["--nx"], "Missing option value"),
    (["--nx", "4x"], "Invalid integer"),
    (["--device", "cpu"], "Only --device gpu is built; the CPU path is the reference executable"),
])
def test_config_errors_use_reference_messages(args, message):
    from dg_hdg_euler_cpp_amd.device import DeviceError

    with pytest.raises(DeviceError) as e:
        parse(args)
    assert str(e.value) == message


def test_cli_reports_errors_like_the_reference():
    import subprocess

    exe = os.path.join(ROOT, "dg_hdg_euler_cpp_amd", "bin", "dg_hdg_2d")
    p = subprocess.run([exe, "--k", "9"], capture_output=True, text=True)
    assert p.returncode == 1 and p.stderr.strip() == "ERROR: Invalid mesh or degree"


def test_1d_cli_reports_errors_like_the_reference():
    """dg_hdg_1d/src/main.cpp:103-106: "ERROR: <message>" and exit status 1."""
    import subprocess

    exe = os.path.join(ROOT, "dg_hdg_euler_cpp_amd", "bin", "dg_hdg_1d")
    p = subprocess.run([exe, "--cells", "2"], capture_output=True, text=True)
    assert p.returncode == 1 and p.stderr.strip() == "ERROR: Invalid mesh or degree"
    p = subprocess.run([exe, "--precond", "block-ilu0"], capture_output=True, text=True)
    assert p.returncode == 1 and p.stderr.strip() == "ERROR: The 1D device path preconditions with block Jacobi"
    p = subprocess.run([exe, "--help"], capture_output=True, text=True)
    assert p.returncode == 0 and "--cells" in p.stdout


def test_vector_pointer_checks_refuse_mismatched_buffers():
    """The ctypes binding refuses, before any kernel runs, buffers that the
    kernels would index out of bounds: host tensors, another dtype, another
    length, non-contiguous views."""
    import torch

    from dg_hdg_euler_cpp_amd.device import DeviceError, _ptr

    class FakeDeviceTensor:
        is_cuda = True

        def __init__(self, dtype, n, contiguous=True):
            self.dtype, self.n, self.contiguous = dtype, n, contiguous

        def numel(self):
            return self.n

        def is_contiguous(self):
            return self.contiguous

        def data_ptr(self):
            return 4096

    assert _ptr(None, torch.float64, 10) is None
    assert _ptr(FakeDeviceTensor(torch.float64, 10), torch.float64, 10).value == 4096
    with pytest.raises(DeviceError, match="HIP device"):
        _ptr(torch.zeros(10, dtype=torch.float64), torch.float64, 10)
    with pytest.raises(DeviceError, match="dtype"):
        _ptr(FakeDeviceTensor(torch.float32, 10), torch.float64, 10)
    with pytest.raises(DeviceError, match="9 entries, expected 10"):
        _ptr(FakeDeviceTensor(torch.float64, 9), torch.float64, 10)
    with pytest.raises(DeviceError, match="contiguous"):
        _ptr(FakeDeviceTensor(torch.float64, 10, contiguous=False), torch.float64, 10)
