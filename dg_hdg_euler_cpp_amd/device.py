"""ctypes binding of the C ABI in ``include/dgh/dgh_device.h``.

Vectors are torch CUDA (HIP) tensors of dtype float64 whose device pointers are
handed to the library; torch is only memory plumbing here. The binding fails
loudly when ``libdgh_device.so`` is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libdgh_device.so")
_lib = None

_D = C.POINTER(C.c_double)


class DeviceError(RuntimeError):
    """An error reported by libdgh_device.so (reference message text)."""


class UpdateStats(C.Structure):
    _fields_ = [("step", C.c_int), ("stage", C.c_int), ("iteration", C.c_int), ("krylov_iterations", C.c_int),
                ("time", C.c_double), ("stage_time", C.c_double), ("dt", C.c_double), ("cfl", C.c_double),
                ("residual_physical_rms", C.c_double), ("damping", C.c_double),
                ("linear_relative_residual", C.c_double), ("dg_linear_relative_residual", C.c_double),
                ("iteration_s", C.c_double), ("inner_iterations", C.c_int), ("outer_iterations", C.c_int),
                ("precision_error", C.c_double)]


class RunStats(C.Structure):
    _fields_ = [("final_time", C.c_double), ("conservation_drift", C.c_double), ("seconds", C.c_double),
                ("max_linear_residual", C.c_double), ("max_dg_linear_residual", C.c_double),
                ("minimum_step", C.c_double), ("maximum_step", C.c_double),
                ("measured_newton_seconds", C.c_double), ("steps", C.c_int), ("newton_iterations", C.c_int),
                ("krylov_iterations", C.c_int), ("measured_updates", C.c_int)]


class AdjointStats(C.Structure):
    _fields_ = [("krylov_iterations", C.c_int), ("linear_solves", C.c_int), ("max_linear_residual", C.c_double)]


def load_library():
    """Load libdgh_device.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DeviceError(f"{LIB_PATH} is missing: build it with `python -m dg_hdg_euler_cpp_amd.build`")
    L = C.CDLL(LIB_PATH)
    L.dgh_last_error.restype = C.c_char_p
    L.dgh_last_error.argtypes = [C.c_void_p]
    L.dgh_config_size.restype = C.c_size_t
    L.dgh_create.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
    L.dgh_destroy.argtypes = [C.c_void_p]
    L.dgh_config_parse.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.c_void_p, C.c_char_p, C.c_size_t]
    _lib = L
    return L


def _ptr(t, dtype, numel):
    """Device pointer of a torch tensor, as c_void_p, after checking it against
    what the C entry point assumes: a contiguous tensor on the HIP device with
    exactly `numel` entries of `dtype`. The kernels index the raw pointer, so a
    CPU tensor, another dtype or a shorter buffer would be read or written out
    of bounds; they are refused here instead. None passes through (optional
    arguments)."""
    if t is None:
        return None
    if not getattr(t, "is_cuda", False):
        raise DeviceError("device vectors must be tensors on the HIP device")
    if t.dtype != dtype:
        raise DeviceError(f"device vector has dtype {t.dtype}, expected {dtype}")
    if t.numel() != numel:
        raise DeviceError(f"device vector has {t.numel()} entries, expected {numel}")
    if not t.is_contiguous():
        raise DeviceError("device vectors must be contiguous")
    return C.c_void_p(t.data_ptr())


def parse_config(args):
    """Reference CLI flags -> opaque dgh_config buffer."""
    L = load_library()
    argv = ["dg_hdg_2d"] + [str(a) for a in args]
    arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
    buf = C.create_string_buffer(L.dgh_config_size())
    err = C.create_string_buffer(512)
    if L.dgh_config_parse(len(argv), arr, buf, err, 512) != 0:
        raise DeviceError(err.value.decode())
    return buf


class DeviceContext:
    """One device-resident Problem + operators (the reference's Problem,
    DGOperator, BlockJacobi, GMRES and TimeIntegrator behind the C ABI)."""

    def __init__(self, args):
        import torch  # memory plumbing only

        self._torch = torch
        L = load_library()
        cfg = parse_config(args)
        h = C.c_void_p()
        if L.dgh_create(cfg, C.byref(h)) != 0:
            raise DeviceError(L.dgh_last_error(None).decode())
        self.h = h
        dims = (C.c_int64 * 10)()
        L.dgh_dims(self.h, dims)
        self.size, self.trace_size, self.cells, self.n, self.faces, self.m, self.fw, self.nx, self.ny, _ = list(dims)
        geo = (C.c_double * 5)()
        L.dgh_geometry(self.h, geo)
        self.hx, self.hy, self.lx, self.ly = geo[0], geo[1], geo[2], geo[3]

    def close(self):
        if getattr(self, "h", None):
            load_library().dgh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- helpers: checked device pointers of the vectors each entry point takes
    def _v(self, t):
        return _ptr(t, self._torch.float64, self.size)

    def _v32(self, t):
        return _ptr(t, self._torch.float32, self.size)

    def _tv(self, t):
        return _ptr(t, self._torch.float64, self.trace_size)

    def _speeds(self, t):
        return _ptr(t, self._torch.float64, 2 * self.faces * self.n)

    def _call(self, name, *args):
        L = load_library()
        self._torch.cuda.synchronize()
        rc = getattr(L, name)(self.h, *args)
        if rc != 0:
            raise DeviceError(L.dgh_last_error(self.h).decode())

    def zeros(self, n=None):
        return self._torch.zeros(self.size if n is None else n, dtype=self._torch.float64, device="cuda")

    def device_bytes(self):
        dims = (C.c_int64 * 10)()
        load_library().dgh_dims(self.h, dims)
        return int(dims[9])

    def basis(self):
        x = (C.c_double * self.n)()
        w = (C.c_double * self.n)()
        load_library().dgh_basis(self.h, x, w)
        return list(x), list(w)

    # ---- state / Problem
    def initial(self):
        u = self.zeros()
        self._call("dgh_initial", self._v(u))
        return u

    def seeded(self, seed):
        u = self.zeros()
        self._call("dgh_seeded", C.c_uint64(seed), self._v(u))
        return u

    def freeze(self, u):
        rate = C.c_double()
        self._call("dgh_freeze", self._v(u), C.byref(rate))
        return rate.value

    def speeds(self):
        out = self.zeros(2 * self.faces * self.n)
        self._call("dgh_speeds", self._speeds(out))
        return out

    def set_step(self, tau, old=None):
        self._call("dgh_set_step", C.c_double(tau), self._v(old))

    def dg_residual(self, u, out=None):
        r = self.zeros() if out is None else out
        self._call("dgh_dg_residual", self._v(u), self._v(r))
        return r

    def admissible(self, u):
        ok = C.c_int()
        self._call("dgh_admissible", self._v(u), C.byref(ok))
        return bool(ok.value)

    def residual_norm(self, r):
        out = C.c_double()
        self._call("dgh_residual_norm", self._v(r), C.byref(out))
        return out.value

    def dual_norm(self, x):
        out = C.c_double()
        self._call("dgh_dual_norm", self._v(x), C.byref(out))
        return out.value

    def state_norm(self, x):
        out = C.c_double()
        self._call("dgh_state_norm", self._v(x), C.byref(out))
        return out.value

    def integral(self, u):
        out = (C.c_double * 4)()
        self._call("dgh_integral", self._v(u), out)
        return list(out)

    def error_l2(self, u, t):
        out = (C.c_double * 4)()
        self._call("dgh_error_l2", self._v(u), C.c_double(t), out)
        return list(out)

    # ---- DG operator / block Jacobi / GMRES
    def dg_refresh(self, u):
        self._call("dgh_dg_refresh", self._v(u))

    def dg_apply(self, x, out=None):
        y = self.zeros() if out is None else out
        self._call("dgh_dg_apply", self._v(x), self._v(y))
        return y

    def dg_apply_transpose(self, w, out=None):
        z = self.zeros() if out is None else out
        self._call("dgh_dg_apply_transpose", self._v(w), self._v(z))
        return z

    def dg_block(self, row):
        import numpy as np

        out = np.zeros(self.m * self.m)
        self._call("dgh_dg_block", C.c_int(row), out.ctypes.data_as(_D))
        return out.reshape(self.m, self.m, order="F")

    def bj_refresh(self):
        self._call("dgh_bj_refresh")

    def bj_apply(self, r, out=None):
        z = self.zeros() if out is None else out
        self._call("dgh_bj_apply", self._v(r), self._v(z))
        return z

    def ilu_refresh(self):
        self._call("dgh_ilu_refresh")

    def ilu_apply(self, r, out=None):
        z = self.zeros() if out is None else out
        self._call("dgh_ilu_apply", self._v(r), self._v(z))
        return z

    def bj_refresh_mixed(self):
        self._call("dgh_bj_refresh_mixed")

    def bj_apply_f32(self, r, out=None):
        z = self._torch.zeros(self.size, dtype=self._torch.float32, device="cuda") if out is None else out
        self._call("dgh_bj_apply_f32", self._v32(r), self._v32(z))
        return z

    def dg_apply_f32(self, x, out=None):
        y = self._torch.zeros(self.size, dtype=self._torch.float32, device="cuda") if out is None else out
        self._call("dgh_dg_apply_f32", self._v32(x), self._v32(y))
        return y

    def dg_fgmres(self, rhs, max_iterations, tolerance):
        x = self.zeros()
        its = C.c_int()
        outer = C.c_int()
        rel = C.c_double()
        self._call("dgh_dg_fgmres", self._v(rhs), self._v(x), C.c_int(max_iterations), C.c_double(tolerance),
                   C.byref(its), C.byref(outer), C.byref(rel))
        return x, its.value, outer.value, rel.value

    def dg_gmres(self, rhs, max_iterations, tolerance):
        x = self.zeros()
        res = self.zeros()
        its = C.c_int()
        rel = C.c_double()
        self._call("dgh_dg_gmres", self._v(rhs), self._v(x), C.c_int(max_iterations), C.c_double(tolerance),
                   C.byref(its), C.byref(rel), self._v(res))
        return x, its.value, rel.value, res

    # ---- condensed HDG
    def tzeros(self):
        return self.zeros(self.trace_size)

    def hdg_refresh_projected(self, u, r):
        self._call("dgh_hdg_refresh_projected", self._v(u), self._v(r))

    def hdg_rhs(self):
        out = self.tzeros()
        self._call("dgh_hdg_rhs", self._tv(out))
        return out

    def hdg_apply(self, x, out=None):
        y = self.tzeros() if out is None else out
        self._call("dgh_hdg_apply", self._tv(x), self._tv(y))
        return y

    def hdg_block(self, face):
        import numpy as np

        out = np.zeros(self.fw * self.fw)
        self._call("dgh_hdg_block", C.c_int(face), out.ctypes.data_as(_D))
        return out.reshape(self.fw, self.fw, order="F")

    def hdg_bj_refresh(self):
        self._call("dgh_hdg_bj_refresh")

    def hdg_bj_apply(self, r, out=None):
        z = self.tzeros() if out is None else out
        self._call("dgh_hdg_bj_apply", self._tv(r), self._tv(z))
        return z

    def hdg_recover(self, dl, out=None):
        du = self.zeros() if out is None else out
        self._call("dgh_hdg_recover", self._tv(dl), self._v(du))
        return du

    def hdg_recovered_residual(self, e, out=None):
        r = self.zeros() if out is None else out
        self._call("dgh_hdg_recovered_residual", self._tv(e), self._v(r))
        return r

    def hdg_gmres(self, rhs, max_iterations, tolerance):
        x = self.tzeros()
        res = self.tzeros()
        its = C.c_int()
        rel = C.c_double()
        self._call("dgh_hdg_gmres", self._tv(rhs), self._tv(x), C.c_int(max_iterations), C.c_double(tolerance),
                   C.byref(its), C.byref(rel), self._tv(res))
        return x, its.value, rel.value, res

    def _run(self, entry, u, capacity):
        run = RunStats()
        upd = (UpdateStats * capacity)()
        count = C.c_int()
        self._call(entry, self._v(u), C.byref(run), upd, C.c_int(capacity), C.byref(count))
        stats = {f: getattr(run, f) for f, _ in RunStats._fields_}
        stats["updates"] = [{f: getattr(upd[i], f) for f, _ in UpdateStats._fields_}
                            for i in range(min(count.value, capacity))]
        return stats

    def advance(self, u, capacity=100000):
        return self._run("dgh_advance", u, capacity)

    def newton_update(self, u):
        """One damped Newton update of the current stage (speeds frozen, step
        set); u is overwritten with the accepted iterate."""
        st = UpdateStats()
        self._call("dgh_newton_update", self._v(u), C.byref(st))
        return {f: getattr(st, f) for f, _ in UpdateStats._fields_}

    # ---- discrete adjoint
    def bj_refresh_transposed(self):
        self._call("dgh_bj_refresh_transposed")

    def bj_apply_transposed(self, r, out=None):
        z = self.zeros() if out is None else out
        self._call("dgh_bj_apply_transposed", self._v(r), self._v(z))
        return z

    def dg_gmres_transposed(self, rhs, max_iterations, tolerance):
        x = self.zeros()
        res = self.zeros()
        its = C.c_int()
        rel = C.c_double()
        self._call("dgh_dg_gmres_transposed", self._v(rhs), self._v(x), C.c_int(max_iterations), C.c_double(tolerance),
                   C.byref(its), C.byref(rel), self._v(res))
        return x, its.value, rel.value, res

    def trajectory_record(self, u, capacity=100000):
        """advance() that checkpoints every stage for adjoint_backward."""
        return self._run("dgh_trajectory_record", u, capacity)

    def trajectory_replay(self, u, capacity=100000):
        """advance() on the recorded step sizes and frozen speeds."""
        return self._run("dgh_trajectory_replay", u, capacity)

    def adjoint_backward(self, final_gradient):
        grad = self.zeros()
        st = AdjointStats()
        self._call("dgh_adjoint_backward", self._v(final_gradient), self._v(grad), C.byref(st))
        return grad, {f: getattr(st, f) for f, _ in AdjointStats._fields_}

    def quadrant_density(self, u, with_gradient=True):
        value = C.c_double()
        g = self.zeros() if with_gradient else None
        self._call("dgh_quadrant_density", self._v(u), C.byref(value), self._v(g) if g is not None else None)
        return (value.value, g) if with_gradient else value.value

    # ---- timing hooks
    def set_timing(self, enabled):
        self._call("dgh_set_timing", C.c_int(1 if enabled else 0))

    def reset_kernel_times(self):
        self._call("dgh_reset_kernel_times")

    def kernel_time(self, name):
        ms = C.c_double()
        launches = C.c_longlong()
        self._call("dgh_kernel_time", name.encode(), C.byref(ms), C.byref(launches))
        return ms.value, launches.value


class Dgh1Config(C.Structure):
    """dgh1_config of include/dgh/dgh1d_device.h (1D euler::Config)."""
    _fields_ = [("cells", C.c_int), ("degree", C.c_int), ("quadrature", C.c_int), ("restart", C.c_int),
                ("max_linear", C.c_int), ("max_newton", C.c_int), ("steps", C.c_int), ("newton_steps", C.c_int),
                ("warmup_updates", C.c_int), ("cfl", C.c_double), ("final_time", C.c_double),
                ("safety", C.c_double), ("nonlinear_tolerance", C.c_double), ("linear_tolerance", C.c_double),
                ("integrator", C.c_int), ("method", C.c_int), ("timing_only", C.c_int), ("prefix", C.c_char * 512),
                ("testcase", C.c_int)]


def _lib1():
    L = load_library()
    if not getattr(L, "_dgh1_ready", False):
        L.dgh1_create.restype = C.c_void_p
        L.dgh1_create.argtypes = [C.POINTER(Dgh1Config)]
        L.dgh1_create_error.restype = C.c_char_p
        L.dgh1_last_error.restype = C.c_char_p
        L.dgh1_last_error.argtypes = [C.c_void_p]
        L.dgh1_destroy.argtypes = [C.c_void_p]
        L.dgh1_size.restype = C.c_longlong
        L.dgh1_size.argtypes = [C.c_void_p]
        L.dgh1_parse_config.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.POINTER(Dgh1Config)]
        L._dgh1_ready = True
    return L


def parse_config_1d(args):
    """Reference 1D CLI flags (dg_hdg_1d/src/config.cpp) -> Dgh1Config."""
    L = _lib1()
    argv = [str(a) for a in args]
    arr = (C.c_char_p * max(1, len(argv)))(*[a.encode() for a in argv])
    cfg = Dgh1Config()
    if L.dgh1_parse_config(len(argv), arr, C.byref(cfg)) != 0:
        raise DeviceError(L.dgh1_create_error().decode())
    return cfg


class DeviceContext1D:
    """The 1D periodic Euler DG path (dg_hdg_1d) behind include/dgh/dgh1d_device.h:
    Problem, DG operator, block Jacobi, GMRES and TimeIntegrator on the device."""

    def __init__(self, args):
        import torch  # memory plumbing only

        self._torch = torch
        L = _lib1()
        self.config = parse_config_1d(args)
        self.h = L.dgh1_create(C.byref(self.config))
        if not self.h:
            raise DeviceError(L.dgh1_create_error().decode())
        self.h = C.c_void_p(self.h)
        self.size = int(L.dgh1_size(self.h))
        self.n = self.config.degree + 1
        self.m = 3 * self.n
        self.cells = self.config.cells

    def close(self):
        if getattr(self, "h", None):
            _lib1().dgh1_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _v(self, t):
        return _ptr(t, self._torch.float64, self.size)

    def _tv(self, t):
        return _ptr(t, self._torch.float64, 3 * self.cells)

    def _call(self, name, *args):
        L = _lib1()
        # torch's fills and copies of the arguments run on torch's stream; the
        # library's own stream does not wait for them
        self._torch.cuda.synchronize()
        if getattr(L, name)(self.h, *args) != 0:
            raise DeviceError(L.dgh1_last_error(self.h).decode())

    def zeros(self):
        return self._torch.zeros(self.size, dtype=self._torch.float64, device="cuda")

    def initial(self):
        u = self.zeros()
        self._call("dgh1_initial", self._v(u))
        return u

    def freeze(self, u):
        rate = C.c_double()
        speeds = (C.c_double * (2 * self.cells))()
        self._call("dgh1_freeze", self._v(u), C.byref(rate), speeds)
        self.speeds = list(speeds)
        return rate.value

    def set_step(self, tau, old):
        self._call("dgh1_set_step", C.c_double(tau), self._v(old))

    def residual(self, u):
        r = self.zeros()
        self._call("dgh1_residual", self._v(u), self._v(r))
        return r

    def admissible(self, u):
        ok = C.c_int()
        self._call("dgh1_admissible", self._v(u), C.byref(ok))
        return bool(ok.value)

    def refresh(self, u):
        self._call("dgh1_refresh", self._v(u))

    def apply(self, x):
        y = self.zeros()
        self._call("dgh1_apply", self._v(x), self._v(y))
        return y

    def block(self, row):
        out = (C.c_double * (self.m * self.m))()
        self._call("dgh1_block", C.c_int(row), out)
        import numpy as np

        return np.array(out).reshape(self.m, self.m, order="F")

    def bj_refresh(self):
        self._call("dgh1_bj_refresh")

    def bj_apply(self, r):
        z = self.zeros()
        self._call("dgh1_bj_apply", self._v(r), self._v(z))
        return z

    def gmres(self, rhs, limit, tol):
        x = self.zeros()
        its, rel = C.c_int(), C.c_double()
        self._call("dgh1_gmres", self._v(rhs), self._v(x), C.c_int(limit), C.c_double(tol), C.byref(its), C.byref(rel))
        return x, its.value, rel.value

    def error_l2(self, u, t):
        e = (C.c_double * 3)()
        self._call("dgh1_error", self._v(u), C.c_double(t), e)
        return list(e)

    def integral(self, u):
        e = (C.c_double * 3)()
        self._call("dgh1_integral", self._v(u), e)
        return list(e)

    def tzeros(self):
        return self._torch.zeros(3 * self.cells, dtype=self._torch.float64, device="cuda")

    def hdg_refresh(self, u, r):
        self._call("dgh1_hdg_refresh", self._v(u), self._v(r))

    def hdg_rhs(self):
        y = self.tzeros()
        self._call("dgh1_hdg_rhs", self._tv(y))
        return y

    def hdg_apply(self, x):
        y = self.tzeros()
        self._call("dgh1_hdg_apply", self._tv(x), self._tv(y))
        return y

    def hdg_block(self, face):
        out = (C.c_double * 9)()
        self._call("dgh1_hdg_block", C.c_int(face), out)
        import numpy as np

        return np.array(out).reshape(3, 3, order="F")

    def hdg_bj_apply(self, r):
        z = self.tzeros()
        self._call("dgh1_hdg_bj_apply", self._tv(r), self._tv(z))
        return z

    def hdg_recover(self, dl):
        du = self.zeros()
        self._call("dgh1_hdg_recover", self._tv(dl), self._v(du))
        return du

    def hdg_recovered_residual(self, dl):
        out = self.zeros()
        self._call("dgh1_hdg_recovered_residual", self._tv(dl), self._v(out))
        return out

    def hdg_gmres(self, rhs, limit, tol):
        x = self.tzeros()
        its, rel = C.c_int(), C.c_double()
        self._call("dgh1_hdg_gmres", self._tv(rhs), self._tv(x), C.c_int(limit), C.c_double(tol), C.byref(its),
                   C.byref(rel))
        return x, its.value, rel.value

    def advance(self, u, capacity=100000):
        run = RunStats()
        ups = (UpdateStats * capacity)()
        count = C.c_int()
        self._call("dgh1_advance", self._v(u), C.byref(run), ups, C.c_int(capacity), C.byref(count))
        out = {k: getattr(run, k) for k, _ in RunStats._fields_}
        out["updates"] = [{k: getattr(ups[i], k) for k, _ in UpdateStats._fields_}
                          for i in range(min(count.value, capacity))]
        return out
