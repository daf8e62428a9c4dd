"""ctypes binding of the CPU oracle (oracle/libdgh_oracle.so).

Test infrastructure: the oracle of the HIP path. Imported only by tests/,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB_PATH = os.path.join(_ROOT, "oracle", "libdgh_oracle.so")
_lib = None

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


def select_library(path):
    """Use another build of the oracle (the threaded one for the CPU baseline);
    call before the first use in the process."""
    global _LIB_PATH
    if _lib is not None and os.path.abspath(path) != os.path.abspath(_LIB_PATH):
        raise RuntimeError("the oracle library is already loaded in this process")
    _LIB_PATH = os.path.abspath(path)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(_LIB_PATH)
        L.orc_error.restype = C.c_char_p
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int, C.POINTER(C.c_char_p)]
        L.orc_destroy.argtypes = [C.c_void_p]
        for name in dir(L):
            pass
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_D)


class OracleError(RuntimeError):
    pass


class Oracle:
    """One reference ``Problem`` plus lazily built operators (CPU, FP64)."""

    def __init__(self, args):
        L = lib()
        argv = ["oracle"] + [str(a) for a in args]
        arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
        self.h = L.orc_create(len(argv), arr)
        if not self.h:
            raise OracleError(L.orc_error().decode())
        dims = np.zeros(8, dtype=np.int64)
        L.orc_dims(C.c_void_p(self.h), dims.ctypes.data_as(C.POINTER(C.c_longlong)))
        self.size, self.trace_size, self.cells, self.n, self.points, self.nx, self.ny, self.faces = [int(x) for x in dims]
        self.m = 4 * self.n * self.n
        geo = np.zeros(5)
        L.orc_geometry(C.c_void_p(self.h), _p(geo))
        self.hx, self.hy, self.lx, self.ly = geo[:4]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(C.c_void_p(self.h))
            self.h = None

    def _call(self, name, *args):
        L = lib()
        rc = getattr(L, name)(C.c_void_p(self.h), *args)
        if rc != 0:
            raise OracleError(L.orc_error().decode())

    # ---- state / setup
    def initial(self):
        u = np.zeros(self.size)
        self._call("orc_initial", _p(u))
        return u

    def seeded(self, seed):
        u = np.zeros(self.size)
        self._call("orc_seeded", C.c_ulonglong(seed), _p(u))
        return u

    def basis(self):
        n = self.n
        out = np.zeros(4 * n + n * n)
        lib().orc_basis(C.c_void_p(self.h), _p(out))
        return dict(x=out[:n], w=out[n:2 * n], left=out[2 * n:3 * n], right=out[3 * n:4 * n],
                    wd=out[4 * n:].reshape(n, n, order="F"))

    def freeze(self, u):
        rate = C.c_double()
        self._call("orc_freeze", _p(np.ascontiguousarray(u)), C.byref(rate))
        return rate.value

    def speeds(self):
        out = np.zeros(2 * self.faces * self.n)
        lib().orc_speeds(C.c_void_p(self.h), _p(out))
        return out.reshape(-1, 2)

    def set_step(self, tau, old=None):
        self._call("orc_set_step", C.c_double(tau), _p(np.ascontiguousarray(old)) if old is not None else None)

    def dg_residual(self, u):
        r = np.zeros(self.size)
        self._call("orc_dg_residual", _p(np.ascontiguousarray(u)), _p(r))
        return r

    def admissible(self, u):
        ok = C.c_int()
        self._call("orc_admissible", _p(np.ascontiguousarray(u)), C.byref(ok))
        return bool(ok.value)

    def residual_norm(self, r):
        out = C.c_double()
        self._call("orc_residual_norm", _p(np.ascontiguousarray(r)), C.byref(out))
        return out.value

    def integral(self, u):
        out = np.zeros(4)
        self._call("orc_integral", _p(np.ascontiguousarray(u)), _p(out))
        return out

    def error_l2(self, u, t):
        out = np.zeros(4)
        self._call("orc_error_l2", _p(np.ascontiguousarray(u)), C.c_double(t), _p(out))
        return out

    def trace(self, u):
        lam = np.zeros(self.trace_size)
        self._call("orc_trace", _p(np.ascontiguousarray(u)), _p(lam))
        return lam

    def hybrid_residual(self, u, lam):
        r = np.zeros(self.size)
        g = np.zeros(self.trace_size)
        self._call("orc_hybrid_residual", _p(np.ascontiguousarray(u)), _p(np.ascontiguousarray(lam)), _p(r), _p(g))
        return r, g

    # ---- DG linearisation
    def dg_refresh(self, u):
        self._call("orc_dg_refresh", _p(np.ascontiguousarray(u)))

    def dg_apply(self, x):
        y = np.zeros(self.size)
        self._call("orc_dg_apply", _p(np.ascontiguousarray(x)), _p(y))
        return y

    def dg_apply_transpose(self, w):
        z = np.zeros(self.size)
        self._call("orc_dg_apply_transpose", _p(np.ascontiguousarray(w)), _p(z))
        return z

    def dg_block(self, row, col):
        out = np.zeros(self.m * self.m)
        self._call("orc_dg_block", C.c_int(row), C.c_int(col), _p(out))
        return out.reshape(self.m, self.m, order="F")

    def volume_cache(self):
        out = np.zeros(self.cells * 3 * self.points * self.points)
        lib().orc_volume_cache(C.c_void_p(self.h), _p(out))
        return out

    def bj_refresh(self):
        self._call("orc_bj_refresh")

    def bj_apply(self, r):
        z = np.zeros(self.size)
        self._call("orc_bj_apply", _p(np.ascontiguousarray(r)), _p(z))
        return z

    def bj_factor(self, block):
        lu = np.zeros(self.m * self.m)
        piv = np.zeros(self.m, dtype=np.int32)
        self._call("orc_bj_factor", C.c_int(block), _p(lu), piv.ctypes.data_as(_I))
        return lu.reshape(self.m, self.m, order="F"), piv

    def dg_gmres(self, rhs, restart, limit, tol):
        x = np.zeros(self.size)
        res = np.zeros(self.size)
        its = C.c_int()
        rel = C.c_double()
        self._call("orc_dg_gmres", _p(np.ascontiguousarray(rhs)), _p(x), C.c_int(restart), C.c_int(limit),
                   C.c_double(tol), C.byref(its), C.byref(rel), _p(res))
        return x, its.value, rel.value, res

    # ---- block ILU(0) over the element graph (block_ilu.cpp); ordering "natural" or "redblack"
    def ilu_refresh(self, ordering="natural"):
        self._call("orc_ilu_refresh", C.c_int({"natural": 0, "redblack": 1}[ordering]))

    def ilu_apply(self, r):
        z = np.zeros(self.size)
        self._call("orc_ilu_apply", _p(np.ascontiguousarray(r)), _p(z))
        return z

    def ilu_gmres(self, rhs, restart, limit, tol):
        x = np.zeros(self.size)
        its = C.c_int()
        rel = C.c_double()
        self._call("orc_ilu_gmres", _p(np.ascontiguousarray(rhs)), _p(x), C.c_int(restart), C.c_int(limit),
                   C.c_double(tol), C.byref(its), C.byref(rel))
        return x, its.value, rel.value

    def dense_jacobian(self):
        out = np.zeros(self.size * self.size)
        self._call("orc_dense_jacobian", _p(out))
        return out.reshape(self.size, self.size, order="F")

    # ---- HDG
    def hdg_refresh_projected(self, u, r):
        self._call("orc_hdg_refresh_projected", _p(np.ascontiguousarray(u)), _p(np.ascontiguousarray(r)))

    def hdg_rhs(self):
        out = np.zeros(self.trace_size)
        lib().orc_hdg_rhs(C.c_void_p(self.h), _p(out))
        return out

    def hdg_apply(self, x):
        y = np.zeros(self.trace_size)
        self._call("orc_hdg_apply", _p(np.ascontiguousarray(x)), _p(y))
        return y

    def hdg_block(self, row, col):
        fw = 4 * self.n
        out = np.zeros(fw * fw)
        self._call("orc_hdg_block", C.c_int(row), C.c_int(col), _p(out))
        return out.reshape(fw, fw, order="F")

    def hdg_recover(self, dl):
        out = np.zeros(self.size)
        self._call("orc_hdg_recover", _p(np.ascontiguousarray(dl)), _p(out))
        return out

    def hdg_recovered_residual(self, e):
        out = np.zeros(self.size)
        self._call("orc_hdg_recovered_residual", _p(np.ascontiguousarray(e)), _p(out))
        return out

    def hdg_factors(self):
        t = 16 * self.n
        x = np.zeros(self.cells * self.m * t)
        s = np.zeros(self.cells * t * t)
        y = np.zeros(self.cells * self.m)
        lib().orc_hdg_factors(C.c_void_p(self.h), _p(x), _p(s), _p(y))
        return x, s, y

    def hdg_bj_refresh(self):
        self._call("orc_hdg_bj_refresh")

    def hdg_bj_apply(self, r):
        z = np.zeros(self.trace_size)
        self._call("orc_hdg_bj_apply", _p(np.ascontiguousarray(r)), _p(z))
        return z

    # ---- whole run
    def advance(self, method, u, cap=100000):
        u = np.array(u, dtype=np.float64, copy=True)
        run = np.zeros(12)
        upd = np.zeros((cap, 13))
        count = C.c_int()
        self._call("orc_advance", method.encode(), _p(u), _p(run), _p(upd), C.c_int(cap), C.byref(count))
        keys = ["final_time", "conservation_drift", "seconds", "max_linear_residual", "max_dg_linear_residual",
                "minimum_step", "maximum_step", "measured_newton_seconds", "steps", "newton_iterations",
                "krylov_iterations", "measured_updates"]
        stats = dict(zip(keys, run.tolist()))
        ukeys = ["step", "stage", "iteration", "krylov_iterations", "time", "stage_time", "dt", "cfl",
                 "residual_physical_rms", "damping", "linear_relative_residual", "dg_linear_relative_residual",
                 "iteration_s"]
        stats["updates"] = [dict(zip(ukeys, row.tolist())) for row in upd[:min(count.value, cap)]]
        return u, stats


class Oracle1D:
    """The 1D DG restatement (dg_hdg_1d/src): one ``Problem`` + DG operator +
    block Jacobi, vectors in the reference layout cell * 3n + v * n + i."""

    def __init__(self, args):
        L = lib()
        L.orc1_create.restype = C.c_void_p
        L.orc1_create.argtypes = [C.c_int, C.POINTER(C.c_char_p)]
        L.orc1_destroy.argtypes = [C.c_void_p]
        L.orc1_last_error.restype = C.c_char_p
        L.orc1_h.restype = C.c_double
        argv = [str(a) for a in args]
        arr = (C.c_char_p * max(1, len(argv)))(*[a.encode() for a in argv])
        self.h = L.orc1_create(len(argv), arr)
        if not self.h:
            raise OracleError(L.orc1_last_error().decode())
        self.size = L.orc1_size(C.c_void_p(self.h))
        self.n = L.orc1_nodes(C.c_void_p(self.h))
        self.m = 3 * self.n
        self.cells = self.size // self.m
        self.dx = L.orc1_h(C.c_void_p(self.h))

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().orc1_destroy(C.c_void_p(self.h))
            except TypeError:  # interpreter shutdown: the module globals are gone
                pass
            self.h = None

    def _call(self, name, *args):
        L = lib()
        rc = getattr(L, name)(C.c_void_p(self.h), *args)
        if rc != 0:
            raise OracleError(L.orc1_last_error().decode())

    def _vec(self, name, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.size)
        self._call(name, _p(x), _p(y))
        return y

    def initial(self):
        u = np.zeros(self.size)
        self._call("orc1_initial", _p(u))
        return u

    def freeze(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        rate = C.c_double()
        speeds = np.zeros(2 * self.cells)
        self._call("orc1_freeze", _p(u), C.byref(rate), _p(speeds))
        self.speeds = speeds.reshape(-1, 2)
        return rate.value

    def set_step(self, tau, old):
        old = np.ascontiguousarray(old, dtype=np.float64)
        self._call("orc1_set_step", C.c_double(tau), _p(old))

    def residual(self, u):
        return self._vec("orc1_residual", u)

    def admissible(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        return bool(lib().orc1_admissible(C.c_void_p(self.h), _p(u)))

    def refresh(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        self._call("orc1_refresh", _p(u))

    def apply(self, x):
        return self._vec("orc1_apply", x)

    def block(self, row, col):
        out = np.zeros(self.m * self.m)
        self._call("orc1_block", C.c_int(row), C.c_int(col), _p(out))
        return out.reshape(self.m, self.m, order="F")

    def bj_refresh(self):
        self._call("orc1_bj_refresh")

    def bj_apply(self, r):
        return self._vec("orc1_bj_apply", r)

    def gmres(self, rhs, restart, limit, tol):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        x = np.zeros(self.size)
        its, rel = C.c_int(), C.c_double()
        self._call("orc1_gmres", _p(rhs), _p(x), C.c_int(restart), C.c_int(limit), C.c_double(tol),
                   C.byref(its), C.byref(rel))
        return x, its.value, rel.value

    def error_l2(self, u, t):
        u = np.ascontiguousarray(u, dtype=np.float64)
        e = np.zeros(3)
        self._call("orc1_error", _p(u), C.c_double(t), _p(e))
        return e

    def integral(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        e = np.zeros(3)
        self._call("orc1_integral", _p(u), _p(e))
        return e

    def _trace(self, name, x, size):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(size)
        self._call(name, _p(x), _p(y))
        return y

    def hdg_refresh(self, u, r):
        u = np.ascontiguousarray(u, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        self._call("orc1_hdg_refresh", _p(u), _p(r))

    def hdg_rhs(self):
        y = np.zeros(3 * self.cells)
        self._call("orc1_hdg_rhs", _p(y))
        return y

    def hdg_apply(self, x):
        return self._trace("orc1_hdg_apply", x, 3 * self.cells)

    def hdg_block(self, row, col):
        out = np.zeros(9)
        self._call("orc1_hdg_block", C.c_int(row), C.c_int(col), _p(out))
        return out.reshape(3, 3, order="F")

    def hdg_bj_apply(self, r):
        return self._trace("orc1_hdg_bj_apply", r, 3 * self.cells)

    def hdg_recover(self, dl):
        return self._trace("orc1_hdg_recover", dl, self.size)

    def hdg_recovered_residual(self, dl):
        return self._trace("orc1_hdg_recovered_residual", dl, self.size)

    def advance(self, u, cap=100000, method="dg"):
        u = np.array(u, dtype=np.float64, copy=True)
        summary = np.zeros(9)
        upd = np.zeros((cap, 12))
        count = C.c_int()
        self._call("orc1_advance", C.c_int(1 if method == "hdg" else 0), _p(u), _p(summary), _p(upd),
                   C.c_int(cap), C.byref(count))
        keys = ["final_time", "steps", "newton_iterations", "krylov_iterations", "conservation_drift",
                "max_linear_residual", "max_dg_linear_residual", "minimum_step", "maximum_step"]
        stats = dict(zip(keys, summary.tolist()))
        ukeys = ["step", "stage", "iteration", "krylov_iterations", "time", "stage_time", "dt", "cfl",
                 "residual_physical_rms", "damping", "linear_relative_residual", "dg_linear_relative_residual"]
        stats["updates"] = [dict(zip(ukeys, row.tolist())) for row in upd[:min(count.value, cap)]]
        return u, stats


def gauss(n):
    x = np.zeros(n)
    w = np.zeros(n)
    rc = lib().orc_gauss(C.c_int(n), _p(x), _p(w))
    if rc:
        raise OracleError(lib().orc_error().decode())
    return x, w


def block_ilu_solve(adjacency, blocks, order, rhs):
    """Block ILU(0) (block_ilu.cpp) of the block matrix {(r, c): blocks[(r, c)]} on
    the graph `adjacency` (rows include their diagonal), rows in `order`; returns
    the preconditioner applied to rhs."""
    rows = len(adjacency)
    width = next(iter(blocks.values())).shape[0]
    cols = [sorted(set(a) | {r}) for r, a in enumerate(adjacency)]
    row_ptr = np.zeros(rows + 1, dtype=np.int32)
    for r in range(rows):
        row_ptr[r + 1] = row_ptr[r] + len(cols[r])
    col_idx = np.array([c for row in cols for c in row], dtype=np.int32)
    data = np.concatenate([np.asarray(blocks[(r, c)], dtype=np.float64).ravel(order="F")
                           for r in range(rows) for c in cols[r]])
    x = np.zeros(rows * width)
    rc = lib().orc_block_ilu_solve(C.c_int(rows), row_ptr.ctypes.data_as(_I), col_idx.ctypes.data_as(_I),
                                   C.c_int(width), _p(data), np.asarray(order, dtype=np.int32).ctypes.data_as(_I),
                                   _p(np.ascontiguousarray(rhs, dtype=np.float64)), _p(x))
    if rc:
        raise OracleError(lib().orc_error().decode())
    return x


def lu_solve(a, b):
    a = np.asfortranarray(a, dtype=np.float64)
    n = a.shape[0]
    x = np.zeros(n)
    rc = lib().orc_lu_solve(C.c_int(n), _p(a.ravel(order="F").copy()), _p(np.ascontiguousarray(b, dtype=np.float64)),
                            _p(x))
    if rc:
        raise OracleError(lib().orc_error().decode())
    return x


def physics(state, axis):
    f = np.zeros(4)
    a = np.zeros(16)
    c = np.zeros(16)
    rc = lib().orc_physics(_p(np.ascontiguousarray(state, dtype=np.float64)), C.c_int(axis), _p(f), _p(a), _p(c))
    if rc:
        raise OracleError(lib().orc_error().decode())
    return f, a.reshape(4, 4), c.reshape(4, 4)


def subtract_weak(degree, quadrature, hx, hy, fx, fy, residual):
    out = np.array(residual, dtype=np.float64, copy=True)
    rc = lib().orc_subtract_weak(C.c_int(degree), C.c_int(quadrature), C.c_double(hx), C.c_double(hy),
                                 _p(np.ascontiguousarray(fx, dtype=np.float64)),
                                 _p(np.ascontiguousarray(fy, dtype=np.float64)), _p(out))
    if rc:
        raise OracleError(lib().orc_error().decode())
    return out


def basis_tables(degree, quadrature):
    n = degree + 1
    Q = quadrature or n
    values = np.zeros(Q * n)
    wd = np.zeros(Q * n)
    qw = np.zeros(Q)
    rc = lib().orc_basis_tables(C.c_int(degree), C.c_int(quadrature), _p(values), _p(wd), _p(qw))
    if rc:
        raise OracleError(lib().orc_error().decode())
    return values.reshape(Q, n, order="F"), wd.reshape(Q, n, order="F"), qw


def uniform(seed, count):
    out = np.zeros(count)
    lib().orc_uniform(C.c_ulonglong(seed), C.c_longlong(count), _p(out))
    return out


def normal(seed, count):
    out = np.zeros(count)
    lib().orc_normal(C.c_ulonglong(seed), C.c_longlong(count), _p(out))
    return out
