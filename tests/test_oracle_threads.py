"""The threaded oracle build (oracle/libdgh_oracle_omp.so, the CPU baseline on
all host cores) gives bit-identical results to the serial build the tests use: same
residual, Jacobian action, block-Jacobi apply, GMRES solution and Newton run.
Each build runs in its own process (the binding loads one library per process).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

SCRIPT = r'''
import sys, numpy as np
sys.path.insert(0, sys.argv[3])
import oracle_py as O
O.select_library(sys.argv[1])
out = {}
o = O.Oracle(["--nx", "6", "--ny", "5", "--k", "8", "--testcase", "seeded", "--cfl", "5", "--integrator", "be"])
u = o.seeded(5)
rate = o.freeze(u)
o.set_step(0.1 / rate, u)
r = o.dg_residual(u)
o.dg_refresh(u)
o.bj_refresh()
x = np.cos(0.1 * np.arange(u.size))
out["rate"], out["r"], out["jv"], out["bj"] = np.array([rate]), r, o.dg_apply(x), o.bj_apply(x)
out["gmres"] = o.dg_gmres(-r, 20, 1500, 1e-12)[0]
run = O.Oracle(["--nx", "12", "--ny", "12", "--k", "3", "--testcase", "wave", "--integrator", "ssp-dirk3",
                "--cfl", "2", "--steps", "1", "--restart", "20"])
out["run"] = run.advance("dg", run.initial())[0]
np.savez(sys.argv[2], **out)
'''


def _dump(lib, path, threads):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    subprocess.run([sys.executable, "-c", SCRIPT, lib, str(path), HERE], check=True, env=env, timeout=600)
    return np.load(path)


def test_threaded_build_is_bit_identical(tmp_path):
    serial = os.path.join(ROOT, "oracle", "libdgh_oracle.so")
    threaded = os.path.join(ROOT, "oracle", "libdgh_oracle_omp.so")
    if not os.path.exists(threaded):
        pytest.skip("threaded oracle not built (make -C oracle)")
    a = _dump(serial, tmp_path / "serial.npz", 1)
    b = _dump(threaded, tmp_path / "threaded.npz", 4)
    for key in a.files:
        assert np.array_equal(a[key], b[key]), key
