"""The FP64 kernels compute psi = 1 - exp(-rho) with a restatement of
glibc's exp (mesobasic_amd/csrc/lbm_exp.h).  Built here for the host with
g++ and compared with the host libm bit for bit; the same code runs on the
GPU, where test_gpu_fp64.py checks it again."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "exp_shim.cpp")


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("exp") / "libexpshim.so")
    subprocess.run(["g++", "-O2", "-std=c++20", "-ffp-contract=off", "-fPIC", "-shared",
                    "-I", os.path.join(ROOT, "mesobasic_amd", "csrc"), SRC, "-o", out, "-lm"],
                   check=True)
    lib = ctypes.CDLL(out)
    lib.exp_shim_mismatches.restype = ctypes.c_int64
    lib.exp_shim_mismatches.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int64,
                                        ctypes.POINTER(ctypes.c_double)]
    return lib


def _count(lib, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    bad = ctypes.c_double(0)
    n = lib.exp_shim_mismatches(x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), x.size,
                                ctypes.byref(bad))
    return n, bad.value


@pytest.mark.parametrize("lo,hi", [(-40.0, 0.0), (-3.5, 0.0), (-511.0, 511.0), (-1e-8, 1e-8)])
def test_exp_restatement_matches_host_libm(shim, lo, hi):
    x = np.random.default_rng(1).uniform(lo, hi, 2_000_000)
    n, bad = _count(shim, x)
    assert n == 0, f"{n} mismatches, first at x={bad!r}"


def test_exp_restatement_near_reduction_boundaries(shim):
    # arguments where x*128/ln2 is within a few ulp of k + 1/2: the
    # rounding of the reduction step decides the table index.
    k = -np.arange(1, 6000)
    x0 = (k + 0.5) / (128 / np.log(2))
    x = (x0[:, None].view(np.int64) + np.arange(-40, 40)[None, :]).view(np.float64).ravel()
    n, bad = _count(shim, x)
    assert n == 0, f"{n} mismatches, first at x={bad!r}"


def test_exp_special_values(shim):
    x = np.array([0.0, -0.0, np.inf, -np.inf, 709.78, 709.79, -708.39, 1e-300, -1e-300])
    n, bad = _count(shim, x)
    assert n == 0, bad
