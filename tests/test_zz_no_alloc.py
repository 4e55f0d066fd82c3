"""ifem_gpu_spmv_noalloc, the SpMV with host vectors that stages through the
library's fixed device array instead of allocating, called through the C ABI
and checked against ifem_gpu_spmv (the same kernel on the same inputs, so the
products agree bit for bit): two calls in a row, the second on a smaller
system (the array still holds the first call's larger contents beyond it),
then the first system again, and the refusals."""
import ctypes as C

import numpy as np
import pytest

from ifem_amd import problems, setup_problem
from ifem_amd._native import lib

pytestmark = pytest.mark.gpu


def _noalloc(s, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.full(s.neq, np.nan)
    rc = lib().ifem_gpu_spmv_noalloc(s.h, x.ctypes.data_as(C.c_void_p),
                                     y.ctypes.data_as(C.c_void_p))
    return rc, y


def test_noalloc_spmv_matches_spmv_on_consecutive_calls():
    big = problems.elasticity_cube(12, 2, seed=3)
    small = problems.poisson_cube(6, 2)
    print(f"[gpu] spmv_noalloc: elasticity neq={big.neq}, then poisson neq={small.neq}", flush=True)
    assert small.neq < big.neq
    with setup_problem(big) as s1, setup_problem(small) as s2:
        for s, seed in ((s1, 1), (s2, 2), (s1, 3)):
            x = np.random.default_rng(seed).standard_normal(s.neq)
            y_ref = s.spmv(x)
            rc, y = _noalloc(s, x)
            assert rc == 0, lib().ifem_gpu_last_error().decode()
            assert np.array_equal(y, y_ref)


def test_noalloc_spmv_refusals():
    print("[gpu] spmv_noalloc refusals", flush=True)
    pb = problems.poisson_cube(4, 2)
    with setup_problem(pb) as s:
        x = np.zeros(s.neq)
        assert lib().ifem_gpu_spmv_noalloc(s.h, x.ctypes.data_as(C.c_void_p), None) == 1
        assert lib().ifem_gpu_spmv_noalloc(None, x.ctypes.data_as(C.c_void_p),
                                           x.ctypes.data_as(C.c_void_p)) == 1
