"""The device's 3x3 Newton solve (magpy_amd/csrc/linalg3.h) built for the host
equals the oracle's dgesv (oracle/sllg_oracle.c, restating the LAPACKE_dgesv
call of lib/optimisation.cpp:134-137) bit for bit.

Both are compiled without floating-point contraction, and the header is the
one the kernels include, so this pins the pivoting logic the implicit kernel
runs: the row interchanges done as conditional swaps, the first maximal pivot
on ties (idamax), interchanges that carry the multipliers (dlaswp), the
substitution order, and b left unchanged on an exactly zero pivot.  Newton
matrices of the implicit midpoint step (I - J/2 for small reduced dt) never
pivot, so the interchange paths are reachable only through cases like these.
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "magpy_amd", "csrc")
DRIVER = r'''
#include "linalg3.h"
extern "C" int l3_solve(double *A, double *b) { return mp::solve3(A, b) ? 0 : 1; }
'''


@pytest.fixture(scope="module")
def l3(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not available")
    d = tmp_path_factory.mktemp("linalg3")
    src = d / "l3.cpp"
    src.write_text(DRIVER)
    so = d / "libl3.so"
    subprocess.run([gxx, "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared",
                    "-I", CSRC, "-o", str(so), str(src)], check=True)
    lib = C.CDLL(str(so))
    lib.l3_solve.restype = C.c_int
    lib.l3_solve.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double)]
    return lib


def both(l3, A, b):
    A = np.ascontiguousarray(A, dtype=np.float64).reshape(3, 3)
    b = np.ascontiguousarray(b, dtype=np.float64)
    a1, b1 = A.copy(), b.copy()
    a2, b2 = A.copy(), b.copy()
    dp = C.POINTER(C.c_double)
    r1 = l3.l3_solve(a1.ctypes.data_as(dp), b1.ctypes.data_as(dp))
    ipiv = np.zeros(3, dtype=np.int32)
    r2 = O.lib().orc_dgesv(3, a2.ctypes.data_as(dp), ipiv.ctypes.data_as(C.POINTER(C.c_int)),
                           b2.ctypes.data_as(dp))
    return r1, b1, r2, b2


def check(l3, A, b):
    r1, x1, r2, x2 = both(l3, A, b)
    assert (r1 == 0) == (r2 == 0)
    assert np.array_equal(x1, x2)
    if r1 != 0:
        assert np.array_equal(x1, np.asarray(b, dtype=np.float64))   # b untouched
    return r1


def test_random_matrices(l3):
    rng = np.random.default_rng(3)
    pivoted = 0
    for _ in range(2000):
        A = rng.normal(size=(3, 3))
        assert check(l3, A, rng.normal(size=3)) == 0
        pivoted += np.argmax(np.abs(A[:, 0])) != 0
    assert pivoted > 1000          # most of them interchange rows


def test_every_row_order(l3):
    rng = np.random.default_rng(4)
    base = rng.normal(size=(3, 3)) + np.diag([5.0, 4.0, 3.0])
    for perm in ([0, 1, 2], [0, 2, 1], [1, 0, 2], [1, 2, 0], [2, 0, 1], [2, 1, 0]):
        A = base[perm]
        b = rng.normal(size=3)
        assert check(l3, A, b) == 0
        # the solution is the permuted system's, i.e. base x = b[inverse perm]
        _, x, _, _ = both(l3, A, b)
        assert np.allclose(A @ x, b, rtol=0, atol=1e-12)


def test_ties_take_the_first_maximal_row(l3):
    for col in ([1.0, -1.0, 1.0], [-2.0, 2.0, 0.5], [0.5, 3.0, -3.0], [0.0, 2.0, 2.0]):
        A = np.array([[col[0], 1.0, 2.0], [col[1], 3.0, -1.0], [col[2], 0.5, 4.0]])
        check(l3, A, [1.0, 2.0, 3.0])


def test_newton_matrices(l3):
    rng = np.random.default_rng(5)
    for _ in range(500):
        A = np.eye(3) - 0.5 * 1e-2 * rng.normal(size=(3, 3))
        assert check(l3, A, rng.normal(size=3)) == 0


@pytest.mark.parametrize("A", [
    np.zeros((3, 3)),
    [[0.0, 1.0, 2.0], [0.0, 3.0, 4.0], [0.0, 5.0, 6.0]],        # zero first column
    [[1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [0.0, 0.0, 1.0]],        # zero pivot at column 1
    [[1.0, 1.0, 0.0], [1.0, 0.0, 1.0], [2.0, 1.0, 1.0]],        # row 3 = row 1 + row 2, exactly
    [[2.0, 0.0, 0.0], [0.0, 3.0, 0.0], [0.0, 0.0, 0.0]],        # zero pivot at column 2
])
def test_singular_leaves_b(l3, A):
    assert check(l3, np.asarray(A, dtype=np.float64), [1.0, -2.0, 0.5]) != 0
