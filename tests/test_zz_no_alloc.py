"""The allocation-free LU solve (hbg_lu_solve_noalloc) against hbg_lu_solve through the C ABI
(GPU): the same factors and solutions on two calls in a row, the second of smaller order (the
shared per-GPU array keeps the first call's contents), a system that needs row interchanges (the
pivoted path in the same array), and the refusal of a size beyond the array."""

import ctypes
import sys

import numpy as np
import pytest

import hydrobem_amd as hb

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def call(name, session, a, b):
    """One LU solve through the C ABI: a (n, n) and b (n, r) complex, column-major copies
    overwritten by the factors and the solution; returns (status, factors, solution)."""
    n = a.shape[0]
    af = np.array(a, dtype=np.complex128, order="F", copy=True)
    bf = np.array(b, dtype=np.complex128, order="F", copy=True)
    st = getattr(hb.library(), name)(session._h, n, af.ctypes.data_as(_dp), bf.shape[1], bf.ctypes.data_as(_dp))
    return st, af, bf


def system(n, r, seed, pivoting=False):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + (0.0 if pivoting else 3.0 * np.sqrt(n)) * np.eye(n)
    if pivoting:  # small diagonal: partial pivoting interchanges rows in every panel
        a[np.arange(n), np.arange(n)] *= 1e-3
    b = rng.normal(size=(n, r)) + 1j * rng.normal(size=(n, r))
    return a, b


@pytest.fixture(scope="module")
def session():
    s = hb.Session(0)
    yield s
    s.close()


def test_two_calls_in_a_row_the_second_smaller(session):
    for n, r, seed in ((700, 10, 1), (300, 4, 2)):
        say(f"case: no-alloc LU n={n} nrhs={r} against hbg_lu_solve")
        a, b = system(n, r, seed)
        st0, f0, x0 = call("hbg_lu_solve", session, a, b)
        st1, f1, x1 = call("hbg_lu_solve_noalloc", session, a, b)
        assert st0 == 0 and st1 == 0
        assert np.array_equal(f0, f1) and np.array_equal(x0, x1), (np.abs(x0 - x1).max(), np.abs(f0 - f1).max())
        assert np.abs(a @ x1 - b).max() / (np.abs(a).max() * np.abs(x1).max()) < 1e-13


def test_system_with_interchanges(session):
    say("case: no-alloc LU n=200 with row interchanges (pivoted path in the array)")
    a, b = system(200, 3, 5, pivoting=True)
    st0, f0, x0 = call("hbg_lu_solve", session, a, b)
    st1, f1, x1 = call("hbg_lu_solve_noalloc", session, a, b)
    assert st0 == 0 and st1 == 0
    assert np.array_equal(x0, x1) and np.array_equal(f0, f1), np.abs(x0 - x1).max()
    assert np.abs(a @ x1 - b).max() / (np.abs(a).max() * np.abs(x1).max()) < 1e-12


def test_sizes_beyond_the_array_are_refused(session):
    say("case: no-alloc LU beyond the array, and a singular system")
    n = 1500
    a = np.eye(n, dtype=np.complex128)
    b = np.ones((n, 32), np.complex128)
    st, _, _ = call("hbg_lu_solve_noalloc", session, a, b)
    assert st == 1 and b"allocation-free maximum" in hb.library().hbg_last_error_message()
    z = np.zeros((64, 64), np.complex128)
    st, _, _ = call("hbg_lu_solve_noalloc", session, z, np.ones((64, 1), np.complex128))
    assert st == 8 and b"zero pivot 1" in hb.library().hbg_last_error_message()
    # the array serves the next call
    a, b = system(128, 2, 7)
    st, _, x = call("hbg_lu_solve_noalloc", session, a, b)
    assert st == 0 and np.abs(a @ x - b).max() / (np.abs(a).max() * np.abs(x).max()) < 1e-13
