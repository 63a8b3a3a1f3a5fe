"""Pins the oracle's blocked dense routines (oracle/ro_dense.c) against LAPACK.

The reference calls zgetrf / zgetrs / zgemm (factorization_module.f90:71,
346-349; robin_tree_module.f90:550-605).  scipy's LAPACK zgetrf picks pivots by
the same rule (izamax: first entry of maximal |re|+|im|), so on generic matrices
the pivot sequence must be identical and the factors equal to rounding.  Sizes
straddle the 32-column panel, the 256-deep GEMM K blocks and the 192 x 96 output
tiles, and the threaded (inner-parallel) paths must give the same bits as the
single-thread ones.
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg.lapack as lapack

from oracle import pyoracle

dp = ctypes.POINTER(ctypes.c_double)
ip = ctypes.POINTER(ctypes.c_int)


class _Z(ctypes.Structure):
    """double complex by value (the x86-64 ABI passes both like two doubles)"""
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


def _lib():
    L = pyoracle.lib()
    L.ro_zgetrf.argtypes = [ctypes.c_int, dp, ip]
    L.ro_zgetrf.restype = ctypes.c_int
    L.ro_zgetrs.argtypes = [ctypes.c_int, dp, ip, ctypes.c_int, dp, ctypes.c_int]
    L.ro_zgemm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _Z, dp, ctypes.c_int, dp,
                           ctypes.c_int, _Z, dp, ctypes.c_int]
    L.ro_dense_set_inner.argtypes = [ctypes.c_int]
    return L


def _p(a):
    return a.ctypes.data_as(dp)


def _rand(rng, *shape):
    return np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def _getrf(L, A):
    lu = np.asfortranarray(A.copy())
    piv = np.zeros(A.shape[0], np.int32)
    info = L.ro_zgetrf(A.shape[0], _p(lu), piv.ctypes.data_as(ip))
    return lu, piv, info


@pytest.mark.parametrize("n", [1, 5, 31, 32, 33, 97, 320, 515])
def test_getrf_matches_lapack(n):
    L = _lib()
    A = _rand(np.random.default_rng(n), n, n)
    lu, piv, info = _getrf(L, A)
    lu_ref, piv_ref, info_ref = lapack.zgetrf(A)
    assert info == 0 and info_ref == 0
    np.testing.assert_array_equal(piv, piv_ref)
    assert np.abs(lu - lu_ref).max() <= 1e-11 * max(1.0, np.abs(lu_ref).max())
    # the factors reproduce the row-permuted matrix
    Lf = np.tril(lu, -1) + np.eye(n)
    U = np.triu(lu)
    PA = A.copy()
    for j, p in enumerate(piv):
        PA[[j, p]] = PA[[p, j]]
    assert np.abs(Lf @ U - PA).max() <= 1e-12 * n * np.abs(A).max()


def test_getrf_reports_zero_pivot():
    L = _lib()
    A = _rand(np.random.default_rng(3), 40, 40)
    A[:, 17] = 0.0     # an exactly zero column: the pivot of column 17 is exactly zero
    _, _, info = _getrf(L, A)
    _, _, info_ref = lapack.zgetrf(A)
    assert info_ref == 18
    assert info == info_ref


@pytest.mark.parametrize("n,nrhs", [(33, 1), (33, 2), (100, 3), (320, 64), (515, 130)])
def test_getrs_solves(n, nrhs):
    L = _lib()
    rng = np.random.default_rng(n + nrhs)
    A = _rand(rng, n, n)
    B = _rand(rng, n, nrhs)
    lu, piv, _ = _getrf(L, A)
    X = np.asfortranarray(B.copy())
    L.ro_zgetrs(n, _p(lu), piv.ctypes.data_as(ip), nrhs, _p(X), n)
    X_ref = np.linalg.solve(A, B)
    assert np.abs(X - X_ref).max() <= 1e-10 * np.abs(X_ref).max()


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (7, 3, 5), (193, 97, 257), (320, 64, 320), (40, 400, 600)])
def test_gemm_matches_numpy(m, n, k):
    L = _lib()
    rng = np.random.default_rng(m * n + k)
    A, B, C = _rand(rng, m, k), _rand(rng, k, n), _rand(rng, m, n)
    alpha, beta = 0.7 - 1.3j, -0.4 + 0.2j
    out = np.asfortranarray(C.copy())
    L.ro_zgemm(m, n, k, _Z(alpha.real, alpha.imag), _p(A), m, _p(B), k,
               _Z(beta.real, beta.imag), _p(out), m)
    ref = alpha * (A @ B) + beta * C
    assert np.abs(out - ref).max() <= 1e-12 * k * np.abs(ref).max()


def test_threaded_paths_give_identical_bits():
    L = _lib()
    rng = np.random.default_rng(11)
    n = 400
    A = _rand(rng, n, n)
    B = _rand(rng, n, 300)
    outs = []
    for inner in (1, 4):
        L.ro_dense_set_inner(inner)
        try:
            lu, piv, _ = _getrf(L, A)
            X = np.asfortranarray(B.copy())
            L.ro_zgetrs(n, _p(lu), piv.ctypes.data_as(ip), B.shape[1], _p(X), n)
            C = np.zeros((n, B.shape[1]), np.complex128, order="F")
            one, zero = _Z(1.0, 0.0), _Z(0.0, 0.0)
            L.ro_zgemm(n, B.shape[1], n, one, _p(A), n, _p(X), n, zero, _p(C), n)
        finally:
            L.ro_dense_set_inner(0)
        outs.append((lu, piv, X, C))
    for a, b in zip(outs[0], outs[1]):
        np.testing.assert_array_equal(a, b)
    # and A X reproduces B
    assert np.abs(outs[0][3] - B).max() <= 1e-9 * np.abs(B).max()
