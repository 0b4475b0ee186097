"""Host stage of the eigen-solver (tridiagonal QL + inverse iteration) against
scipy.linalg.eigh_tridiagonal.  CPU only: calls an internal host entry point."""
import ctypes as C

import numpy as np
import pytest
import scipy.linalg as sl

from opsec_amd import lib


def tridiag_top(d, e, nev):
    L = lib.load()
    fn = L.opsec_diag_tridiag_top
    fn.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    fn.restype = C.c_int
    n = len(d)
    d = np.ascontiguousarray(d, float)
    e = np.ascontiguousarray(e, float)
    ev = np.zeros(nev)
    Z = np.zeros((nev, n))
    assert fn(n, d.ctypes.data, e.ctypes.data, nev, ev.ctypes.data, Z.ctypes.data) == 0
    return ev, Z


@pytest.mark.parametrize("n,nev,seed", [(1, 1, 0), (2, 1, 1), (50, 10, 2), (400, 120, 3)])
def test_random_tridiagonal(n, nev, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=n)
    e = rng.normal(size=max(n - 1, 0))
    ev, Z = tridiag_top(d, e, nev)
    w, V = sl.eigh_tridiagonal(d, e)
    np.testing.assert_allclose(ev, w[::-1][:nev], rtol=1e-12, atol=1e-12)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    np.testing.assert_allclose(Z @ T @ Z.T, np.diag(ev), atol=1e-10)
    np.testing.assert_allclose(Z @ Z.T, np.eye(nev), atol=1e-10)


def test_split_and_degenerate_tridiagonal():
    # two identical unreduced blocks (exactly degenerate spectrum) plus a split
    rng = np.random.default_rng(7)
    blk_d = rng.normal(size=30)
    blk_e = rng.normal(size=29)
    d = np.concatenate([blk_d, blk_d, [0.3]])
    e = np.concatenate([blk_e, [0.0], blk_e, [0.0]])
    nev = 40
    ev, Z = tridiag_top(d, e, nev)
    w = np.sort(np.linalg.eigvalsh(np.diag(d) + np.diag(e, 1) + np.diag(e, -1)))[::-1]
    np.testing.assert_allclose(ev, w[:nev], atol=1e-12)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    np.testing.assert_allclose(Z @ Z.T, np.eye(nev), atol=1e-9)
    np.testing.assert_allclose(Z @ T @ Z.T, np.diag(ev), atol=1e-9)


@pytest.mark.parametrize("nev", [1, 7, 30])
def test_wilkinson_pairs(nev):
    # W_101^+: the largest eigenvalues come in pairs equal to far below
    # rounding, so nev = 7 splits a pair
    m = 50
    d = np.abs(np.arange(2 * m + 1) - m).astype(float)
    e = np.ones(2 * m)
    ev, Z = tridiag_top(d, e, nev)
    w = sl.eigh_tridiagonal(d, e, eigvals_only=True)[::-1]
    np.testing.assert_allclose(ev, w[:nev], rtol=1e-13, atol=1e-12)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    np.testing.assert_allclose(Z @ Z.T, np.eye(nev), atol=1e-9)
    np.testing.assert_allclose(Z @ T @ Z.T, np.diag(ev), atol=1e-9)


def test_tight_cluster_returns_every_pair():
    # 60 eigenvalues within a few ulps of 1: the Sturm counts inside the
    # cluster are not monotone, yet exactly nev orthonormal pairs come back
    rng = np.random.default_rng(5)
    n, nev = 60, 17
    d = 1.0 + 1e-15 * rng.normal(size=n)
    e = 1e-15 * (1.0 + rng.random(n - 1))
    ev, Z = tridiag_top(d, e, nev)
    assert ev.shape == (nev,) and np.all(np.diff(ev) <= 0)
    np.testing.assert_allclose(ev, 1.0, atol=1e-13)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    np.testing.assert_allclose(Z @ Z.T, np.eye(nev), atol=1e-9)
    assert np.abs(T @ Z.T - Z.T * ev).max() < 1e-13
