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
