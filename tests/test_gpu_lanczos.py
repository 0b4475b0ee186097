"""GPU block Lanczos: the reference's eigen known answer, numpy eigh on
random matrices, and the KL modes of a periodic box (degenerate multiplets)."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from opsec_amd import lib, synth
from tests import oracles as O

pytestmark = pytest.mark.gpu
GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz"))


def jx_matrix(n):
    j = (n - 1) / 2
    A = np.zeros((n, n))
    for a in range(n):
        for b in range(n):
            m = j - b
            if a == b - 1:
                A[a, b] = 0.5 * math.sqrt(j * (j + 1) - m * (m + 1))
            elif a == b + 1:
                A[a, b] = 0.5 * math.sqrt(j * (j + 1) - m * (m - 1))
    return 0.5 * (A + A.T)


def test_jx_known_answer(ctx):
    # tests/testeig.cpp:78-121: the 6 largest eigenvalues of J_x (j = 5) are 5..0
    A = jx_matrix(11)
    ev, V, nconv = ctx.lanczos_dense(A, 6, tol=1e-10)
    np.testing.assert_allclose(ev, [5, 4, 3, 2, 1, 0], atol=1e-10)
    assert nconv == 6
    np.testing.assert_allclose(V @ A @ V.T, np.diag(ev), atol=1e-10)


@pytest.mark.parametrize("n,nev,block", [(300, 20, 16), (700, 150, 64), (1000, 60, 0)])
def test_random_symmetric_vs_eigh(ctx, n, nev, block):
    rng = np.random.default_rng(n)
    Q = np.linalg.qr(rng.normal(size=(n, n)))[0]
    lam = np.sort(rng.uniform(-1, 1, n) ** 3)[::-1] * 10
    A = (Q * lam) @ Q.T
    A = 0.5 * (A + A.T)
    ev, V, nconv = ctx.lanczos_dense(A, nev, tol=1e-9, block=block)
    w = np.linalg.eigvalsh(A)[::-1][:nev]
    np.testing.assert_allclose(ev, w, atol=1e-8 * abs(w).max())
    assert nconv == nev
    R = A @ V.T - V.T * ev
    assert np.abs(R).max() < 1e-7 * abs(w).max()
    np.testing.assert_allclose(V @ V.T, np.eye(nev), atol=1e-9)


def test_klt_box_multiplets(ctx):
    """KL modes of the periodic 8^3 box: S is 3-level circulant, so its
    spectrum is the real DFT of the generator row, in cubic multiplets."""
    n = 8
    cells = synth.box_cells(n, 0.0, 200.0)
    sv = lib.survey_box(n, 0.0, 200.0)
    lo, hi = synth.linear_bands(4, 0.02, 0.2)
    ctx.model_real(GOLD["pk_k"], GOLD["pk_P"], lo, hi, lib.OPSEC_XI_TABLE)
    ctx.cells_upload(cells, sv)
    ctx.pair_tables([0], ctx.quad(lib.OPSEC_QUAD_GM))
    N = len(cells)
    S = ctx.assemble(0, N)
    nev = 100
    ev, V, nconv = ctx.klt(nev, tol=1e-8)
    w = np.linalg.eigvalsh(S)[::-1]
    np.testing.assert_allclose(ev, w[:nev], atol=1e-9 * w[0])
    # circulant spectrum: eigenvalues = DFT of the first row (generator)
    gen = S[0].reshape(n, n, n)
    dft = np.sort(np.fft.fftn(gen).real.ravel())[::-1]
    np.testing.assert_allclose(w, dft, atol=1e-10 * w[0])
    # B^T S B = diag(lambda), B^T B = I (the klt/comma consistency invariant)
    np.testing.assert_allclose(V @ S @ V.T, np.diag(ev), atol=1e-9 * w[0])
    np.testing.assert_allclose(V @ V.T, np.eye(nev), atol=1e-9)
    assert nconv == nev


def test_auto_krylov_retry_keeps_and_improves(ctx):
    """A slowly decaying spectrum defeats the first Krylov space (maxdim =
    3 nev + 4 b): the automatic retries with twice the space converge more
    modes, and the result never has fewer converged modes than the first pass.
    (A numpy block Lanczos of the same construction converges 0, 20 and 40 of
    the 40 modes with spaces of 192, 368 and 736.)"""
    n, nev, block = 3000, 40, 16
    rng = np.random.default_rng(11)
    Q = np.linalg.qr(rng.normal(size=(n, n)))[0]
    lam = np.exp(-20.0 * np.arange(n) / n)
    A = (Q * lam) @ Q.T
    A = 0.5 * (A + A.T)
    first = 3 * nev + 4 * block
    ev1, _, nconv1 = ctx.lanczos_dense(A, nev, tol=1e-9, block=block, maxdim=first)
    ev, V, nconv = ctx.lanczos_dense(A, nev, tol=1e-9, block=block)
    print(f"flat spectrum: first pass {nconv1} of {nev} converged, automatic retries {nconv}")
    assert nconv1 < nev
    assert nconv > nconv1
    w = np.sort(lam)[::-1][:nev]
    np.testing.assert_allclose(ev[:nconv], w[:nconv], atol=1e-8)


def test_dense_path_for_whole_space(ctx):
    """A Krylov space that would span the whole space goes to the dense
    eigen-solver: N prime (no block size divides it), all modes converged."""
    n, nev = 211, 50
    rng = np.random.default_rng(13)
    Q = np.linalg.qr(rng.normal(size=(n, n)))[0]
    lam = np.linspace(3.0, -1.0, n)
    A = (Q * lam) @ Q.T
    A = 0.5 * (A + A.T)
    ev, V, nconv = ctx.lanczos_dense(A, nev, tol=1e-10, block=0)
    assert nconv == nev
    np.testing.assert_allclose(ev, lam[:nev], atol=1e-11 * 3)
    np.testing.assert_allclose(V @ A @ V.T, np.diag(ev), atol=1e-10)
