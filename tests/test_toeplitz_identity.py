"""The identity behind the Toeplitz projection, on the CPU.

The projection kernels' index maps and weights (csrc/project.hip) restated in
numpy: cells in grid order g = x G + slab position (x of a box with the (y, z)
slab, phi of a shell with the (r, mu) arcs), the length-2 n1 DFT along that axis
as 2 n1 real batches (batch 0 and odd b the real part of k = (b + 1) / 2, even
b > 0 its imaginary part), That_k from the pair table (2-level Toeplitz on a
box, a dense arc x arc matrix on a shell), the inverse weights c_k / L, and the
stage-2 product over the grid.  Against B'^T S B' with S formed entry by entry
from the same table (the cache keys of src/sig.cpp:109-112 and the arc cache of
src/sig.cpp:151-199), on irregular, shuffled cell subsets, the shortest axes,
and a shell table that is not symmetric in its two arcs.  A wrong sign of the
imaginary parts, a wrong c_k, a transposed map or a dropped zero at d = n1
fails these by O(1).
"""
import numpy as np
import pytest


def _toeplitz_project(Bw, gmap, n1, G, that_of_k):
    """B'^T S B' by the DFT route: Bw is N x M in cell order, gmap maps the
    n1 G grid positions to cells (-1: none)."""
    L, M = 2 * n1, Bw.shape[1]
    Bg = np.zeros((n1 * G, M))
    ok = gmap >= 0
    Bg[ok] = Bw[gmap[ok]]
    Bx = Bg.reshape(n1, G, M)
    x = np.arange(n1)
    Y = np.zeros((n1, G, M))
    for b in range(2 * n1):
        k = (b + 1) >> 1
        imag = b > 0 and b % 2 == 0
        ang = 2 * np.pi * x * k / L
        fwd = -np.sin(ang) if imag else np.cos(ang)       # Bhat: R_k or I_k
        Bhat = np.tensordot(fwd, Bx, axes=(0, 0))
        Yhat = that_of_k(k) @ Bhat
        ck = 1.0 if k in (0, n1) else 2.0
        inv = -ck * np.sin(ang) / L if imag else ck * np.cos(ang) / L
        Y += inv[:, None, None] * Yhat[None]
    return Bg.T @ Y.reshape(n1 * G, M)


def _cells_subset(n1, G, drop, rng):
    pos = np.arange(n1 * G)
    pos = pos[rng.random(pos.size) >= drop]
    return rng.permutation(pos)          # cell a sits at grid position pos[a]


@pytest.mark.parametrize("dims,drop", [((6, 5, 4), 0.2), ((1, 3, 5), 0.0), ((2, 4, 3), 0.3), ((5, 1, 7), 0.1)])
def test_box_identity(dims, drop):
    n1, n2, n3 = dims
    G = n2 * n3
    rng = np.random.default_rng(sum(dims))
    Q = rng.normal(size=(n1, n2, n3))                    # Q(|dx|, |dy|, |dz|)
    pos = _cells_subset(n1, G, drop, rng)
    N, M = pos.size, 7
    x, r = pos // G, pos % G
    y, z = r // n3, r % n3
    S = Q[np.abs(x[:, None] - x[None]), np.abs(y[:, None] - y[None]), np.abs(z[:, None] - z[None])]
    Bw = rng.normal(size=(N, M))
    gmap = -np.ones(n1 * G, dtype=int)
    gmap[pos] = np.arange(N)
    ys, zs = np.arange(G) // n3, np.arange(G) % n3
    L = 2 * n1

    def that(k):
        d = np.arange(1, n1)
        qhat = Q[0] + 2 * np.tensordot(np.cos(2 * np.pi * d * k / L), Q[1:], axes=(0, 0))
        return qhat[np.abs(ys[:, None] - ys[None]), np.abs(zs[:, None] - zs[None])]

    got = _toeplitz_project(Bw, gmap, n1, G, that)
    ref = Bw.T @ S @ Bw
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("dims,drop", [((3, 4, 8), 0.25), ((2, 3, 1), 0.0), ((4, 2, 2), 0.1), ((2, 2, 9), 0.3)])
def test_shell_identity(dims, drop):
    nr, nmu, nphi = dims
    G = nr * nmu
    rng = np.random.default_rng(100 + sum(dims))
    Q = rng.normal(size=(G, G, nphi))                     # Q(h_a, h_b, |dphi|), not symmetric in h
    pos = _cells_subset(nphi, G, drop, rng)               # grid order g = dphi G + h
    N, M = pos.size, 6
    phi, h = pos // G, pos % G
    S = Q[h[:, None], h[None], np.abs(phi[:, None] - phi[None])]
    Bw = rng.normal(size=(N, M))
    gmap = -np.ones(nphi * G, dtype=int)
    gmap[pos] = np.arange(N)
    L = 2 * nphi

    def that(k):
        d = np.arange(1, nphi)
        return Q[:, :, 0] + 2 * np.tensordot(Q[:, :, 1:], np.cos(2 * np.pi * d * k / L), axes=(2, 0))

    got = _toeplitz_project(Bw, gmap, nphi, G, that)
    ref = Bw.T @ S @ Bw
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_a_dropped_imaginary_sign_is_caught():
    """The same restatement with the imaginary parts' inverse sign flipped
    misses the dense product by O(1): the checks above can fail."""
    n1, G, M = 4, 6, 5
    rng = np.random.default_rng(7)
    Q = rng.normal(size=(G, G, n1))
    pos = rng.permutation(n1 * G)
    h, phi = pos % G, pos // G
    S = Q[h[:, None], h[None], np.abs(phi[:, None] - phi[None])]
    Bw = rng.normal(size=(pos.size, M))
    gmap = -np.ones(n1 * G, dtype=int)
    gmap[pos] = np.arange(pos.size)
    L, x = 2 * n1, np.arange(n1)
    Bx = np.zeros((n1 * G, M))
    Bx[gmap >= 0] = Bw[gmap[gmap >= 0]]
    Bx = Bx.reshape(n1, G, M)
    Y = np.zeros((n1, G, M))
    for b in range(2 * n1):
        k = (b + 1) >> 1
        imag = b > 0 and b % 2 == 0
        ang = 2 * np.pi * x * k / L
        That = Q[:, :, 0] + 2 * np.tensordot(Q[:, :, 1:], np.cos(2 * np.pi * np.arange(1, n1) * k / L), axes=(2, 0))
        Yhat = That @ np.tensordot(-np.sin(ang) if imag else np.cos(ang), Bx, axes=(0, 0))
        ck = 1.0 if k in (0, n1) else 2.0
        inv = ck * np.sin(ang) / L if imag else ck * np.cos(ang) / L    # sign flipped
        Y += inv[:, None, None] * Yhat[None]
    got = Bx.reshape(n1 * G, M).T @ Y.reshape(n1 * G, M)
    ref = Bw.T @ S @ Bw
    assert np.abs(got - ref).max() > 1e-3 * np.abs(ref).max()
