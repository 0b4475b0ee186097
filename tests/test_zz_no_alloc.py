"""The allocation-free Gram entry point (opsec_fisher_gram_noalloc) against
opsec_estimate_run, through the C ABI: the same F, f and q bits, on two calls
in a row (the second smaller, so it reuses the scratch the first one left),
in FP64 and with FP32 traces, and its argument checks."""
import ctypes as C

import numpy as np
import pytest

from opsec_amd import lib

pytestmark = pytest.mark.gpu


def _inputs(M, npar, seed):
    rng = np.random.default_rng(seed)
    lam = np.sort(rng.uniform(0.05, 5, M))[::-1].copy()
    D = rng.normal(size=(npar, M * (M + 1) // 2))
    y = np.sqrt(1 + lam) * rng.normal(size=M)
    prior = rng.uniform(0.5, 2.0, npar)
    return lam, D, y, prior


def _noalloc(ctx, lam, D, y, prior, prec):
    M, npar = len(lam), len(prior)
    bufs = []

    def up(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctx.dev_alloc(a.nbytes)
        ctx.h2d(p, a)
        bufs.append(p)
        return p

    dl, dD, dy, dp = up(lam), up(D), up(y), up(prior)
    dF, df, dq = (ctx.dev_alloc(npar * npar * 8), ctx.dev_alloc(npar * 8), ctx.dev_alloc(npar * 8))
    bufs += [dF, df, dq]
    try:
        ctx.check(ctx.lib.opsec_fisher_gram_noalloc(ctx.h, M, npar, C.c_void_p(dl), C.c_void_p(dD), C.c_void_p(dy),
                                                    C.c_void_p(dp), prec, C.c_void_p(dF), C.c_void_p(df),
                                                    C.c_void_p(dq)), "fisher_gram_noalloc")
        ctx.sync()
        F, f, q = np.empty((npar, npar)), np.empty(npar), np.empty(npar)
        ctx.d2h(F, dF)
        ctx.d2h(f, df)
        ctx.d2h(q, dq)
    finally:
        for b in bufs:
            ctx.dev_free(b)
    return F, f, q


@pytest.mark.parametrize("prec", [lib.OPSEC_PREC_FP64, lib.OPSEC_PREC_FP32])
def test_noalloc_gram_matches_estimate_run(ctx, prec):
    # first call: 50 parameters (64 x 64 Gram), 2500 modes; second: smaller
    for M, npar, seed in ((2500, 50, 1), (700, 9, 2)):
        lam, D, y, prior = _inputs(M, npar, seed)
        F, f, q = _noalloc(ctx, lam, D, y, prior, prec)
        ref = ctx.estimate(lam, D, y, prior, lib.OPSEC_MIX_IDENTITY, prec)
        assert np.array_equal(F, ref["F"]), (M, npar)
        assert np.array_equal(f, ref["f"]), (M, npar)
        assert np.array_equal(q, ref["q"]), (M, npar)
        assert np.abs(F).max() > 0 and np.all(np.isfinite(F))


def test_noalloc_gram_rejects_bad_arguments(ctx):
    lam, D, y, prior = _inputs(10, 63, 3)
    with pytest.raises(lib.OpsecError, match="Nparams must be 1..62"):
        _noalloc(ctx, lam, D, y, prior, lib.OPSEC_PREC_FP64)
    p = C.c_void_p(ctx.dev_alloc(64))
    try:
        rc = ctx.lib.opsec_fisher_gram_noalloc(ctx.h, 10, 2, p, None, p, p, lib.OPSEC_PREC_FP64, p, p, p)
        assert rc == -1 and b"NULL" in ctx.lib.opsec_last_error(ctx.h)
    finally:
        ctx.dev_free(p.value)
