"""bem_zgetrf_local_noalloc (the allocation-free variant of the block-local
pivoted LU) against bem_zgetrf_local through the C ABI: the same factors,
interchanges, growth and status, bit for bit, on two calls in a row, the
second with a smaller matrix; the context's device footprint does not change
across a call; sizes beyond BEM_NOALLOC_MAX_N are refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _matrix(n, seed, zero_col=None):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    A *= np.exp(rng.normal(size=(n, 1)))   # row scales: interchanges inside the blocks
    if zero_col is not None:
        A[:, zero_col] = 0.0
    return np.asfortranarray(A)


def _noalloc(ctx, A):
    """Factors through bem_zgetrf_local_noalloc; the device buffers are the
    caller's (allocated before the call), and the context's device bytes are
    read on both sides of the call."""
    n = A.shape[0]
    dA = ctx._dev(A)
    dp = ctx._dev(np.zeros(n + 1, dtype=np.int32))
    g = ctypes.c_double(0.0)
    try:
        before = ctx.device_bytes()
        rc = ctx.lib.bem_zgetrf_local_noalloc(ctx.h, dA, n, n, dp, ctypes.byref(g))
        after = ctx.device_bytes()
        LU = np.zeros_like(A)
        piv = np.zeros(n + 1, dtype=np.int32)
        ctx._check(ctx.lib.bem_memcpy_d2h(ctx.h, LU.ctypes.data, dA, LU.nbytes))
        ctx._check(ctx.lib.bem_memcpy_d2h(ctx.h, piv.ctypes.data, dp, piv.nbytes))
        return rc, LU, piv[:n], g.value, before, after
    finally:
        ctx.lib.bem_device_free(ctx.h, dA)
        ctx.lib.bem_device_free(ctx.h, dp)


def test_noalloc_lu_matches_local_lu():
    import bem_amd
    from bem_amd import BemError
    with bem_amd.Context(0, lanes=1) as ctx:
        for n, seed in ((1000, 1), (333, 2)):   # the second call smaller
            print(f"noalloc vs local: n={n}", flush=True)
            A = _matrix(n, seed)
            LU_ref, piv_ref, g_ref = ctx.zgetrf_local_host(A)
            rc, LU, piv, g, before, after = _noalloc(ctx, A)
            assert rc == 0
            assert before == after
            assert np.array_equal(LU, LU_ref) and np.array_equal(piv, piv_ref) and g == g_ref
        # an exactly zero column: the same status from both
        print("noalloc vs local: singular block", flush=True)
        A = _matrix(200, 3, zero_col=70)
        with pytest.raises(BemError) as e:
            ctx.zgetrf_local_host(A)
        rc, *_ = _noalloc(ctx, A)
        assert rc == e.value.code


def test_noalloc_rejects_oversize():
    import bem_amd
    from bem_amd.runtime import BEM_E_UNSUPPORTED, BEM_NOALLOC_MAX_N
    n = BEM_NOALLOC_MAX_N + 1
    with bem_amd.Context(0, lanes=1) as ctx:
        dA = ctx._dev(np.zeros((n, n), dtype=np.complex128, order="F"))
        dp = ctx._dev(np.zeros(n + 1, dtype=np.int32))
        try:
            rc = ctx.lib.bem_zgetrf_local_noalloc(ctx.h, dA, n, n, dp, None)
            assert rc == BEM_E_UNSUPPORTED
            assert b"BEM_NOALLOC_MAX_N" in ctx.lib.bem_last_error(ctx.h)
        finally:
            ctx.lib.bem_device_free(ctx.h, dA)
            ctx.lib.bem_device_free(ctx.h, dp)
