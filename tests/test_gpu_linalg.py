"""Dense complex128 kernels of the HIP library (ZGEMM on FP64 MFMA,
tournament-pivoted LU + triangular solves) against numpy, through the C ABI."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mnk", [(1, 1, 1), (16, 16, 4), (64, 64, 16), (65, 130, 17), (200, 64, 300), (257, 31, 64)])
def test_zgemm(gpu_ctx, mnk):
    m, n, k = mnk
    rng = np.random.default_rng(m * 7 + n)
    A = rng.normal(size=(m, k)) + 1j * rng.normal(size=(m, k))
    B = rng.normal(size=(k, n)) + 1j * rng.normal(size=(k, n))
    C = rng.normal(size=(m, n)) + 1j * rng.normal(size=(m, n))
    al, be = 0.7 - 0.2j, -1.0 + 0.5j
    got = gpu_ctx.zgemm_host(A, B, C, al, be)
    ref = al * (A @ B) + be * C
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max() * max(1, k) ** 0.5


@pytest.mark.parametrize("n", [5, 64, 200, 333, 700])
def test_lu_solve(gpu_ctx, n):
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    # rows scaled so that partial pivoting is exercised
    A *= np.exp(rng.normal(size=(n, 1)))
    X = rng.normal(size=(n, 7)) + 1j * rng.normal(size=(n, 7))
    B = A @ X
    got = gpu_ctx.zgesv_host(A, B)
    resid = np.abs(A @ got - B).max() / (np.abs(A).max() * np.abs(got).max() * n)
    assert resid < 1e-14
    assert np.abs(got - X).max() < 1e-8 * np.abs(X).max()




def test_lu_factors(gpu_ctx):
    """P L U reconstructs A (interchanges applied in order, LAPACK convention)."""
    n = 150
    rng = np.random.default_rng(5)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    LU, piv = gpu_ctx.zgetrf_host(A)
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    PA = A.copy()
    for i in range(n):
        j = piv[i]
        assert i <= j < n
        PA[[i, j]] = PA[[j, i]]
    assert np.abs(L @ U - PA).max() < 1e-12 * np.abs(A).max() * n
