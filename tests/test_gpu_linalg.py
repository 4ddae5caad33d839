"""Dense complex128 kernels of the HIP library (ZGEMM on FP64 MFMA,
tournament-pivoted LU + triangular solves, the block-local pivoted LU of the
solve) against numpy, through the C ABI."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mnk", [(1, 1, 1), (16, 16, 4), (64, 64, 16), (65, 130, 17), (200, 64, 300), (257, 31, 64),
                                 (3000, 8, 2900), (777, 3, 5000), (64, 1, 63), (130, 8, 0)])
def test_zgemm(gpu_ctx, mnk):
    """MFMA tiles for wide products; at most 8 columns the split-K skinny
    kernel (partials added in a fixed order)."""
    m, n, k = mnk
    rng = np.random.default_rng(m * 7 + n)
    A = rng.normal(size=(m, k)) + 1j * rng.normal(size=(m, k))
    B = rng.normal(size=(k, n)) + 1j * rng.normal(size=(k, n))
    C = rng.normal(size=(m, n)) + 1j * rng.normal(size=(m, n))
    al, be = 0.7 - 0.2j, -1.0 + 0.5j
    got = gpu_ctx.zgemm_host(A, B, C, al, be)
    ref = al * (A @ B) + be * C
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max() * max(1, k) ** 0.5
    if n <= 8:   # deterministic: the same bits on a repeat
        assert np.array_equal(gpu_ctx.zgemm_host(A, B, C, al, be), got)


@pytest.mark.parametrize("n,nrhs", [(5, 7), (64, 7), (200, 7), (333, 7), (700, 7), (700, 1), (1000, 8), (333, 20),
                                    (1000, 9)])
def test_lu_solve(gpu_ctx, n, nrhs):
    """Substitutions through the diagonal-block inverses against numpy: up to 8
    right-hand sides a launch per 64-row block (every workgroup forms the
    block's solution, the others update their rows below it), more through
    blocked GEMMs."""
    rng = np.random.default_rng(n + nrhs)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    # rows scaled so that partial pivoting is exercised
    A *= np.exp(rng.normal(size=(n, 1)))
    X = rng.normal(size=(n, nrhs)) + 1j * rng.normal(size=(n, nrhs))
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


def _block_local_lu(A, nb=64):
    """Right-looking LU whose pivot for column j is searched only in rows
    [j, end of j's nb-block) (first maximum of |re|+|im|), with interchanges
    of whole rows and elimination of every row below: what the blocked GPU
    factorisation computes in exact arithmetic."""
    A = np.array(A, dtype=np.complex128)
    n = A.shape[0]
    piv = np.zeros(n, dtype=np.int64)
    for j in range(n):
        end = min((j // nb + 1) * nb, n)
        col = np.abs(A[j:end, j].real) + np.abs(A[j:end, j].imag)
        p = j + int(np.argmax(col))
        piv[j] = p
        if p != j:
            A[[j, p]] = A[[p, j]]
        A[j + 1:, j] /= A[j, j]
        A[j + 1:, j + 1:] -= np.outer(A[j + 1:, j], A[j, j + 1:])
    return A, piv


@pytest.mark.parametrize("n", [5, 64, 200, 333])
def test_block_local_lu(gpu_ctx, n):
    """bem_zgetrf_local (the solve's default LU, pivoting inside each 64 x 64
    diagonal block) against the numpy restatement above: the same
    interchanges, factors within 1e-12, P A = L U, and the growth factor
    max(1, max|L|) max|U| / max|A| it reports."""
    rng = np.random.default_rng(100 + n)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    A *= np.exp(rng.normal(size=(n, 1)))   # rows scaled so that pivoting matters
    LU, piv, growth = gpu_ctx.zgetrf_local_host(A)
    ref, rpiv = _block_local_lu(A)
    assert np.array_equal(piv, rpiv)
    assert all((i // 64) * 64 <= piv[i] < min((i // 64 + 1) * 64, n) for i in range(n))
    scale = np.abs(A).max()
    assert np.abs(LU - ref).max() <= 1e-12 * scale * n
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    PA = A.copy()
    for i in range(n):
        PA[[i, piv[i]]] = PA[[piv[i], i]]
    assert np.abs(L @ U - PA).max() <= 1e-12 * scale * n
    l1 = lambda M: (np.abs(M.real) + np.abs(M.imag)).max()
    want = max(1.0, l1(np.tril(LU, -1))) * l1(U) / l1(A)
    assert abs(growth - want) <= 1e-12 * growth


def test_block_local_lu_singular_block(gpu_ctx):
    """A zero column inside the third diagonal block: BEM_E_SINGULAR."""
    from bem_amd import BEM_E_SINGULAR
    from bem_amd.runtime import BemError
    n = 150
    rng = np.random.default_rng(7)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    A[:, 130] = 0.0
    with pytest.raises(BemError) as e:
        gpu_ctx.zgetrf_local_host(A)
    assert e.value.code == BEM_E_SINGULAR


ZTILES = [32, 64, 256, 1032, 2032, 1033, 1034, 1064, 2064, 1065, 2034, 1066, 1128, 1129]


@pytest.mark.parametrize("tile", ZTILES)
def test_zgemm_tile_variants(gpu_ctx, tile):
    """Every complex128 tile variant the autotuner may pick, in every shape
    class (narrow, rank update, general)."""
    saved = gpu_ctx.gemm_tiles()
    try:
        gpu_ctx.gemm_tiles([tile, tile, tile] + saved[3:])
        rng = np.random.default_rng(tile)
        for m, n, k in [(200, 64, 64), (257, 190, 64), (300, 301, 129)]:
            A = rng.normal(size=(m, k)) + 1j * rng.normal(size=(m, k))
            B = rng.normal(size=(k, n)) + 1j * rng.normal(size=(k, n))
            C = rng.normal(size=(m, n)) + 1j * rng.normal(size=(m, n))
            got = gpu_ctx.zgemm_host(A, B, C, 1.0 - 0.5j, 0.25j)
            ref = (1.0 - 0.5j) * (A @ B) + 0.25j * C
            assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max() * k ** 0.5, (tile, m, n, k)
    finally:
        gpu_ctx.gemm_tiles(saved)


CTILES = [1, 2, 3, 4, 5, 11, 12, 13, 14, 15, 16, 17, 21, 23, 26]


@pytest.mark.parametrize("tile", CTILES)
def test_cgemm_tile_variants(gpu_ctx, tile):
    """complex64 variants: four real products (1..5) and three (11..15), whose
    error bound u (|Ar| + |Ai|)(|Br| + |Bi|) per term is of the same order."""
    saved = gpu_ctx.gemm_tiles()
    try:
        gpu_ctx.gemm_tiles(saved[:3] + [tile, tile, tile])
        rng = np.random.default_rng(10 + tile)
        for m, n, k in [(200, 64, 64), (257, 190, 64), (300, 301, 129)]:
            A = (rng.normal(size=(m, k)) + 1j * rng.normal(size=(m, k))).astype(np.complex64)
            B = (rng.normal(size=(k, n)) + 1j * rng.normal(size=(k, n))).astype(np.complex64)
            got = gpu_ctx.cgemm_host(A, B).astype(np.complex128)
            ref = A.astype(np.complex128) @ B.astype(np.complex128)
            assert np.abs(got - ref).max() <= 8e-6 * np.abs(ref).max() * k ** 0.5, (tile, m, n, k)
    finally:
        gpu_ctx.gemm_tiles(saved)


def test_autotune_cache_roundtrip(tmp_path):
    """The autotuner measures every variant once, writes its JSON cache, and a
    second context loads the same table from it without measuring."""
    import json
    import bem_amd
    path = str(tmp_path / "tiles.json")
    with bem_amd.Context(0, lanes=1) as ctx:
        tiles = ctx.autotune(path)
    d = json.load(open(path))
    assert d["zgemm"] == tiles[:3] and d["cgemm"] == tiles[3:]
    assert all(t in ZTILES for t in tiles[:3]) and all(t in CTILES for t in tiles[3:])
    with bem_amd.Context(0, lanes=1) as ctx2:
        assert ctx2.autotune(path) == tiles
        assert ctx2.gemm_tiles() == tiles
