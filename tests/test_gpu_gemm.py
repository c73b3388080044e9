"""The batched complex GEMM (3M on MFMA) against numpy complex128 products."""
import ctypes

import numpy as np
import pytest

import wtomo3d_amd as W

pytestmark = pytest.mark.gpu
P = lambda a: a.ctypes.data_as(ctypes.c_void_p)


def _run(prec, M, N, K, mode, lda=None, ldb=None, ldc=None, perm=False, seed=0, var=-1):
    rng = np.random.default_rng(seed)
    ct = np.complex64 if prec == 32 else np.complex128
    lda, ldb, ldc = lda or M, ldb or K, ldc or M
    A = (rng.standard_normal((K, lda)) + 1j * rng.standard_normal((K, lda))).astype(ct)   # col-major M x K
    B = (rng.standard_normal((N, ldb)) + 1j * rng.standard_normal((N, ldb))).astype(ct)
    C = (rng.standard_normal((N, ldc)) + 1j * rng.standard_normal((N, ldc))).astype(ct)
    bp = rng.permutation(K).astype(np.int32) if perm else None
    Am = A[:, :M].T.astype(np.complex128)
    Bm = B[:, :K].T.astype(np.complex128)
    if perm:
        Bm = Bm[bp, :]
    Cm = C[:, :M].T.astype(np.complex128)
    ref = {0: Cm - Am @ Bm, 1: Am @ Bm, 2: Cm + Am @ Bm, 3: -(Am @ Bm)}[mode]
    Cout = C.copy()
    ms = ctypes.c_double()
    rc = W.lib().wt_gemm_host_var(prec, var, M, N, K, P(A), lda, P(B), ldb, P(Cout), ldc, mode,
                                  P(bp) if perm else None, 1, ctypes.byref(ms))
    assert rc == 0
    got = Cout[:, :M].T.astype(np.complex128)
    scale = np.abs(Am) @ np.abs(Bm) + np.abs(Cm)
    err = np.max(np.abs(got - ref) / scale)
    assert np.array_equal(Cout[:, M:], C[:, M:])  # padding rows untouched
    return err


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("shape", [(1, 1, 1), (64, 64, 64), (130, 70, 9), (257, 129, 300), (33, 500, 17),
                                   (600, 64, 200), (513, 37, 70)])
@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_gemm_modes_shapes(prec, shape, mode):
    M, N, K = shape
    err = _run(prec, M, N, K, mode, lda=M + 3, ldb=K + 1, ldc=M + 5)
    assert err <= (2e-6 if prec == 32 else 1e-14)


@pytest.mark.parametrize("prec", [32, 64])
def test_gemm_row_permuted_b(prec):
    err = _run(prec, 64, 200, 64, 1, perm=True, seed=3)
    assert err <= (2e-6 if prec == 32 else 1e-14)


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("var", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("shape", [(1, 1, 1), (64, 64, 64), (130, 70, 9), (257, 129, 300), (65, 65, 8)])
def test_gemm_tile_variants(prec, var, shape):
    """Every tile variant (the autotuner picks among them per launch) gives
    the product to the precision's rounding; mode SUB with padded leading
    dimensions.  Variants 4 and 5 in fp32 are the three-way bf16 splits."""
    M, N, K = shape
    err = _run(prec, M, N, K, 0, lda=M + 1, ldb=K + 2, ldc=M + 3, seed=11, var=var)
    assert err <= (2e-6 if prec == 32 else 1e-14)


@pytest.mark.parametrize("var", [4, 5])
@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("shape", [(300, 260, 2000), (129, 257, 33), (1000, 64, 512)])
def test_split_products_as_accurate_as_fp32(var, mode, shape):
    """The bf16 three-way splits (fp32 variants 4 and 5) against complex128:
    the error, relative to |A||B| + |C| entry by entry, stays within 1.5x
    that of the fp32 matrix-core product (variant 0) and below 2e-6, on long
    K and ragged tiles, in every mode."""
    M, N, K = shape
    e0 = _run(32, M, N, K, mode, lda=M + 2, ldb=K + 1, ldc=M + 1, seed=5, var=0)
    e4 = _run(32, M, N, K, mode, lda=M + 2, ldb=K + 1, ldc=M + 1, seed=5, var=var)
    assert e4 <= 2e-6 and e4 <= 1.5 * e0 + 1e-8, (e4, e0)


@pytest.mark.parametrize("var", [4, 5])
def test_split_row_permuted_b_and_extreme_scales(var):
    """The split variants with a row-permuted B, and with operands spanning
    2^-60 to 2^60 (every split part stays normal in bf16's fp32 exponent
    range)."""
    assert _run(32, 64, 200, 64, 1, perm=True, seed=3, var=var) <= 2e-6
    rng = np.random.default_rng(9)
    M, N, K = 70, 90, 300
    A = (rng.standard_normal((K, M)) * 2.0 ** rng.integers(-60, 60, (K, M))).astype(np.float32)
    A = (A + 1j * rng.standard_normal((K, M))).astype(np.complex64)
    B = (rng.standard_normal((N, K)) + 1j * rng.standard_normal((N, K))).astype(np.complex64)
    C = np.zeros((N, M), np.complex64)
    ms = ctypes.c_double()
    assert W.lib().wt_gemm_host_var(32, var, M, N, K, P(A), M, P(B), K, P(C), M, 1, None, 1, ctypes.byref(ms)) == 0
    Am, Bm = A.T.astype(np.complex128), B.T.astype(np.complex128)
    err = np.max(np.abs(C.T - Am @ Bm) / (np.abs(Am) @ np.abs(Bm)))
    assert err <= 2e-6
