"""Split-FP16 MFMA GEMM (C = alpha A^T B + beta C) against FP64 products.

Each FP32 operand becomes hi + lo FP16 planes under a power-of-two scale; the
three products a_hi b_hi + a_hi b_lo + a_lo b_hi are accumulated in FP32.  The
error bound is normwise: |C - C_ref| <= tol * |alpha| * (|A|^T |B|), with tol a
few times 2^-21 (the dropped lo*lo term and the hi/lo rounding) plus the FP32
accumulation.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP16X3 = 2


def _normwise(out, ref, A, B, alpha, beta, C0):
    bound = abs(alpha) * (np.abs(A).T @ np.abs(B)) + abs(beta) * np.abs(C0)
    return float(np.max(np.abs(out - ref) / np.maximum(bound, 1e-300)))


@pytest.mark.parametrize("M,N,K", [(3, 5, 7), (37, 150, 45), (128, 128, 32), (129, 257, 33),
                                   (256, 256, 64), (1000, 777, 1300),
                                   # few tiles and a long K: split into K-chunks summed in a second pass
                                   (300, 200, 9000), (60, 60, 50000), (1830, 260, 4100)])
def test_gemm16_shapes(ctx, M, N, K):
    rng = np.random.default_rng(M + 10 * N + 100 * K)
    A, B, C0 = rng.normal(size=(K, M)), rng.normal(size=(K, N)), rng.normal(size=(M, N))
    alpha, beta = 0.7, -1.3
    ref = alpha * (A.T @ B) + beta * C0
    out = ctx.gemm_test(FP16X3, "T", "N", alpha, A, B, beta, C0)
    assert _normwise(out, ref, A, B, alpha, beta, C0) < 4e-6


def test_gemm16_accuracy_close_to_fp32(ctx):
    rng = np.random.default_rng(5)
    M, N, K = 300, 200, 2000
    A, B = rng.normal(size=(K, M)), rng.normal(size=(K, N))
    ref = A.T @ B
    z = np.zeros((M, N))
    e16 = np.max(np.abs(ctx.gemm_test(FP16X3, "T", "N", 1.0, A, B, 0.0, z) - ref))
    e32 = np.max(np.abs(ctx.gemm_test(1, "T", "N", 1.0, A, B, 0.0, z) - ref))
    scale = np.max(np.abs(A).T @ np.abs(B))
    assert e16 / scale < 2e-6
    assert e16 < 30 * e32


@pytest.mark.parametrize("sa,sb", [(1e-20, 1e20), (1e15, 1e-12), (3e-30, 7e-3)])
def test_gemm16_scale_invariance(ctx, sa, sb):
    """Per-operand power-of-two scales keep the relative accuracy at any magnitude."""
    rng = np.random.default_rng(11)
    A, B = rng.normal(size=(64, 96)), rng.normal(size=(64, 80))
    ref = A.T @ B
    out = ctx.gemm_test(FP16X3, "T", "N", 1.0 / (sa * sb), A * sa, B * sb, 0.0, np.zeros((96, 80)))
    assert _normwise(out, ref, A, B, 1.0, 0.0, np.zeros((96, 80))) < 4e-6


def test_gemm16_wide_dynamic_range(ctx):
    """Columns spanning 12 orders of magnitude: small entries lose precision only
    relative to the largest, which the normwise bound allows."""
    rng = np.random.default_rng(3)
    K, M, N = 200, 64, 64
    A = rng.normal(size=(K, M)) * np.logspace(-12, 0, M)[None, :]
    B = rng.normal(size=(K, N))
    ref = A.T @ B
    out = ctx.gemm_test(FP16X3, "T", "N", 1.0, A, B, 0.0, np.zeros((M, N)))
    # column-wise relative error stays small for columns within ~1e-4 of the largest
    big = np.abs(A).max(axis=0) > 1e-4
    err = np.abs(out - ref)[big] / (np.abs(A).T @ np.abs(B))[big]
    assert np.max(err) < 4e-6


def test_gemm16_zero_operand_and_beta_zero_nan(ctx):
    rng = np.random.default_rng(2)
    A, B = np.zeros((40, 70)), rng.normal(size=(40, 90))
    C0 = rng.normal(size=(70, 90))
    out = ctx.gemm_test(FP16X3, "T", "N", 1.0, A, B, 2.0, C0)
    assert np.max(np.abs(out - 2.0 * C0)) <= 1e-6 * np.max(np.abs(C0))
    A = rng.normal(size=(40, 70))
    out = ctx.gemm_test(FP16X3, "T", "N", 1.0, A, B, 0.0, np.full((70, 90), np.nan))
    assert np.all(np.isfinite(out))
    # the same through the split-K pass (long K, one tile)
    A, B = rng.normal(size=(6000, 70)), rng.normal(size=(6000, 90))
    out = ctx.gemm_test(FP16X3, "T", "N", 1.0, A, B, 0.0, np.full((70, 90), np.nan))
    assert _normwise(out, A.T @ B, A, B, 1.0, 0.0, np.zeros((70, 90))) < 4e-6


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("M,N,K", [(3, 5, 7), (70, 130, 65), (300, 257, 129)])
def test_gemm16_transposed_split(ctx, ta, tb, M, N, K):
    """Operands that are not k-contiguous pass through the LDS-tiled transposing split."""
    rng = np.random.default_rng(M + N + K)
    A = rng.normal(size=(K, M) if ta == "T" else (M, K))
    B = rng.normal(size=(N, K) if tb == "T" else (K, N))
    C0 = rng.normal(size=(M, N))
    opA = A.T if ta == "T" else A
    opB = B.T if tb == "T" else B
    ref = 0.5 * (opA @ opB) + 0.25 * C0
    out = ctx.gemm_test(FP16X3, ta, tb, 0.5, A, B, 0.25, C0)
    assert _normwise(out, ref, opA.T, opB, 0.5, 0.25, C0) < 4e-6


@pytest.mark.parametrize("K,ld", [(3000, 3003), (3000, 3004), (2001, 2004), (1600, 1601)])
def test_gemm16_tall_operands_with_padded_leading_dimension(ctx, K, ld):
    """Tall operands (K >= 1536 rows) take the column-blocked split: leading dimensions above
    the row count, odd ones (no 16-byte loads) and a last group of 8 rows cut by the end."""
    rng = np.random.default_rng(K + ld)
    M, N = 50, 70
    A, B, C0 = rng.normal(size=(ld, M)), rng.normal(size=(ld, N)), rng.normal(size=(M, N))
    A[K:] = np.nan  # rows beyond K are not part of the operands
    B[K:] = np.nan
    alpha, beta = 1.1, 0.4
    ref = alpha * (A[:K].T @ B[:K]) + beta * C0
    out = ctx.gemm_test(FP16X3, "T", "N", alpha, A, B, beta, C0, K=K)
    assert np.all(np.isfinite(out))
    assert _normwise(out, ref, A[:K], B[:K], alpha, beta, C0) < 4e-6
