"""The flop counts bench.py prices the MP2 stages with (``mp2_flop_counts``) against the products the
device issues: every GEMM of the CPU restatements of tei_transform_dev and mp2_grad_dev (the same M, N,
K and batch counts as the device calls) is counted as 2 M N K per batch entry, at sizes where the
terms of the closed forms differ from each other, with o < v (the benchmarked case), o > v and o = v
(the transform's three contraction orders for (ia|jb))."""
import pytest

import bench
import test_mp2grad_layout as adjoint_seq
import test_transform_layout as transform_seq


def _count_gemms(monkeypatch, module):
    counted = [0]
    inner = adjoint_seq.gemm

    def gemm(ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch=1):
        counted[0] += 2 * M * N * K * batch
        return inner(ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch)

    monkeypatch.setattr(module, "gemm", gemm)
    return counted


@pytest.mark.parametrize("n,o", [(7, 2), (9, 3), (8, 5), (8, 4)])
def test_transform_flops_match_the_product_sequence(monkeypatch, n, o):
    v = n - o
    g, C, _ = adjoint_seq._system(n, o, 4)
    Co, Cv = C[:, :o], C[:, o:]
    counted = _count_gemms(monkeypatch, transform_seq)
    transform_seq.device_sequence(g, Co, Cv, Co, Cv)
    assert counted[0] == bench.mp2_flop_counts(n, o, v)[0]


@pytest.mark.parametrize("n,o", [(7, 2), (9, 3), (8, 5), (8, 4)])
def test_adjoint_flops_match_the_product_sequence(monkeypatch, n, o):
    v = n - o
    g, C, eps = adjoint_seq._system(n, o, 4)
    counted = _count_gemms(monkeypatch, adjoint_seq)
    adjoint_seq.device_sequence(g, C, eps, o)
    assert counted[0] == bench.mp2_flop_counts(n, o, v)[1]
