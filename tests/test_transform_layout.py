"""CPU check of the dense AO -> MO transform's product sequence (csrc/transform.hip, tei_transform_dev).

Every GEMM the device issues is restated on flat column-major arrays with the same M, N, K, leading
dimensions and batch strides, through every order the dispatch can take (mu then lambda, sigma then
nu, or both outer indices first and then nu and lambda in either order), and checked against the
oracle's transform (Transformation.jl:39-93; ``tei_transform``) on a tensor with no index symmetry
and four distinct coefficient blocks."""
import numpy as np
import pytest

import jues_oracle as O
from test_mp2grad_layout import gemm


def device_sequence(g, C1, C2, C3, C4, stride_error=0):
    """(pq|rs) as [p,q,r,s] through the device's products; stride_error plants a wrong batch stride
    in the last product (a negative control for this check)."""
    n = g.shape[0]
    d1, d2, d3, d4 = C1.shape[1], C2.shape[1], C3.shape[1], C4.shape[1]
    n2, n3 = n * n, n ** 3
    gf = g.ravel(order="F")
    c1, c2, c3, c4 = (c.ravel(order="F").copy() for c in (C1, C2, C3, C4))
    out = np.zeros(d1 * d2 * d3 * d4)
    d12 = d1 * d2
    if d1 <= d4 and d3 < d4:
        pn = d1 * n
        pnr = pn * d3
        X1, X2, X3 = np.zeros(d1 * n3), np.zeros(pnr * n), np.zeros(pnr * d4)
        gemm(1, 0, d1, n3, n, 1.0, c1, n, 0, gf, n, 0, 0.0, X1, d1, 0)
        gemm(0, 0, pn, d3, n, 1.0, X1, pn, pn * n, c3, n, 0, 0.0, X2, pn, pnr, n)
        gemm(0, 0, pnr, d4, n, 1.0, X2, pnr, 0, c4, n, 0, 0.0, X3, pnr, 0)
        gemm(0, 0, d1, d2, n, 1.0, X3, d1, pn + stride_error, c2, n, 0, 0.0, out, d1, d12, d3 * d4)
    elif d1 > d4 and d2 < d1:
        Y1, Y2, X3 = np.zeros(n3 * d4), np.zeros(n * d2 * n * d4), np.zeros(d12 * n * d4)
        gemm(0, 0, n3, d4, n, 1.0, gf, n3, 0, c4, n, 0, 0.0, Y1, n3, 0)
        gemm(0, 0, n, d2, n, 1.0, Y1, n, n2, c2, n, 0, 0.0, Y2, n, n * d2, n * d4)
        gemm(1, 0, d1, d2 * n * d4, n, 1.0, c1, n, 0, Y2, n, 0, 0.0, X3, d1, 0)
        gemm(0, 0, d12, d3, n, 1.0, X3, d12, d12 * n + stride_error, c3, n, 0, 0.0, out, d12, d12 * d3, d4)
    else:
        X2 = np.zeros(d1 * n2 * d4)
        if d1 <= d4:
            X1 = np.zeros(d1 * n3)
            gemm(1, 0, d1, n3, n, 1.0, c1, n, 0, gf, n, 0, 0.0, X1, d1, 0)
            gemm(0, 0, d1 * n2, d4, n, 1.0, X1, d1 * n2, 0, c4, n, 0, 0.0, X2, d1 * n2, 0)
        else:
            Y1 = np.zeros(n3 * d4)
            gemm(0, 0, n3, d4, n, 1.0, gf, n3, 0, c4, n, 0, 0.0, Y1, n3, 0)
            gemm(1, 0, d1, n2 * d4, n, 1.0, c1, n, 0, Y1, n, 0, 0.0, X2, d1, 0)
        if d3 <= d2:
            X3 = np.zeros(d1 * n * d3 * d4)
            gemm(0, 0, d1 * n, d3, n, 1.0, X2, d1 * n, d1 * n2, c3, n, 0, 0.0, X3, d1 * n, d1 * n * d3, d4)
            gemm(0, 0, d1, d2, n, 1.0, X3, d1, d1 * n + stride_error, c2, n, 0, 0.0, out, d1, d12, d3 * d4)
        else:
            X3 = np.zeros(d12 * n * d4)
            gemm(0, 0, d1, d2, n, 1.0, X2, d1, d1 * n, c2, n, 0, 0.0, X3, d1, d12, n * d4)
            gemm(0, 0, d12, d3, n, 1.0, X3, d12, d12 * n + stride_error, c3, n, 0, 0.0, out, d12, d12 * d3, d4)
    return out.reshape((d1, d2, d3, d4), order="F")


# (d1, d2, d3, d4) covering every order: mu then lam (d1 <= d4, d3 < d4), sig then nu (d1 > d4,
# d2 < d1), and the two outer indices first followed by d3 <= d2 or d3 > d2, from either side
@pytest.mark.parametrize("dims", [(2, 5, 3, 4), (5, 2, 4, 3), (3, 2, 5, 4), (2, 5, 4, 4), (4, 5, 2, 3),
                                  (4, 4, 5, 3), (6, 6, 6, 6)])
def test_transform_sequence_matches_oracle(dims):
    n = 6
    rng = np.random.default_rng(11 + sum(dims))
    g = rng.standard_normal((n, n, n, n))
    Cs = [rng.standard_normal((n, d)) for d in dims]
    ref = O.tei_transform(g, *Cs)
    got = device_sequence(g, *Cs)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("dims", [(2, 5, 3, 4), (5, 2, 4, 3), (3, 2, 5, 4), (2, 5, 4, 4)])
def test_transform_sequence_detects_stride_error(dims):
    n = 6
    rng = np.random.default_rng(5)
    g = rng.standard_normal((n, n, n, n))
    Cs = [rng.standard_normal((n, d)) for d in dims]
    ref = O.tei_transform(g, *Cs)
    got = device_sequence(g, *Cs, stride_error=-1)
    assert np.max(np.abs(got - ref)) > 1e-3 * np.max(np.abs(ref))
