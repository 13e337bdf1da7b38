"""CPU check of the dense MP2 adjoint's product sequence (csrc/mp2grad.hip, mp2_grad_dev).

Every GEMM and permutation the device issues is restated here on flat column-major arrays with the
same M, N, K, leading dimensions and batch strides, so the buffer layouts of the half-transforms
(H12 = (mu nu|jb), H34 = (ia|lam sig)), the four dE/dC slots and the back-transform to dE/dg are
checked against the oracle's einsum gradient (RMP2.jl:33-37 differentiated; ``rmp2_grad``) on a
small tensor with no index symmetry."""
import numpy as np
import pytest
from numpy.lib.stride_tricks import as_strided

import jues_oracle as O


def _mat(buf, off, rows, cols, ld):
    """rows x cols column-major view of the flat buffer at element offset off, leading dimension ld."""
    s = buf.itemsize
    return as_strided(buf[off:], shape=(rows, cols), strides=(s, ld * s))


def gemm(ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch=1):
    """C_z = alpha op(A_z) op(B_z) + beta C_z, z < batch (the library's gemm argument order)."""
    for z in range(batch):
        a = _mat(A, z * sA, K, M, lda).T if ta else _mat(A, z * sA, M, K, lda)
        b = _mat(B, z * sB, N, K, ldb).T if tb else _mat(B, z * sB, K, N, ldb)
        c = _mat(C, z * sC, M, N, ldc)
        c[...] = alpha * (a @ b) + (beta * c if beta != 0 else 0.0)


def permute(src, dims, perm, dst, alpha=1.0, beta=0.0):
    """dst = alpha permutedims(src, perm) + beta dst (out dim k = in dim perm[k]), column-major."""
    x = src[:int(np.prod(dims))].reshape(dims, order="F").transpose(perm)
    out = dst[:x.size].reshape(x.shape, order="F")
    out[...] = alpha * x + beta * out


@pytest.mark.parametrize("n,o,seed", [(12, 3, 0), (11, 4, 1)])
def test_dense_adjoint_sequence_matches_oracle(n, o, seed):
    rng = np.random.default_rng(seed)
    v = n - o
    g = rng.standard_normal((n, n, n, n))                 # no (mu nu) / (lam sig) / pair symmetry
    C = rng.standard_normal((n, n))
    eps = np.sort(rng.uniform(-2, 2, n))
    eps[o:] += 3.0
    E, dC_ref, dg_ref = O.rmp2_grad(g, C, eps, o)
    Co, Cv = np.asfortranarray(C[:, :o]).ravel(order="F"), np.asfortranarray(C[:, o:]).ravel(order="F")
    gf = g.ravel(order="F")
    n2, n3, ov = n * n, n ** 3, o * v
    # G = dE/dK, [i,a,j,b]
    K = O.tei_transform(g, C[:, :o], C[:, o:], C[:, :o], C[:, o:])
    D = O.form_Dijab(eps, o, v).transpose(0, 2, 1, 3)
    G = (2.0 * (2.0 * K - K.transpose(0, 3, 2, 1)) / D).ravel(order="F")
    z = lambda m: np.zeros(m)

    # H12[mu,nu,j,b]: X[mu,nu,j,sig] = sum_lam g Co (batched over sig), then X Cv
    X, H12 = z(n2 * o * n), z(n2 * ov)
    gemm(0, 0, n2, o, n, 1.0, gf, n2, n3, Co, n, 0, 0.0, X, n2, n2 * o, n)
    gemm(0, 0, n2 * o, v, n, 1.0, X, n2 * o, 0, Cv, n, 0, 0.0, H12, n2 * o, 0)
    # H34[i,a,lam,sig]: Z[i,nu,lam,sig] = Co^T g, then Z Cv batched over (lam,sig)
    Z, H34 = z(o * n3), z(n2 * ov)
    gemm(1, 0, o, n3, n, 1.0, Co, n, 0, gf, n, 0, 0.0, Z, o, 0)
    gemm(0, 0, o, v, n, 1.0, Z, o, o * n, Cv, n, 0, 0.0, H34, o, ov, n2)
    dCo, dCv = z(n * o), z(n * v)
    # slot 1
    Gt = z(o * n * ov)
    gemm(0, 1, o, n, v, 1.0, G, o, ov, Cv, n, 0, 0.0, Gt, o, o * n, ov)
    gemm(0, 1, n, o, n * ov, 1.0, H12, n, 0, Gt, o, 0, 0.0, dCo, n, 0)
    # slot 2
    CoT, Gp, M2 = z(o * n), z(ov * ov), z(ov * o * n)
    permute(Co, (n, o), (1, 0), CoT)
    permute(G, (o, v, ov), (0, 2, 1), Gp)
    gemm(0, 0, o, ov, n, 1.0, CoT, o, 0, H12, n2, n, 0.0, M2, o, o * ov, n)
    gemm(1, 0, n, v, o * ov, 1.0, M2, o * ov, 0, Gp, o * ov, 0, 0.0, dCv, n, 0)
    # slot 3
    on = o * n
    Gq, G3q, W, ones, dC3t = z(ov * ov), z(ov * o * n), z(o * n2), np.ones(n), z(on)
    permute(G, (ov, o, v), (1, 0, 2), Gq)
    gemm(0, 1, o * ov, n, v, 1.0, Gq, o * ov, 0, Cv, n, 0, 0.0, G3q, o * ov, 0)
    gemm(0, 0, o, n, ov, 1.0, G3q, o, o * ov, H34, ov, ov * n, 0.0, W, o, on, n)
    gemm(0, 0, on, 1, n, 1.0, W, on, 0, ones, n, 0, 0.0, dC3t, on, 0)
    permute(dC3t, (o, n), (1, 0), dCo, 1.0, 1.0)
    # slot 4
    M4 = z(ov * o * n)
    gemm(0, 0, ov, o, n, 1.0, H34, ov, ov * n, Co, n, 0, 0.0, M4, ov, ov * o, n)
    gemm(1, 0, n, v, ov * o, 1.0, M4, ov * o, 0, G, ov * o, 0, 1.0, dCv, n, 0)
    dC = np.concatenate([dCo.reshape(n, o, order="F"), dCv.reshape(n, v, order="F")], axis=1)
    assert np.max(np.abs(dC - dC_ref)) < 1e-10 * np.max(np.abs(dC_ref))
    # back-transform: P2[(i,nu),lam,b] batched over b, Q, then dg = Co Q
    P2, Q, dg = z(on * n * v), z(on * n2), z(n2 * n2)
    gemm(0, 1, on, n, o, 1.0, Gt, on, on * o, Co, n, 0, 0.0, P2, on, on * n, v)
    gemm(0, 1, on * n, n, v, 1.0, P2, on * n, 0, Cv, n, 0, 0.0, Q, on * n, 0)
    gemm(0, 0, n, n3, o, 1.0, Co, n, 0, Q, o, 0, 0.0, dg, n, 0)
    assert np.max(np.abs(dg.reshape(n, n, n, n, order="F") - dg_ref)) < 1e-10 * np.max(np.abs(dg_ref))


def test_sequence_check_catches_a_wrong_stride():
    """The restatement is not vacuous: slot 4 with the batch stride of H34 off by one block of
    (ia) misses the oracle's dE/dCv."""
    rng = np.random.default_rng(3)
    n, o = 10, 3
    v = n - o
    g = rng.standard_normal((n, n, n, n))
    C = rng.standard_normal((n, n))
    eps = np.sort(rng.uniform(-2, 2, n))
    eps[o:] += 3.0
    ov = o * v
    gf, Co, Cv = g.ravel(order="F"), C[:, :o].ravel(order="F"), C[:, o:].ravel(order="F")
    K = O.tei_transform(g, C[:, :o], C[:, o:], C[:, :o], C[:, o:])
    D = O.form_Dijab(eps, o, v).transpose(0, 2, 1, 3)
    G = (2.0 * (2.0 * K - K.transpose(0, 3, 2, 1)) / D).ravel(order="F")
    Z, H34 = np.zeros(o * n ** 3), np.zeros(n * n * ov)
    gemm(1, 0, o, n ** 3, n, 1.0, Co, n, 0, gf, n, 0, 0.0, Z, o, 0)
    gemm(0, 0, o, v, n, 1.0, Z, o, o * n, Cv, n, 0, 0.0, H34, o, ov, n * n)

    def slot4(sA):
        M4, dCv = np.zeros(ov * o * n), np.zeros(n * v)
        gemm(0, 0, ov, o, n, 1.0, H34, ov, sA, Co, n, 0, 0.0, M4, ov, ov * o, n - 1)
        gemm(1, 0, n, v, ov * o, 1.0, M4, ov * o, 0, G, ov * o, 0, 0.0, dCv, n, 0)
        return dCv

    # the oracle's slot-4 term alone: h4 contracted with g
    h4 = np.einsum("iajb,mi,na,lj->mnlb", G.reshape(o, v, o, v, order="F"), C[:, :o], C[:, o:], C[:, :o],
                   optimize=True)
    ref = np.einsum("mnlb,mnls->sb", h4, g, optimize=True)
    # (n - 1 batches of sig: the last sig is dropped in both so the comparison stays one-to-one)
    good, bad = slot4(ov * n), slot4(ov * n - ov)
    ref_trunc = ref.copy()
    ref_trunc[n - 1] = 0.0
    assert np.max(np.abs(good.reshape(n, v, order="F") - ref_trunc)) < 1e-10 * np.max(np.abs(ref))
    assert np.max(np.abs(bad.reshape(n, v, order="F") - ref_trunc)) > 1e-3 * np.max(np.abs(ref))
