"""CPU check of the dense MP2 adjoint's product sequence (csrc/mp2grad.hip, mp2_grad_dev).

Every GEMM and permutation the device issues is restated here on flat column-major arrays with the
same M, N, K, leading dimensions and batch strides: the forward (ia|jb) transform with its kept
intermediates (i nu|lam b) and (i nu|j b), the extra half-transform (mu a|j sig), the four dE/dC
positions and the back-transform to dE/dg, checked against the oracle's einsum gradient (RMP2.jl:33-37
differentiated; ``rmp2_grad``) on small tensors with no index symmetry."""
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


def gemm_batch_sum(ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, batch):
    """C = alpha sum_z op(A_z) op(B_z) + beta C: the products to a workspace, then summed over z."""
    W = np.zeros(batch * M * N)
    gemm(ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, 0.0, W, M, M * N, batch)
    c = _mat(C, 0, M, N, ldc)
    c[...] = W.reshape((M, N, batch), order="F").sum(axis=2) + (beta * c if beta != 0 else 0.0)


def permute(src, dims, perm, dst, alpha=1.0, beta=0.0):
    """dst = alpha permutedims(src, perm) + beta dst (out dim k = in dim perm[k]), column-major."""
    x = src[:int(np.prod(dims))].reshape(dims, order="F").transpose(perm)
    out = dst[:x.size].reshape(x.shape, order="F")
    out[...] = alpha * x + beta * out


def _system(n, o, seed):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, n, n, n))                 # no (mu nu) / (lam sig) / pair symmetry
    C = rng.standard_normal((n, n))
    eps = np.sort(rng.uniform(-2, 2, n))
    eps[o:] += 3.0
    return g, C, eps


def device_sequence(g, C, eps, o, pos4_stride_error=0):
    """dE/dC and dE/dg through the device's products; G from the forward K as mp2_g_kernel forms it."""
    n = C.shape[0]
    v = n - o
    n2, n3, ov = n * n, n ** 3, o * v
    nov, on, N = n * ov, o * n, ov * ov
    gf = g.ravel(order="F")
    Co, Cv = C[:, :o].ravel(order="F").copy(), C[:, o:].ravel(order="F").copy()
    z = np.zeros
    # forward (ia|jb) with X2 = (i nu|lam b) as [i,nu,lam,b] kept, the last two quarters batched
    # (X3 = (i nu|j b) as [i,nu,j,b]) and Y3 = X3 as [nu,i,b,j] kept
    Z1, X2, X3, Y3, K = z(o * n3), z(nov * n), z(nov * o), z(nov * o), z(N)
    gemm(1, 0, o, n3, n, 1.0, Co, n, 0, gf, n, 0, 0.0, Z1, o, 0)
    gemm(0, 0, o * n2, v, n, 1.0, Z1, o * n2, 0, Cv, n, 0, 0.0, X2, o * n2, 0)
    gemm(0, 0, on, o, n, 1.0, X2, on, on * n, Co, n, 0, 0.0, X3, on, on * o, v)
    gemm(0, 0, o, v, n, 1.0, X3, o, on, Cv, n, 0, 0.0, K, o, ov, o * v)
    permute(X3, (o, n, o, v), (1, 0, 3, 2), Y3)
    K4 = K.reshape(o, v, o, v, order="F")
    D = O.form_Dijab(eps, o, v).transpose(0, 2, 1, 3)
    G = (2.0 * (2.0 * K4 - K4.transpose(0, 3, 2, 1)) / D).ravel(order="F")
    dCo, dCv = z(n * o), z(n * v)
    # position 2
    Gp2 = z(N)
    permute(G, (o, v, o, v), (0, 3, 2, 1), Gp2)
    gemm(0, 0, n, v, ov * o, 1.0, Y3, n, 0, Gp2, ov * o, 0, 0.0, dCv, n, 0)
    # position 3
    Gt = z(on * ov)
    gemm(0, 1, o, n, v, 1.0, G, o, ov, Cv, n, 0, 0.0, Gt, o, on, ov)
    gemm_batch_sum(1, 0, n, o, on, 1.0, X2, on, on * n, Gt, on, on * o, 0.0, dCo, n, v)
    # (mu a|j sig): X[mu,nu,j,sig] batched over sig, then Y[mu,a,(j,sig)] batched over (j,sig)
    X, Y = z(n3 * o), z(nov * n)
    gemm(0, 0, n2, o, n, 1.0, gf, n2, n3, Co, n, 0, 0.0, X, n2, n2 * o, n)
    gemm(0, 0, n, v, n, 1.0, X, n, n2, Cv, n, 0, 0.0, Y, n, n * v, on)
    # position 1
    T1 = z(ov * on)
    gemm(0, 1, ov * o, n, v, 1.0, G, ov * o, 0, Cv, n, 0, 0.0, T1, ov * o, 0)
    gemm(0, 1, n, o, nov, 1.0, Y, n, 0, T1, o, 0, 1.0, dCo, n, 0)
    # position 4 (pos4_stride_error shifts the second product's leading dimension: a planted mistake)
    Yt = z(ov * on + n * v)
    gemm(1, 0, o, nov, n, 1.0, Co, n, 0, Y, n, 0, 0.0, Yt, o, 0)
    gemm(1, 0, n, v, ov * o, 1.0, Yt, ov * o + pos4_stride_error, 0, G, ov * o, 0, 1.0, dCv, n, 0)
    # back-transform
    P2, Q, dg = z(on * n * v), z(on * n2), z(n2 * n2)
    gemm(0, 1, on, n, o, 1.0, Gt, on, on * o, Co, n, 0, 0.0, P2, on, on * n, v)
    gemm(0, 1, on * n, n, v, 1.0, P2, on * n, 0, Cv, n, 0, 0.0, Q, on * n, 0)
    gemm(0, 0, n, n3, o, 1.0, Co, n, 0, Q, o, 0, 0.0, dg, n, 0)
    dC = np.concatenate([dCo.reshape(n, o, order="F"), dCv.reshape(n, v, order="F")], axis=1)
    return dC, dg.reshape(n, n, n, n, order="F")


@pytest.mark.parametrize("n,o,seed", [(12, 3, 0), (11, 4, 1), (9, 6, 2)])
def test_dense_adjoint_sequence_matches_oracle(n, o, seed):
    g, C, eps = _system(n, o, seed)
    _, dC_ref, dg_ref = O.rmp2_grad(g, C, eps, o)
    dC, dg = device_sequence(g, C, eps, o)
    assert np.max(np.abs(dC - dC_ref)) < 1e-10 * np.max(np.abs(dC_ref))
    assert np.max(np.abs(dg - dg_ref)) < 1e-10 * np.max(np.abs(dg_ref))


def test_sequence_check_catches_a_wrong_stride():
    """The restatement is not vacuous: position 4's product with a leading dimension off by one
    misses the oracle's dE/dCv."""
    n, o = 10, 3
    g, C, eps = _system(n, o, 3)
    _, dC_ref, _ = O.rmp2_grad(g, C, eps, o)
    dC, _ = device_sequence(g, C, eps, o, pos4_stride_error=1)
    assert np.max(np.abs(dC[:, o:] - dC_ref[:, o:])) > 1e-3 * np.max(np.abs(dC_ref))
    assert np.max(np.abs(dC[:, :o] - dC_ref[:, :o])) < 1e-10 * np.max(np.abs(dC_ref))
