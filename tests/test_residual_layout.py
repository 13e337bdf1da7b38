"""The restructured RCCD residual that the device evaluates, restated in NumPy and
checked against the oracle's term-by-term restatement of RCCD.jl:242-256.

The device forms the residual from (ov) x (ov) products in the layout [i,a,j,b]:
    WJ = K + K U/2 - Kx T/2,   WE = -J + Kx Tx/2 + D/2,
    D[kc,jb] = delta_kj Fae[b,c] - delta_cb Fmi[k,j],
    H0 = U WJ + T WE,          Y = Tx WE,
    S[ia,jb] = H0[ia,jb] + Y[ib,ja] + H0[jb,ia] + Y[ja,ib]   (summed by the update kernel),
    R = <ij|ab> + S[ia,jb] + sum_mn t[mn,ab] W'[mn,ij] + sum_ef t[ij,ef] <ab|ef>,
    W' = <mn|ij> + sum_ef <mn|ef> t[ij,ef],
with K[ia,jb] = (ia|jb), Kx[ia,jb] = (ib|ja), J[ia,jb] = (ij|ab), T[ia,jb] = t[i,j,a,b],
Tx[ia,jb] = t[i,j,b,a], U = 2T - Tx.  The Fock-like terms t.Fae^T - Fmi^T.t have no product
of their own: T D and (Tx D)^x each carry half of them up to a transpose.  A dropped term, a
wrong sign or a transposed index in any of these fails the comparison."""
import numpy as np
import pytest

import jues_oracle as O
from jues import synthetic


def _system(n, o, naux, seed):
    s = synthetic.make_system(n, o, n - o, naux, seed, name="layout")
    return s, O.rccd_integrals_from_factors(s.B, s.Co, s.Cv)


def _amplitudes(o, v, seed, scale=0.05):
    t = np.random.default_rng(seed).normal(size=(o, o, v, v)) * scale
    return 0.5 * (t + t.transpose(1, 0, 3, 2))  # t[i,j,a,b] = t[j,i,b,a]


def _device_residual(t, I, fock_in_we=True):
    oovv, ovov, oooo, vvvv = I["oovv"], I["ovov"], I["oooo"], I["vvvv"]
    o, v = t.shape[0], t.shape[2]
    mm = lambda A, B: np.einsum("iakc,kcjb->iajb", A, B)  # (ov) x (ov) product
    # layouts [i,a,j,b]
    K = np.einsum("ijab->iajb", oovv)   # (ia|jb) = <ij|ab>
    Kx = np.einsum("ijba->iajb", oovv)  # (ib|ja)
    J = ovov                            # (ij|ab), stored [i,a,j,b]
    T = np.einsum("ijab->iajb", t)
    Tx = np.einsum("ijba->iajb", t)
    U = 2 * T - Tx
    Ut = 2 * t - t.transpose(1, 0, 2, 3)  # U in [i,j,a,b]
    Fae = -np.einsum("mnef,mnaf->ae", oovv, Ut)
    Fmi = np.einsum("mnef,inef->mi", oovv, Ut)
    WJ = K + mm(K, U) / 2 - mm(Kx, T) / 2
    WE = -J + mm(Kx, Tx) / 2
    if fock_in_we:
        D = np.zeros((o, v, o, v))
        for k in range(o):
            D[k, :, k, :] += Fae.T          # D[k,c,k,b] = Fae[b,c]
        for b in range(v):
            D[:, b, :, b] -= Fmi            # D[k,b,j,b] = -Fmi[k,j]
        WE = WE + D / 2
    H0 = mm(U, WJ) + mm(T, WE)
    Y = mm(Tx, WE)
    # S as the update kernel sums it (H0[jb,ia] and Y[ja,ib] are its LDS-transposed terms)
    S = H0 + Y.transpose(0, 3, 2, 1) + H0.transpose(2, 3, 0, 1) + Y.transpose(2, 1, 0, 3)
    Wp = oooo + np.einsum("mnef,ijef->mnij", oovv, t)
    return (oovv + S.transpose(0, 2, 1, 3) + np.einsum("mnab,mnij->ijab", t, Wp)
            + np.einsum("ijef,abef->ijab", t, vvvv))


@pytest.mark.parametrize("n,o,naux,seed", [(9, 3, 14, 5), (11, 4, 16, 6), (7, 1, 9, 7)])
def test_device_residual_layout_matches_oracle(n, o, naux, seed):
    s, I = _system(n, o, naux, seed)
    t = _amplitudes(s.o, s.v, seed + 100)
    ref = O.rccd_residual_tensor(t, I)
    got = _device_residual(t, I)
    assert np.max(np.abs(got - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_fock_terms_are_carried_by_we():
    """Without the D/2 dressing of WE the Fock-like terms are missing: the comparison has
    the power to see them."""
    s, I = _system(9, 3, 14, 5)
    t = _amplitudes(s.o, s.v, 105)
    ref = O.rccd_residual_tensor(t, I)
    assert np.max(np.abs(_device_residual(t, I, fock_in_we=False) - ref)) > 1e-4 * np.max(np.abs(ref))
