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


def _pairs(n, strict):
    """Pairs (p<=q), or (p<q), in the device's order: index p + q(q+1)/2 (p + q(q-1)/2)."""
    return [(p, q) for q in range(n) for p in range(q if strict else q + 1)]


def _packed_ladders(t, I):
    """Both ladders on packed pairs, as the device forms them, unpacked the way the update
    kernel reads them:
      Tp[(i<=j),(e<=f)] = t[ijef] + t[ijfe] (e<f), t[ijee];  Tm[(i<j),(e<f)] = t[ijef] - t[ijfe]
      Vp[(e<=f),(a<=b)] = (<ab|ef> + <ab|fe>)/2;           Vm[(e<f),(a<b)] = (<ab|ef> - <ab|fe>)/2
      KP[(k<=l),(c<=d)] = K[klcd] + K[kldc] (k<l), K[kkcd];  KM = K[klcd] - K[kldc],  K = <kl|cd>
      W+^T[(ij),(kl)] = O+ + Tp KP^T,  W-^T = O- + Tm KM^T  (O+- packs <kl|ij> like KP / KM)
      Lp = Tp Vp + W+^T Tq / 2,  Lm = Tm Vm + W-^T Tm / 2  (Tq: Tp with its (a,a) columns doubled)
      R[ijab] += Lp[(ij),(ab)] + s_ij s_ab Lm[(ij),(ab)]  (s = +1 for p<q, -1 for p>q, 0 for p=q)"""
    oovv, oooo, vvvv = I["oovv"], I["oooo"], I["vvvv"]
    o, v = t.shape[0], t.shape[2]
    ole, olt, vle, vlt = _pairs(o, False), _pairs(o, True), _pairs(v, False), _pairs(v, True)
    mat = lambda rows, cols, f: np.array([[f(r, c) for c in cols] for r in rows]).reshape(len(rows), len(cols))
    Tp = mat(ole, vle, lambda ij, ef: t[ij + ef] if ef[0] == ef[1] else t[ij + ef] + t[ij + ef[::-1]])
    Tm = mat(olt, vlt, lambda ij, ef: t[ij + ef] - t[ij + ef[::-1]])
    Vp = mat(vle, vle, lambda ef, ab: 0.5 * (vvvv[ab + ef] + vvvv[ab + ef[::-1]]))
    Vm = mat(vlt, vlt, lambda ef, ab: 0.5 * (vvvv[ab + ef] - vvvv[ab + ef[::-1]]))
    KP = mat(ole, vle, lambda kl, cd: oovv[kl + cd] if kl[0] == kl[1] else oovv[kl + cd] + oovv[kl + cd[::-1]])
    KM = mat(olt, vlt, lambda kl, cd: oovv[kl + cd] - oovv[kl + cd[::-1]])
    WPT = mat(ole, ole, lambda ij, kl: oooo[kl + ij] if kl[0] == kl[1] else oooo[kl + ij] + oooo[kl[::-1] + ij])
    WMT = mat(olt, olt, lambda ij, kl: oooo[kl + ij] - oooo[kl[::-1] + ij])
    WPT = WPT + Tp @ KP.T
    WMT = WMT + Tm @ KM.T
    Tq = Tp * np.array([2.0 if a == b else 1.0 for (a, b) in vle])[None, :]
    Lp = Tp @ Vp + 0.5 * WPT @ Tq
    Lm = Tm @ Vm + 0.5 * WMT @ Tm
    le = lambda p, q: min(p, q) + max(p, q) * (max(p, q) + 1) // 2
    lt = lambda p, q: min(p, q) + max(p, q) * (max(p, q) - 1) // 2
    sg = lambda p, q: (p < q) - (p > q)
    R = np.zeros((o, o, v, v))
    for i in range(o):
        for j in range(o):
            for a in range(v):
                for b in range(v):
                    r = Lp[le(i, j), le(a, b)]
                    if sg(i, j) and sg(a, b):
                        r += sg(i, j) * sg(a, b) * Lm[lt(i, j), lt(a, b)]
                    R[i, j, a, b] = r
    return R


def _device_residual(t, I, fock_in_we=True, packed=False):
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
    if packed:
        ladders = _packed_ladders(t, I)
    else:
        Wp = oooo + np.einsum("mnef,ijef->mnij", oovv, t)
        ladders = np.einsum("mnab,mnij->ijab", t, Wp) + np.einsum("ijef,abef->ijab", t, vvvv)
    return oovv + S.transpose(0, 2, 1, 3) + ladders


@pytest.mark.parametrize("packed", [False, True])
@pytest.mark.parametrize("n,o,naux,seed", [(9, 3, 14, 5), (11, 4, 16, 6), (7, 1, 9, 7)])
def test_device_residual_layout_matches_oracle(n, o, naux, seed, packed):
    s, I = _system(n, o, naux, seed)
    t = _amplitudes(s.o, s.v, seed + 100)
    ref = O.rccd_residual_tensor(t, I)
    got = _device_residual(t, I, packed=packed)
    assert np.max(np.abs(got - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_fock_terms_are_carried_by_we():
    """Without the D/2 dressing of WE the Fock-like terms are missing: the comparison has
    the power to see them."""
    s, I = _system(9, 3, 14, 5)
    t = _amplitudes(s.o, s.v, 105)
    ref = O.rccd_residual_tensor(t, I)
    assert np.max(np.abs(_device_residual(t, I, fock_in_we=False) - ref)) > 1e-4 * np.max(np.abs(ref))
