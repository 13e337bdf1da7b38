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


def _device_residual_delta(t, t0, I):
    """The difference R(t) - R(t0) as the device's defect-correction stage forms it from
    d = t - t0, the layouts of t and the anchor's intermediates WJ0, WE0, W'0:
      dWJ = K U(d)/2 - Kx T(d)/2,  dWE = Kx Tx(d)/2 + D(dFae, dFmi)/2  (Fae, Fmi are linear in t),
      dH0 = T(d) WE0 + T(t) dWE + U(d) WJ0 + U(t) dWJ,   dY = Tx(d) WE0 + Tx(t) dWE,
      dS summed as S, the pp-ladder of d, and the hole-hole difference
      t.W'(t) - t0.W'0 = t.dW' + d.W'0 with dW' = <mn|ef> d[ij,ef]."""
    oovv, ovov, oooo, vvvv = I["oovv"], I["ovov"], I["oooo"], I["vvvv"]
    o, v = t.shape[0], t.shape[2]
    mm = lambda A, B: np.einsum("iakc,kcjb->iajb", A, B)  # noqa: E731
    K = np.einsum("ijab->iajb", oovv)
    Kx = np.einsum("ijba->iajb", oovv)
    J = ovov
    lay = lambda x: (np.einsum("ijab->iajb", x), np.einsum("ijba->iajb", x))  # noqa: E731
    d = t - t0
    (T0, Tx0), (Td, Txd) = lay(t0), lay(d)
    U0, Ud = 2 * T0 - Tx0, 2 * Td - Txd
    Ut0 = 2 * t0 - t0.transpose(1, 0, 2, 3)
    Fae0 = -np.einsum("mnef,mnaf->ae", oovv, Ut0)
    Fmi0 = np.einsum("mnef,inef->mi", oovv, Ut0)
    Dm = lambda Fa, Fm: (np.einsum("kj,cb->kcjb", np.eye(o), Fa.T)  # noqa: E731
                         - np.einsum("cb,kj->kcjb", np.eye(v), Fm))
    WJ0 = K + mm(K, U0) / 2 - mm(Kx, T0) / 2
    WE0 = -J + mm(Kx, Tx0) / 2 + Dm(Fae0, Fmi0) / 2
    Utd = 2 * d - d.transpose(1, 0, 2, 3)
    dFae = -np.einsum("mnef,mnaf->ae", oovv, Utd)
    dFmi = np.einsum("mnef,inef->mi", oovv, Utd)
    dWJ = mm(K, Ud) / 2 - mm(Kx, Td) / 2
    dWE = mm(Kx, Txd) / 2 + Dm(dFae, dFmi) / 2
    Tt, Txt, Ut = T0 + Td, Tx0 + Txd, U0 + Ud
    dH = mm(Td, WE0) + mm(Tt, dWE) + mm(Ud, WJ0) + mm(Ut, dWJ)
    dY = mm(Txd, WE0) + mm(Txt, dWE)
    dS = dH + dY.transpose(0, 3, 2, 1) + dH.transpose(2, 3, 0, 1) + dY.transpose(2, 1, 0, 3)
    W0 = oooo + np.einsum("mnef,ijef->mnij", oovv, t0)
    dW = np.einsum("mnef,ijef->mnij", oovv, d)
    hh = np.einsum("mnab,mnij->ijab", t, dW) + np.einsum("mnab,mnij->ijab", d, W0)
    return dS.transpose(0, 2, 1, 3) + np.einsum("ijef,abef->ijab", d, vvvv) + hh


@pytest.mark.parametrize("n,o,naux,seed", [(9, 3, 14, 5), (8, 1, 10, 6)])
def test_defect_correction_delta_residual(n, o, naux, seed):
    """R(t0) + [R(t) - R(t0)] as the defect-correction stage forms the bracket equals R(t)
    (RCCD.jl:242-256 restated) to 1e-12, for a t far from t0; dropping either half of a
    product's difference fails it."""
    s, I = _system(n, o, naux, seed)
    t0 = _amplitudes(s.o, s.v, seed + 400)
    t = t0 + _amplitudes(s.o, s.v, seed + 500, scale=0.03)
    R0 = O.rccd_residual_tensor(t0, I)
    ref = O.rccd_residual_tensor(t, I)
    dR = _device_residual_delta(t, t0, I)
    assert np.max(np.abs(R0 + dR - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
    # the linearised difference (t0 in place of t in the second halves) misses the d^2 terms
    lin = _device_residual_delta(t0 + 0.0 * t, t0, I)
    assert np.max(np.abs(lin)) < 1e-15
    assert np.max(np.abs(R0 + dR - O.rccd_residual_tensor(t0, I))) > 1e-6


def test_fock_terms_are_carried_by_we():
    """Without the D/2 dressing of WE the Fock-like terms are missing: the comparison has
    the power to see them."""
    s, I = _system(9, 3, 14, 5)
    t = _amplitudes(s.o, s.v, 105)
    ref = O.rccd_residual_tensor(t, I)
    assert np.max(np.abs(_device_residual(t, I, fock_in_we=False) - ref)) > 1e-4 * np.max(np.abs(ref))


# ---------------------------------------------------------------- RCCSD

def _device_rccsd(t1, t2, B, Co, Cv):
    """The device's RCCSD residual sequence (mRCCSD equations), restated with chemist blocks
    g_xxxx[p,q,r,s] = (pq|rs): tau = t2 + t1 t1, tau' = t2 + t1 t1/2; Fae / Fmi from tau' plus
    their T1 terms, Fme; the T1 residual; Fb = Fae - t1.Fme/2, Fj = Fmi + t1.Fme/2; both ladders
    on tau (W' with its T1 terms); Wabef's T1 part -sum_m t1[mb] Z[ij,ma] + (ia<->jb); the ring
    with U2 = U - 2 tt^x, Q = Tx + 2 tt^x and the T1 terms of WJ / WE, the Fock-like terms in
    WE; the T1^2 and T1-linear terms; S summed as the update kernel does."""
    g = lambda a, b, c, d: O.mo_from_factors(B, a, b, c, d)
    ovov, oovv, ovvv, ooov = g(Co, Cv, Co, Cv), g(Co, Co, Cv, Cv), g(Co, Cv, Cv, Cv), g(Co, Co, Co, Cv)
    oooo_p = g(Co, Co, Co, Co).transpose(0, 2, 1, 3)  # <mn|ij>
    vvvv_p = g(Cv, Cv, Cv, Cv).transpose(0, 2, 1, 3)  # <ab|ef>
    o, v = t1.shape
    es = np.einsum
    mm = lambda A, Bm: es("iakc,kcjb->iajb", A, Bm)
    Kpp = ovov.transpose(0, 2, 1, 3)  # <ij|ab> = (ia|jb), [i,j,a,b]
    Kov = ovov                        # (ia|jb), [i,a,j,b]
    Kx = ovov.transpose(0, 3, 2, 1)   # (ib|ja), [i,a,j,b]
    Jov = oovv.transpose(0, 2, 1, 3)  # (ij|ab), [i,a,j,b]
    tt = es("ia,jb->ijab", t1, t1)
    tau, taup = t2 + tt, t2 + 0.5 * tt
    ttx = es("ib,ja->iajb", t1, t1)   # tt^x[ia,jb] = t1[ib] t1[ja]
    Upp = 2 * taup - taup.transpose(1, 0, 2, 3)
    Tov = t2.transpose(0, 2, 1, 3)
    Ttov = es("ijba->iajb", t2)
    Uov = 2 * Tov - Ttov
    Fae = -es("mnef,mnaf->ae", Kpp, Upp)
    Fmi = es("mnef,inef->mi", Kpp, Upp)
    Fae += 2 * es("mf,mfae->ae", t1, ovvv) - es("mf,maef->ae", t1, ovvv)
    Fmi += 2 * es("ne,mine->mi", t1, ooov) - es("ne,nime->mi", t1, ooov)
    Fme = 2 * es("nf,menf->me", t1, Kov) - es("nf,nemf->me", t1, Kov)
    n1 = (es("ie,ae->ia", t1, Fae) - es("ma,mi->ia", t1, Fmi) + es("me,iame->ia", Fme, Uov)
          + 2 * es("me,iame->ia", t1, Kov) - es("me,maie->ia", t1, Jov)
          - 2 * es("mane,mine->ia", Tov, ooov) + es("mane,nime->ia", Tov, ooov)
          + 2 * es("iemf,mfae->ia", Tov, ovvv) - es("iemf,meaf->ia", Tov, ovvv))
    Fae = Fae - 0.5 * es("mb,me->be", t1, Fme)
    Fmi = Fmi + 0.5 * es("je,me->mj", t1, Fme)
    # ladders on tau
    Wp = oooo_p + es("je,mine->mnij", t1, ooov) + es("ie,njme->mnij", t1, ooov)
    Wp = Wp + es("mnef,ijef->mnij", Kpp, tau)
    L = es("ijef,abef->ijab", tau, vvvv_p) + es("mnab,mnij->ijab", tau, Wp)
    Z = es("ijef,mfae->ijma", tau, ovvv)
    Q = -es("mb,ijma->ijab", t1, Z)
    X2 = Q + Q.transpose(1, 0, 3, 2)
    # ring
    Qr = Ttov + 2 * ttx
    U2 = Uov - 2 * ttx
    WJ = Kov + 0.5 * mm(Kov, U2) - 0.5 * mm(Kx, Tov)
    WJ = WJ + es("jf,mebf->mejb", t1, ovvv) - es("nb,njme->mejb", t1, ooov)
    WE = -Jov + 0.5 * mm(Kx, Qr)
    WE = WE - es("jf,mfbe->mejb", t1, ovvv) + es("nb,mjne->mejb", t1, ooov)
    D = np.zeros((o, v, o, v))
    for k in range(o):
        D[k, :, k, :] += Fae.T
    for b in range(v):
        D[:, b, :, b] -= Fmi
    WE = WE + 0.5 * D
    H = mm(Tov, WE) + mm(Uov, WJ)
    Y = mm(Ttov, WE)
    Qa = es("ie,mejb->imjb", t1, Kov)
    H = H - es("ma,imjb->iajb", t1, Qa)
    Qb = es("ie,maje->imaj", t1, Jov)
    H = H - es("mb,imaj->iajb", t1, Qb)
    H = H + es("ie,jbae->iajb", t1, ovvv) - es("ma,mijb->iajb", t1, ooov)
    S = H + Y.transpose(0, 3, 2, 1) + H.transpose(2, 3, 0, 1) + Y.transpose(2, 1, 0, 3)
    n2 = Kpp + S.transpose(0, 2, 1, 3) + L + X2
    return n1, n2


@pytest.mark.parametrize("n,o,naux,seed", [(9, 3, 14, 8), (8, 2, 12, 9)])
def test_device_rccsd_sequence_matches_oracle(n, o, naux, seed):
    s = synthetic.make_system(n, o, n - o, naux, seed, name="layout")
    I = O.rccsd_integrals_from_factors(s.B, s.Co, s.Cv)
    rng = np.random.default_rng(seed + 200)
    t1 = rng.normal(size=(s.o, s.v)) * 0.05
    t2 = _amplitudes(s.o, s.v, seed + 300)
    r1, r2 = O.rccsd_update(t1, t2, I)
    d1, d2 = _device_rccsd(t1, t2, s.B, s.Co, s.Cv)
    assert np.max(np.abs(d1 - r1)) < 1e-12 * max(1.0, np.max(np.abs(r1)))
    assert np.max(np.abs(d2 - r2)) < 1e-12 * max(1.0, np.max(np.abs(r2)))


def _gemm_F(buf_a, lda, off_a, ta, buf_b, ldb, off_b, tb, M, N, K, ldc):
    """C (M x N, leading dimension ldc) = op(A) op(B) on flat column-major buffers, as the
    device's gemm reads them."""
    def mat(buf, ld, off, rows, cols):
        return np.lib.stride_tricks.as_strided(buf[off:], shape=(rows, cols), strides=(8, 8 * ld))
    A = mat(buf_a, lda, off_a, K, M).T if ta else mat(buf_a, lda, off_a, M, K)
    B = mat(buf_b, ldb, off_b, N, K).T if tb else mat(buf_b, ldb, off_b, K, N)
    return A @ B


def _factor_ovvv_terms(t1, t2, B, Co, Cv, mrccsd):
    """The device's (ia|bc) terms through the factors (ccsd.hip with AO factors), each GEMM
    restated on the flat column-major buffers with the device's leading dimensions, batch
    strides and interleaved outputs: returns the Fae, n1, Lz, WJ, WE and H contributions."""
    o, v = t1.shape
    Lov = O.mo_ov_factors(B, Co, Cv)                          # [m,e,P]
    naux = Lov.shape[2]
    Lvv = np.einsum("ma,mnP,nb->abP", Cv, B, Cv)              # [a,e,P]
    f = lambda x: np.ascontiguousarray(np.asarray(x).ravel(order="F"))  # noqa: E731
    lov, lvv, T1 = f(Lov), f(Lvv), f(t1)
    lovT, lovp = f(Lov.transpose(1, 0, 2)), f(Lov.transpose(0, 2, 1))
    ov, vv, oo = o * v, v * v, o * o
    # B1[m,P,e] (ldc = o naux, batch stride o) and Y2[m,e,P] (ldc = o, batch stride ov)
    B1, Y2, Y3 = np.zeros(ov * naux), np.zeros(ov * naux), np.zeros(oo * naux)
    for P in range(naux):
        c = _gemm_F(T1, o, 0, False, lvv, v, vv * P, True, o, v, v, 0)
        for e in range(v):
            B1[o * P + o * naux * e: o * P + o * naux * e + o] = c[:, e]
        Y2[ov * P: ov * (P + 1)] = f(c)
        Y3[oo * P: oo * (P + 1)] = f(_gemm_F(lov, o, ov * P, False, T1, o, 0, True, o, o, v, 0))
    g = _gemm_F(lov, ov, 0, True, T1, ov, 0, False, naux, 1, ov, 0)
    Fae = 2 * _gemm_F(lvv, vv, 0, False, f(g), naux, 0, False, vv, 1, naux, 0).reshape(v, v, order="F")
    A_, B_ = (lovp, B1) if mrccsd else (B1, lovp)
    Fae -= _gemm_F(A_, o * naux, 0, True, B_, o * naux, 0, False, v, v, o * naux, 0)
    Tov = f(t2.transpose(0, 2, 1, 3))
    W1 = f(_gemm_F(Tov, ov, 0, False, lov, ov, 0, False, ov, naux, ov, 0))
    n1 = 2 * _gemm_F(W1, o, 0, False, lvv, v, 0, True, o, v, v * naux, 0)
    W2 = np.zeros(ov * naux)
    for fi in range(v):                                       # batch over f, ldc = ov, stride o
        c = _gemm_F(Tov, o, oo * v * fi, False, lovT, ov, 0, False, o, naux, ov, 0)
        for P in range(naux):
            W2[o * fi + ov * P: o * fi + ov * P + o] = c[:, P]
    n1 -= _gemm_F(W2, o, 0, False, lvv, v, 0, True, o, v, v * naux, 0)
    # Z[i,j,a,m] over slabs of m, then Lz[(ija),b] = -Z[(ija),m] t1[m,b]
    tau = f(t2 + np.einsum("ia,jb->ijab", t1, t1))
    Z = np.zeros(oo * v * o)
    ms = 2
    for m0 in range(0, o, ms):
        cnt = min(ms, o - m0)
        Zs = _gemm_F(lovT, ov, v * m0, False, lvv, vv, 0, True, v * cnt, vv, naux, 0)
        Zsp = f(Zs.reshape(v, cnt, v, v, order="F").transpose(3, 0, 2, 1))
        Z[oo * v * m0: oo * v * (m0 + cnt)] = f(_gemm_F(tau, oo, 0, False, Zsp, vv, 0, False, oo, v * cnt, vv, 0))
    Lz = -_gemm_F(Z, oo * v, 0, False, T1, o, 0, False, oo * v, v, o, 0).reshape(o, o, v, v, order="F")
    WJ = _gemm_F(lov, ov, 0, False, Y2, ov, 0, True, ov, ov, naux, 0).reshape(o, v, o, v, order="F")
    Qm = _gemm_F(Y3, oo, 0, False, lvv, vv, 0, True, oo, vv, naux, 0).reshape(o, o, v, v, order="F")
    WE = -Qm.transpose(0, 3, 1, 2)
    H = _gemm_F(Y2, ov, 0, False, lov, ov, 0, True, ov, ov, naux, 0).reshape(o, v, o, v, order="F")
    return Fae, n1, Lz, WJ, WE, H


@pytest.mark.parametrize("mrccsd", [True, False])
def test_factored_ovvv_terms_equal_stored_block(mrccsd):
    """The (ia|bc) terms of the factor route equal the stored-block contractions of the device
    sequence above (and so, through it, mRCCSD.jl / RCCSD.jl) to 1e-12, on factors without
    (mu,nu) symmetry; swapping the two Fae variants breaks the Fae check."""
    s = synthetic.make_system(9, 3, 6, 13, 31, name="fac")
    rng = np.random.default_rng(5)
    B = np.asfortranarray(s.B + 0.05 * rng.normal(size=s.B.shape))
    t1 = rng.normal(size=(s.o, s.v)) * 0.1
    t2 = _amplitudes(s.o, s.v, 77)
    ovvv = O.mo_from_factors(B, s.Co, s.Cv, s.Cv, s.Cv)
    es = np.einsum
    tau = t2 + es("ia,jb->ijab", t1, t1)
    Tov = t2.transpose(0, 2, 1, 3)
    ref_Fae = 2 * es("mf,mfae->ae", t1, ovvv) - es("mf,maef->ae" if mrccsd else "mf,meaf->ae", t1, ovvv)
    other = 2 * es("mf,mfae->ae", t1, ovvv) - es("mf,meaf->ae" if mrccsd else "mf,maef->ae", t1, ovvv)
    ref_n1 = 2 * es("iemf,mfae->ia", Tov, ovvv) - es("iemf,meaf->ia", Tov, ovvv)
    ref_Lz = -es("mb,ijma->ijab", t1, es("ijef,mfae->ijma", tau, ovvv))
    ref_WJ = es("jf,mebf->mejb", t1, ovvv)
    ref_WE = -es("jf,mfbe->mejb", t1, ovvv)
    ref_H = es("ie,jbae->iajb", t1, ovvv)
    got = _factor_ovvv_terms(t1, t2, B, s.Co, s.Cv, mrccsd)
    for name, g, r in zip(("Fae", "n1", "Lz", "WJ", "WE", "H"), got,
                          (ref_Fae, ref_n1, ref_Lz, ref_WJ, ref_WE, ref_H)):
        assert np.max(np.abs(g - r)) < 1e-12 * max(1.0, np.max(np.abs(r))), name
    assert np.max(np.abs(got[0] - other)) > 1e-4 * np.max(np.abs(other))


def test_factored_k_ring_product():
    """With K[ia,jb] = (ia|jb) = sum_P Lov[ia,P] Lov[jb,P], the ring product K U equals
    Lov (U Lov)^T for the symmetric U = 2T - Tx of pair-symmetric amplitudes (the device's
    two naux-wide products); it needs U symmetric, which a non-symmetric U breaks."""
    s = synthetic.make_system(10, 3, 7, 15, 41, name="klr")
    rng = np.random.default_rng(2)
    B = np.asfortranarray(s.B + 0.05 * rng.normal(size=s.B.shape))
    o, v = s.o, s.v
    Lov = O.mo_ov_factors(B, s.Co, s.Cv).reshape(o * v, -1, order="F")
    K = O.mo_from_factors(B, s.Co, s.Cv, s.Co, s.Cv).reshape(o * v, o * v, order="F")
    t = _amplitudes(o, v, 9)
    T = t.transpose(0, 2, 1, 3).reshape(o * v, o * v, order="F")
    Tx = np.einsum("ijba->iajb", t).reshape(o * v, o * v, order="F")
    U = 2 * T - Tx
    assert np.max(np.abs(U - U.T)) < 1e-14
    lr = Lov @ (U @ Lov).T
    assert np.max(np.abs(lr - K @ U)) < 1e-12 * np.max(np.abs(K @ U))
    Un = U + 0.1 * rng.normal(size=U.shape)
    assert np.max(np.abs(Lov @ (Un @ Lov).T - K @ Un)) > 1e-4


# ---------------------------------------------------------------- recomputed ladder

@pytest.mark.parametrize("nf", [1, 2, 4, 6])
def test_recomputed_ladder_slabs_equal_stored_ladder(nf):
    """The ladder with <ab|ef> rebuilt from the factors one slab of f at a time: per slab,
    X[e,a,b,f-f0] = sum_P Lt[ea,P] L[bf,P] = (ae|bf) (L[ae,P] the MO factors, rows a + v e;
    Lt the same with a and e exchanged), its rows of V+- (pairs e<=f / e<f with f in the slab)
    are ((ae|bf) +- (af|be))/2, and they are the contiguous K-range f0(f0+1)/2 ... of the
    packed product T+- V+-.  Summed over slabs this is the stored-V ladder.  (The factors of
    real orbitals are symmetric in a, e, so L and Lt hold the same numbers here: the test pins
    the slab offsets and the packing, not which of the two the device reads.)"""
    s = synthetic.make_system(9, 3, 6, 14, 10, name="layout")
    o, v, B, Cv = s.o, s.v, s.B, s.Cv
    t = _amplitudes(o, v, 410)
    vvvv = O.mo_from_factors(B, Cv, Cv, Cv, Cv).transpose(0, 2, 1, 3)  # <ab|ef>
    ole, olt, vle, vlt = _pairs(o, False), _pairs(o, True), _pairs(v, False), _pairs(v, True)
    Tp = np.array([[t[ij + ef] if ef[0] == ef[1] else t[ij + ef] + t[ij + ef[::-1]] for ef in vle] for ij in ole])
    Tm = np.array([[t[ij + ef] - t[ij + ef[::-1]] for ef in vlt] for ij in olt])
    Vp = np.array([[0.5 * (vvvv[ab + ef] + vvvv[ab + ef[::-1]]) for ab in vle] for ef in vle])
    Vm = np.array([[0.5 * (vvvv[ab + ef] - vvvv[ab + ef[::-1]]) for ab in vlt] for ef in vlt])
    L3 = np.einsum("ma,mnP,ne->aeP", Cv, B, Cv)
    naux = L3.shape[2]
    L = L3.reshape(v * v, naux, order="F")                      # row a + v e
    Lt = L3.transpose(1, 0, 2).reshape(v * v, naux, order="F")  # row e + v a holds L[a,e]
    Lp = np.zeros((len(ole), len(vle)))
    Lm = np.zeros((len(olt), len(vlt)))
    for f0 in range(0, v, nf):
        f1 = min(v, f0 + nf)
        X = (Lt @ L[v * f0:v * f1].T).reshape(v, v, v, f1 - f0, order="F")  # X[e,a,b,f-f0]
        r0p, r1p = f0 * (f0 + 1) // 2, f1 * (f1 + 1) // 2
        r0m, r1m = f0 * (f0 - 1) // 2, f1 * (f1 - 1) // 2
        Vps = np.array([[0.5 * (X[e, a, b, f - f0] + X[e, b, a, f - f0]) for (a, b) in vle]
                        for (e, f) in vle[r0p:r1p]])
        Lp += Tp[:, r0p:r1p] @ Vps
        if r1m > r0m:
            Vms = np.array([[0.5 * (X[e, a, b, f - f0] - X[e, b, a, f - f0]) for (a, b) in vlt]
                            for (e, f) in vlt[r0m:r1m]])
            Lm += Tm[:, r0m:r1m] @ Vms
    assert np.max(np.abs(Lp - Tp @ Vp)) < 1e-12 * np.max(np.abs(Tp @ Vp))
    assert np.max(np.abs(Lm - Tm @ Vm)) < 1e-12 * np.max(np.abs(Tm @ Vm))
