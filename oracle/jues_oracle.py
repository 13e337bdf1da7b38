"""CPU FP64 restatement of JuES's correlation hot path -- TEST INFRASTRUCTURE ONLY.

This module is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import it; the product path
(``libjues_hip.so`` and the ``jues`` host package) never routes through it.

Every function restates one reference function from its source text (paths
relative to the JuES.jl tree), index for index, with NumPy ``einsum`` /
``reshape(order="F")`` standing in for TensorOperations ``@tensor`` and
``BLAS.gemm!``.  Arrays are indexed exactly like the Julia arrays (0-based):
``oovv[i,j,a,b]`` here is ``oovv[i+1,j+1,a+1,b+1]`` there.

How each part is pinned (the checks live in ``tests/test_oracle.py``):

* ``tei_transform`` / ``get_eri``: brute-force 8-loop sum at n<=4, the exact
  3-index factor route ``(pq|rs) = sum_P (C^T B_P C)_pq (C^T B_P C)_rs``,
  identity-C and Frobenius-norm invariants, chem<->phys permutations.
* ``rmp2``: the literal 4-loop form of RMP2.jl, rank-1 closed form,
  homogeneity ``E(lambda B) = lambda^4 E(B)``.
* ``rccd_*``: two independent restatements (RCCD.jl ``@tensoropt`` form and
  mRCCD.jl GEMM form) agree to 1e-12; first Jacobi step from t=0 gives E_MP2;
  the converged residual vanishes; fixed point independent of guess and DIIS;
  small-coupling limit ``E_CCD(lambda g) -> lambda^2 E_MP2(g)``.
* ``rccsd_*``: mRCCSD.jl 6-block and RCCSD.jl 15-block restatements agree;
  with ooov = ovvv = 0 RCCSD reduces to RCCD exactly.
* ``rmp2_grad``: central finite differences.
* ``make_bmunu`` / ``df_rmp2`` / ``df_rccd_*`` (DF.jl, DF-RMP2.jl, DF-RCCD.jl):
  b . J^(1/2) round trip and the fitted (mu nu|lam sig) = pqP J^-1 pqP^T closed
  form; DF-MP2 equals ``rmp2`` on the fitted four-index integrals; one DF sweep
  equals ``rccd_step`` (RCCD.jl) on the same integrals to 1e-12, while the
  WmBeJ term of DF-RCCD.jl:250 as written does not; the MP2 guess's energy is
  the DF-MP2 energy.
* Reference known answers (water STO-3G) through ``oracle/sto3g.py``.
"""
from __future__ import annotations

import numpy as np

_es = lambda spec, *ops: np.einsum(spec, *ops, optimize=True)  # noqa: E731


def _pd(A, perm):
    """Julia ``permutedims(A, perm)`` (1-based perm)."""
    return np.transpose(A, [p - 1 for p in perm])


def _rs(A, *shape):
    """Julia column-major ``reshape``."""
    return np.reshape(A, shape, order="F")


def _tperm(A, lhs):
    """Julia ``@tensor X[lhs...] := A[1,2,3,4]`` (pure relabelling)."""
    inv = [0] * len(lhs)
    for k, l in enumerate(lhs):
        inv[k] = l - 1
    return np.transpose(A, inv)


# ---------------------------------------------------------------------------
# Backend/Transformation.jl:39-93 and Backend/IntegralTransformation.jl:38-101
# ---------------------------------------------------------------------------

def tei_transform(gao, C1, C2, C3, C4):
    """``out[i,a,j,b] = sum C1[mu,i] C2[nu,a] C3[lam,j] C4[sig,b] g[mu,nu,lam,sig]``.

    Quarter order sigma -> lambda -> nu -> mu as in Transformation.jl:69-91.
    """
    temp = _es("sb,mnls->mnlb", C4, gao)        # :71-73
    temp2 = _es("lj,mnlb->mnjb", C3, temp)      # :77-79
    temp = _es("na,mnjb->majb", C2, temp2)      # :83-85
    return _es("mi,majb->iajb", C1, temp)       # :89-91


def tei_transform_bruteforce(gao, C1, C2, C3, C4):
    """Eight nested loops; only for n <= 4."""
    n = gao.shape[0]
    d1, d2, d3, d4 = C1.shape[1], C2.shape[1], C3.shape[1], C4.shape[1]
    out = np.zeros((d1, d2, d3, d4))
    for p in range(d1):
        for q in range(d2):
            for r in range(d3):
                for s in range(d4):
                    acc = 0.0
                    for m in range(n):
                        for nu in range(n):
                            for l in range(n):
                                for sg in range(n):
                                    acc += (C1[m, p] * C2[nu, q] * C3[l, r] * C4[sg, s]
                                            * gao[m, nu, l, sg])
                    out[p, q, r, s] = acc
    return out


def mo_from_factors(B, C1, C2, C3, C4):
    """Exact 3-index route: ``(pq|rs) = sum_P (C1^T B_P C2)_pq (C3^T B_P C4)_rs``."""
    L = _es("mp,mnP,nq->pqP", C1, B, C2)
    R = _es("mp,mnP,nq->pqP", C3, B, C4)
    return _es("pqP,rsP->pqrs", L, R)


def get_eri(gao, Cao, Cav, eri_string, notation="phys", fcn=0, Cbo=None, Cbv=None):
    """IntegralTransformation.jl:38-101.  Upper-case letters take the alpha orbitals
    (``O`` -> Cao, ``V`` -> Cav), lower-case the beta ones (``o`` -> Cbo, ``v`` -> Cbv),
    :52-68; ``fcn`` drops the first fcn columns of either occupied block (:53,57).
    With ``Cbo``/``Cbv`` omitted the reference is restricted and beta = alpha."""
    if len(eri_string) != 4:
        raise ValueError(f"Invalid string given to get_eri: {eri_string}")
    Cbo = Cao if Cbo is None else Cbo
    Cbv = Cav if Cbv is None else Cbv
    s = eri_string
    if notation == "phys":
        s = s[0] + s[2] + s[1] + s[3]                       # :46-48
    blocks = {"O": Cao[:, fcn:], "o": Cbo[:, fcn:], "V": Cav, "v": Cbv}
    Cs = []
    for ch in s:                                            # :52-68
        if ch not in blocks:
            raise ValueError(f"Invalid string given to get_eri: {eri_string}")
        Cs.append(blocks[ch])
    V = tei_transform(gao, *Cs)                             # :94
    if notation == "phys":
        V = np.transpose(V, (0, 2, 1, 3))                   # :96-98
    return np.ascontiguousarray(V)


# ---------------------------------------------------------------------------
# MollerPlesset/RMP2.jl:11-45 and CoupledCluster/Denominators.jl:14-33
# ---------------------------------------------------------------------------

def form_Dijab(eps, o, v):
    """Denominators.jl:14-33: ``D[i,j,a,b] = e_i + e_j - e_{o+a} - e_{o+b}``."""
    eo, ev = eps[:o], eps[o:o + v]
    return (eo[:, None, None, None] + eo[None, :, None, None]
            - ev[None, None, :, None] - ev[None, None, None, :])


def form_Dia(eps, o, v):
    """mRCCSD.jl:64: ``D[i,a] = e_i - e_{o+a}``."""
    return eps[:o, None] - eps[None, o:o + v]


def rmp2(moeri, eps, o):
    """RMP2.jl:27-41 vectorised: ``sum K[ijab](2K[ijab] - K[ijba]) / D``."""
    v = moeri.shape[2]
    D = form_Dijab(eps, o, v)
    return float(np.sum(moeri * (2.0 * moeri - moeri.transpose(0, 1, 3, 2)) / D))


def rmp2_loops(moeri, eps, o):
    """RMP2.jl:27-41 literally (b, a, j, i loop order); small inputs only."""
    v = moeri.shape[2]
    dmp2 = 0.0
    for b in range(v):
        for a in range(v):
            for j in range(o):
                for i in range(o):
                    dmp2 += (moeri[i, j, a, b] * (2 * moeri[i, j, a, b] - moeri[i, j, b, a])
                             / (eps[i] + eps[j] - eps[o + a] - eps[o + b]))
    return dmp2


def rmp2_grad(gao, C, eps, o):
    """Analytic gradient of the RMP2 energy w.r.t. C (n x n) and g (n^4).

    ``E = sum K_iajb (2K_iajb - K_ibja)/D_ijab`` with ``K = (ia|jb)``.  Every K
    entry independent: ``G_iajb = dE/dK_iajb = 2(2K_iajb - K_ibja)/D_ijab``;
    ``dE/dg = sum G C_mu,i C_nu,a C_lam,j C_sig,b``;
    ``dE/dC[:,occ]`` and ``dE/dC[:,vir]`` collect the four positions C enters.
    No symmetry of g is assumed.  Returns (E, dE/dC, dE/dg).
    """
    Co, Cv = C[:, :o], C[:, o:]
    v = Cv.shape[1]
    K = tei_transform(gao, Co, Cv, Co, Cv)                 # (ia|jb) as [i,a,j,b]
    D = form_Dijab(eps, o, v).transpose(0, 2, 1, 3)        # D as [i,a,j,b]
    Kx = K.transpose(0, 3, 2, 1)                           # K_ibja as [i,a,j,b]
    E = float(np.sum(K * (2 * K - Kx) / D))
    G = 2.0 * (2.0 * K - Kx) / D
    dg = _es("iajb,mi,na,lj,sb->mnls", G, Co, Cv, Co, Cv)
    # positions: C1 (occ, index mu), C2 (vir, nu), C3 (occ, lam), C4 (vir, sig)
    h1 = _es("iajb,na,lj,sb->inls", G, Cv, Co, Cv)
    dCo = _es("inls,mnls->mi", h1, gao)
    h3 = _es("iajb,mi,na,sb->mnjs", G, Co, Cv, Cv)
    dCo += _es("mnjs,mnls->lj", h3, gao)
    h2 = _es("iajb,mi,lj,sb->mals", G, Co, Co, Cv)
    dCv = _es("mals,mnls->na", h2, gao)
    h4 = _es("iajb,mi,na,lj->mnlb", G, Co, Cv, Co)
    dCv += _es("mnlb,mnls->sb", h4, gao)
    dC = np.concatenate([dCo, dCv], axis=1)
    return E, dC, dg


def mo_ov_factors(B, Co, Cv):
    """L[i,a,P] = sum_mu,nu Co[mu,i] B[mu,nu,P] Cv[nu,a]: (ia|jb) = sum_P L[ia,P] L[jb,P]."""
    n, _, naux = B.shape
    X = np.tensordot(Co, B, axes=([0], [0]))          # [i, nu, P]
    return np.einsum("inP,na->iaP", X, Cv, optimize=True)


def rmp2_factors(B, C, eps, o):
    """RMP2.jl:33-37 on (ia|jb) built from 3-index factors (no n^4 tensor)."""
    Co, Cv = C[:, :o], C[:, o:]
    v = Cv.shape[1]
    L = mo_ov_factors(B, Co, Cv)
    naux = L.shape[2]
    K = (L.reshape(o * v, naux, order="F") @ L.reshape(o * v, naux, order="F").T).reshape(o, v, o, v, order="F")
    D = form_Dijab(eps, o, v).transpose(0, 2, 1, 3)
    return float(np.sum(K * (2 * K - K.transpose(0, 3, 2, 1)) / D))


def rmp2_factors_slabs(B, C, eps, o):
    """RMP2.jl:33-37 from 3-index factors, one occupied index j at a time: for fixed j,
    K[i,a,j,b] = sum_P L[i,a,P] L[j,b,P] is an (ov) x v slab and the exchange K[i,b,j,a]
    lies in the same slab, so memory stays at o v^2 (the o=80, v=420 K is 9 GB)."""
    Co, Cv = C[:, :o], C[:, o:]
    v = Cv.shape[1]
    L = mo_ov_factors(B, Co, Cv)
    naux = L.shape[2]
    Lm = L.reshape(o * v, naux, order="F")
    eo, ev = eps[:o], eps[o:o + v]
    E = 0.0
    for j in range(o):
        Kj = (Lm @ L[j].T).reshape(o, v, v, order="F")          # [i,a,b] = (ia|jb)
        D = eo[:, None, None] + eo[j] - ev[None, :, None] - ev[None, None, :]
        E += float(np.sum(Kj * (2.0 * Kj - Kj.transpose(0, 2, 1)) / D))
    return E


def rmp2_grad_factors(B, C, eps, o):
    """dE/dC of the RMP2 energy through the factor route: with K = L L^T and
    G = dE/dK (as in ``rmp2_grad``), dE/dL = (G + G^T) L and L = Co^T B_P Cv gives
    dE/dCo[mu,i] = sum_aP dL[i,a,P] (B_P Cv)[mu,a], dE/dCv[nu,a] = sum_iP dL[i,a,P] (Co^T B_P)[i,nu].
    No symmetry of B in (mu,nu) is assumed.  Returns (E, dE/dC)."""
    Co, Cv = C[:, :o], C[:, o:]
    v = Cv.shape[1]
    L = mo_ov_factors(B, Co, Cv)
    naux = L.shape[2]
    Lm = L.reshape(o * v, naux, order="F")
    K = (Lm @ Lm.T).reshape(o, v, o, v, order="F")
    D = form_Dijab(eps, o, v).transpose(0, 2, 1, 3)
    Kx = K.transpose(0, 3, 2, 1)
    E = float(np.sum(K * (2 * K - Kx) / D))
    G = (2.0 * (2.0 * K - Kx) / D).reshape(o * v, o * v, order="F")
    dL = ((G + G.T) @ Lm).reshape(o, v, naux, order="F")
    Z = np.einsum("mnP,na->maP", B, Cv, optimize=True)    # (B_P Cv)[mu,a]
    W = np.tensordot(Co, B, axes=([0], [0]))             # (Co^T B_P)[i,nu]
    dCo = np.einsum("iaP,maP->mi", dL, Z, optimize=True)
    dCv = np.einsum("iaP,inP->na", dL, W, optimize=True)
    return E, np.concatenate([dCo, dCv], axis=1)


def rmp2_grad_eps(K, eps, o):
    """dE/deps of the RMP2 energy (RMP2.jl:33-37) from K = (ia|jb) stored as [i,a,j,b].

    With W = K (2K - K_ibja) / D^2 and D = e_i + e_j - e_{o+a} - e_{o+b}, each orbital
    energy enters D through two slots: dE/de_k = -(sum_{a,jb} W[k,a,jb] + sum_{ia,b} W[ia,k,b])
    for occupied k and dE/de_{o+c} = sum_{i,jb} W[i,c,jb] + sum_{ia,j} W[ia,j,c] for virtual c.
    No symmetry of K is assumed.  Returns an (o + v) vector, occupied first."""
    v = K.shape[1]
    D = form_Dijab(eps, o, v).transpose(0, 2, 1, 3)
    W = K * (2.0 * K - K.transpose(0, 3, 2, 1)) / D ** 2
    r = W.sum(axis=(2, 3)) + W.sum(axis=(0, 1))          # [i,a]: both slots of (ia)
    return np.concatenate([-r.sum(axis=1), r.sum(axis=0)])


def rmp2_energy_from_C(gao, C, eps, o):
    Co, Cv = C[:, :o], C[:, o:]
    K = tei_transform(gao, Co, Cv, Co, Cv)
    return rmp2(K.transpose(0, 2, 1, 3), eps, o)


# ---------------------------------------------------------------------------
# CoupledCluster/RCCD.jl (the @tensoropt form)
# ---------------------------------------------------------------------------

def make_rccd_integrals(gao, Cao, Cav):
    """RCCD.jl:85-97: physicist blocks oovv, ovov, ovvo, oooo, vvvv."""
    oovv = np.transpose(tei_transform(gao, Cao, Cav, Cao, Cav), (0, 2, 1, 3))
    vvvv = np.transpose(tei_transform(gao, Cav, Cav, Cav, Cav), (0, 2, 1, 3))
    ovvo = np.transpose(tei_transform(gao, Cao, Cav, Cav, Cao), (0, 2, 1, 3))
    ovov = np.transpose(tei_transform(gao, Cao, Cao, Cav, Cav), (0, 2, 1, 3))
    oooo = np.transpose(tei_transform(gao, Cao, Cao, Cao, Cao), (0, 2, 1, 3))
    return dict(oovv=oovv, ovov=ovov, ovvo=ovvo, oooo=oooo, vvvv=vvvv)


def rccd_integrals_from_factors(B, Cao, Cav):
    blk = lambda A, Bc, Cc, Dc: np.transpose(mo_from_factors(B, A, Bc, Cc, Dc), (0, 2, 1, 3))
    return dict(oovv=blk(Cao, Cav, Cao, Cav), ovov=blk(Cao, Cao, Cav, Cav),
                ovvo=blk(Cao, Cav, Cav, Cao), oooo=blk(Cao, Cao, Cao, Cao),
                vvvv=blk(Cav, Cav, Cav, Cav))


def ccd_energy(t, oovv):
    """RCCD.jl:99-117: ``E = sum oovv[ijab](2t[ijab] - t[jiab])``."""
    return float(np.sum(oovv * (2.0 * t - t.transpose(1, 0, 2, 3))))


def rccd_residual_tensor(t, I):
    """RCCD.jl:134-139 + form_T2 :242-256 before the division by D (:301)."""
    oovv, ovov, ovvo, oooo, vvvv = I["oovv"], I["ovov"], I["ovvo"], I["oooo"], I["vvvv"]
    Fae = -_es("mnef,mnaf->ae", oovv, 2 * t - t.transpose(1, 0, 2, 3))           # :178-180
    Fmi = _es("mnef,inef->mi", oovv, 2 * t - t.transpose(0, 1, 3, 2))            # :215-217
    Wmnij = oooo + _es("ijef,mnef->mnij", t, oovv) / 2                           # :320-322
    Wabef = _es("mnab,mnef->abef", t, oovv) / 2 + vvvv                           # :358-360
    WmBeJ = (ovvo + _es("mnef,njfb->mbej", oovv, 2 * t - t.transpose(1, 0, 2, 3)) / 2
             - _es("nmef,njfb->mbej", oovv, t) / 2)                              # :399-402
    WmBEj = -ovov.transpose(0, 1, 3, 2) + _es("jnfb,nmef->mbej", t, oovv) / 2    # :444-446
    R = (oovv + _es("ijae,be->ijab", t, Fae) + _es("jibe,ae->ijab", t, Fae)
         - _es("imab,mj->ijab", t, Fmi) - _es("mjab,mi->ijab", t, Fmi)
         + _es("mnab,mnij->ijab", t, Wmnij)
         + _es("ijef,abef->ijab", t, Wabef)
         + 2 * _es("imae,mbej->ijab", t, WmBeJ)
         - _es("miae,mbej->ijab", t, WmBeJ)
         + _es("imae,mbej->ijab", t, WmBEj)
         + _es("mibe,maej->ijab", t, WmBEj)
         + _es("mjae,mbei->ijab", t, WmBEj)
         + 2 * _es("jmbe,maei->ijab", t, WmBeJ)
         - _es("mjbe,maei->ijab", t, WmBeJ)
         + _es("jmbe,maei->ijab", t, WmBEj))
    return R


def rccd_step(t, I, D):
    """One Jacobi update ``t_new = R(t) ./ D`` (RCCD.jl:301)."""
    return rccd_residual_tensor(t, I) / D


# ---------------------------------------------------------------------------
# CoupledCluster/mRCCD.jl: the explicit GEMM formulation (second restatement)
# ---------------------------------------------------------------------------

def _gemm(tA, tB, alpha, A, B, beta, C):
    opA = A.T if tA == "T" else A
    opB = B.T if tB == "T" else B
    return alpha * (opA @ opB) + beta * C


def mrccd_residual_tensor(t, I):
    """mRCCD.jl:159-164 + 265-488 (the ``./ Dijab`` of :391 left out)."""
    oovv, ovov, ovvo, oooo, vvvv = I["oovv"], I["ovov"], I["ovvo"], I["oooo"], I["vvvv"]
    o, v = t.shape[0], t.shape[2]
    # form_Fae! :265-275
    X = 2 * _rs(_pd(t, (3, 1, 2, 4)), v, o * o * v) - _rs(_pd(t, (3, 2, 1, 4)), v, o * o * v)
    Fae = _gemm("N", "N", -1.0, X, _rs(_pd(oovv, (1, 2, 4, 3)), o * o * v, v), 0.0, 0)
    # form_Fmi! :285-299
    X = 2 * _rs(_pd(t, (2, 3, 4, 1)), o * v * v, o) - _rs(_pd(t, (2, 4, 3, 1)), o * v * v, o)
    Fmi = _gemm("N", "N", 1.0, _rs(oovv, o, o * v * v), X, 0.0, 0)
    # form_Wmnij! :403-415
    Wmnij = _rs(_gemm("N", "T", 0.5, _rs(oovv, o * o, v * v), _rs(t, o * o, v * v), 1.0,
                      _rs(oooo, o * o, o * o)), o, o, o, o)
    # form_Wabef! :424-436
    Wabef = _rs(_gemm("T", "N", 0.5, _rs(t, o * o, v * v), _rs(oovv, o * o, v * v), 1.0,
                      _rs(vvvv, v * v, v * v)), v, v, v, v)
    # form_WmBeJ! :446-464
    W = _rs(_pd(ovvo, (1, 3, 4, 2)), o * v, o * v)
    t1 = 2 * _rs(_pd(t, (1, 3, 2, 4)), o * v, o * v) - _rs(_pd(t, (2, 3, 1, 4)), o * v, o * v)
    W = _gemm("N", "N", 0.5, _rs(_pd(oovv, (1, 3, 2, 4)), o * v, o * v), t1, 1.0, W)
    W = _gemm("T", "N", -0.5, _rs(_pd(oovv, (1, 4, 2, 3)), o * v, o * v),
              _rs(_pd(t, (1, 3, 2, 4)), o * v, o * v), 1.0, W)
    WmBeJ = _pd(_rs(W, o, v, o, v), (1, 4, 2, 3))
    # form_WmBEj! :475-488
    W = -_pd(ovov, (1, 2, 4, 3))
    W = _rs(_pd(W, (1, 3, 4, 2)), o * v, o * v)
    W = _gemm("N", "N", 0.5, _rs(_pd(oovv, (2, 3, 1, 4)), o * v, o * v),
              _rs(_pd(t, (2, 3, 1, 4)), o * v, o * v), 1.0, W)
    WmBEj = _pd(_rs(W, o, v, o, v), (1, 4, 2, 3))

    # form_T2 :301-394
    d = oovv.copy()
    d = _rs(_gemm("N", "T", 1.0, _rs(t, o * o * v, v), Fae, 1.0, _rs(d, o * o * v, v)), o, o, v, v)
    X = _tperm(d, (2, 1, 4, 3))
    X = _gemm("N", "T", 1.0, _rs(t, o * o * v, v), Fae, 1.0, _rs(X, o * o * v, v))
    d = _pd(_rs(X, o, o, v, v), (2, 1, 4, 3))
    X = _rs(_tperm(d, (1, 3, 4, 2)), o * v * v, o)
    X = _gemm("T", "N", -1.0, _rs(_pd(t, (2, 1, 3, 4)), o, o * v * v), Fmi, 1.0, X)
    d = _pd(_rs(X.T, o, o, v, v), (2, 1, 3, 4))
    X = _rs(_tperm(d, (2, 3, 4, 1)), o * v * v, o)
    X = _gemm("T", "N", -1.0, _rs(t, o, o * v * v), Fmi, 1.0, X)
    d = _rs(X.T, o, o, v, v)
    X = _rs(d, o * o, v * v)
    X = _gemm("T", "N", 1.0, _rs(Wmnij, o * o, o * o), _rs(t, o * o, v * v), 1.0, X)
    X = _gemm("N", "T", 1.0, _rs(t, o * o, v * v), _rs(Wabef, v * v, v * v), 1.0, X)
    d = _rs(X, o, o, v, v)

    X = _rs(_tperm(d, (1, 3, 4, 2)), o * v, v * o)
    scr2 = _rs(_tperm(t, (1, 3, 2, 4)), o * v, o * v)
    scr3 = 2 * _rs(_tperm(WmBeJ, (1, 3, 2, 4)), o * v, v * o)
    it = _rs(_tperm(WmBEj, (1, 3, 2, 4)), o * v, v * o)
    scr3 = scr3 + it
    X = _gemm("N", "N", 1.0, scr2, scr3, 1.0, X)
    scr3 = (scr3 - it) * 0.5
    scr2 = _rs(_pd(t, (1, 4, 2, 3)), o * v, o * v)
    X = _gemm("T", "N", -1.0, scr2, scr3, 1.0, X)
    d = _pd(_rs(X, o, v, v, o), (1, 4, 2, 3))

    X = _rs(_pd(d, (1, 4, 3, 2)), o * v, v * o)
    scr2 = _rs(_pd(t, (1, 4, 2, 3)), o * v, o * v)
    scr4 = _rs(_pd(WmBEj, (1, 3, 2, 4)), o * v, o * v)
    tmp = _gemm("T", "N", 1.0, scr2, scr4, 0.0, 0)
    X = X + tmp
    d = _pd(_rs(X, o, v, v, o), (1, 4, 3, 2))
    scr1 = _rs(_pd(d, (2, 3, 4, 1)), o * v, v * o) + tmp
    d = _pd(_rs(scr1, o, v, v, o), (4, 1, 2, 3))
    scr1 = _rs(_pd(d, (2, 4, 3, 1)), o * v, v * o)
    scr2 = _rs(_pd(t, (1, 3, 2, 4)), o * v, o * v)
    scr1 = _gemm("N", "N", 1.0, scr2, scr4, 1.0, scr1)
    scr2 = 2 * scr2 - _rs(_pd(t, (2, 3, 1, 4)), o * v, o * v)
    scr1 = _gemm("N", "N", 1.0, scr2, scr3, 1.0, scr1)
    return _pd(_rs(scr1, o, v, v, o), (4, 1, 3, 2))


# ---------------------------------------------------------------------------
# DIIS as in mRCCD.jl:64-67, 171-208 (kept in FP64 here)
# ---------------------------------------------------------------------------

class DIIS:
    """Pulay extrapolation with mRCCD.jl's bookkeeping.

    ``max_vecs`` is the number of (amplitude, error) pairs kept; mRCCD.jl's
    ``max_diis = 6`` (a list that also holds the starting guess) keeps 5.
    """

    def __init__(self, max_vecs=5):
        self.max_vecs = max_vecs
        self.vals, self.errs = [], []

    def push(self, t_new, err):
        self.vals.append(np.array(t_new, copy=True))
        self.errs.append(np.array(err, copy=True))
        if len(self.vals) > self.max_vecs:                  # :180-183
            del self.vals[0]
            del self.errs[0]

    def extrapolate(self):
        m = len(self.errs)
        B = -np.ones((m + 1, m + 1))                        # :185-186
        B[m, m] = 0.0
        for a in range(m):
            for b in range(m):
                B[a, b] = np.vdot(self.errs[a].ravel(), self.errs[b].ravel())   # :188-192
        B[:m, :m] /= np.max(np.abs(B[:m, :m]))             # :194
        rhs = np.zeros(m + 1)
        rhs[m] = -1.0                                       # :195-196
        c = np.linalg.solve(B, rhs)                         # :198 (inv(B)*r)
        out = np.zeros_like(self.vals[0])
        for k in range(m):                                  # :199-202
            out += c[k] * self.vals[k]
        return out, c[:m]


def solve_rccd(I, eps, o, *, guess="zero", diis=True, max_vecs=5, maxit=100,
               e_conv=1e-13, r_conv=1e-12, residual=rccd_residual_tensor, history=None):
    """Fixed point of ``t = R(t)/D``; Jacobi (RCCD.jl:55-74) or DIIS (mRCCD.jl:68-109).

    ``guess``: "zero" (mRCCD.jl:245), "ovov" ((ij|ab)/D, RCCD.jl:145-160) or
    "mp2" (oovv/D).  Converged when |dE| < e_conv and ||t_new - t||_2 < r_conv.
    """
    v = eps.shape[0] - o
    D = form_Dijab(eps, o, v)
    oovv = I["oovv"]
    if guess == "zero":
        t = np.zeros_like(oovv)
    elif guess == "ovov":
        t = I["ovov"].transpose(0, 2, 1, 3) / D
    elif guess == "mp2":
        t = oovv / D
    else:
        raise ValueError(guess)
    d = DIIS(max_vecs) if diis else None
    E_old = ccd_energy(t, oovv)
    for it in range(maxit):
        t_new = residual(t, I) / D
        err = t_new - t
        E_new = ccd_energy(t_new, oovv)
        rn = float(np.linalg.norm(err))
        if history is not None:
            history.append((E_new, rn))
        if d is not None:
            d.push(t_new, err)
            t_next, _ = d.extrapolate()
        else:
            t_next = t_new
        dE = E_new - E_old
        E_old = E_new
        t = t_next
        if abs(dE) < e_conv and rn < r_conv:
            break
    return ccd_energy(t, oovv), t, it + 1


# ---------------------------------------------------------------------------
# CoupledCluster/mRCCSD.jl (6 physicist blocks)
# ---------------------------------------------------------------------------

def make_rccsd_integrals(gao, Cao, Cav):
    """mRCCSD.jl:102-110 via get_eri (physicist)."""
    g = lambda s: get_eri(gao, Cao, Cav, s)
    return dict(oooo=g("oooo"), ooov=g("ooov"), oovv=g("oovv"), ovov=g("ovov"),
                ovvv=g("ovvv"), vvvv=g("vvvv"))


def rccsd_integrals_from_factors(B, Cao, Cav):
    C = {"o": Cao, "v": Cav}
    def phys(s):
        chem = s[0] + s[2] + s[1] + s[3]
        X = mo_from_factors(B, *[C[ch] for ch in chem])
        return np.ascontiguousarray(np.transpose(X, (0, 2, 1, 3)))
    return {s: phys(s) for s in ("oooo", "ooov", "oovv", "ovov", "ovvv", "vvvv")}


def ccsd_energy(oovv, t1, t2):
    """mRCCSD.jl:111-117 with tiatia = t1 (x) t1."""
    tt = _es("ma,nf->mnaf", t1, t1)
    return float(np.sum(oovv * (2 * t2 + 2 * tt - t2.transpose(1, 0, 2, 3) - tt.transpose(1, 0, 2, 3))))


def rccsd_update(t1, t2, I):
    """mRCCSD.jl cciter :118-135 (intermediates :137-201, T1 :202-212, T2 :213-241)."""
    oooo, ooov, oovv, ovov, ovvv, vvvv = (I["oooo"], I["ooov"], I["oovv"], I["ovov"],
                                          I["ovvv"], I["vvvv"])
    o, v = t1.shape
    tt = _es("ma,nf->mnaf", t1, t1)
    # form_Fae! :137-149 (GEMM form restated as contractions)
    _ovvv = (2 * _rs(_pd(ovvv, (1, 3, 2, 4)), o * v, v * v)
             - _rs(_pd(ovvv, (1, 4, 3, 2)), o * v, v * v))
    Fae = _rs(_rs(t1, 1, o * v) @ _ovvv, v, v)
    _t = _rs(_pd(t2, (3, 1, 2, 4)), v, o * o * v) + 0.5 * _rs(_pd(tt, (3, 1, 2, 4)), v, o * o * v)
    _o = (2 * _rs(_pd(oovv, (1, 2, 4, 3)), o * o * v, v)
          - _rs(_pd(oovv, (2, 1, 4, 3)), o * o * v, v))
    Fae = Fae - _t @ _o
    # form_Fmi! :150-162
    _ooov = (2 * _rs(_pd(ooov, (2, 4, 1, 3)), o * v, o * o)
             - _rs(_pd(ooov, (1, 4, 2, 3)), o * v, o * o))
    Fmi = _rs(_rs(t1, 1, o * v) @ _ooov, o, o)
    _o = 2 * _rs(oovv, o, v * v * o) - _rs(_pd(oovv, (1, 2, 4, 3)), o, v * v * o)
    _t = _rs(_pd(t2, (2, 3, 4, 1)), v * v * o, o) + 0.5 * _rs(_pd(tt, (2, 3, 4, 1)), v * v * o, o)
    Fmi = Fmi + _o @ _t
    # form_Fme! :163-168
    Fme = _es("nf,mnef->me", t1, 2 * oovv - oovv.transpose(1, 0, 2, 3))
    # form_Wmnij! :169-176
    Wmnij = (oooo + _es("je,mnie->mnij", t1, ooov) + _es("ie,nmje->mnij", t1, ooov)
             + 0.5 * _es("ijef,mnef->mnij", t2 + tt, oovv))
    # form_Wabef! :177-184
    Wabef = (vvvv - _es("mb,mafe->abef", t1, ovvv) - _es("ma,mbef->abef", t1, ovvv)
             + 0.5 * _es("mnab,mnef->abef", t2 + tt, oovv))
    # form_WmBeJ! :185-193
    WmBeJ = (oovv.transpose(0, 3, 2, 1) + _es("jf,mbef->mbej", t1, ovvv)
             - _es("nb,nmje->mbej", t1, ooov)
             - 0.5 * _es("jnfb,mnef->mbej", t2 + 2 * tt, oovv)
             + 0.5 * _es("njfb,mnef->mbej", t2, 2 * oovv - oovv.transpose(1, 0, 2, 3)))
    # form_WmBEj! :194-201
    WmBEj = (-ovov.transpose(0, 1, 3, 2) - _es("jf,mbfe->mbej", t1, ovvv)
             + _es("nb,mnje->mbej", t1, ooov)
             + _es("jnfb,nmef->mbej", 0.5 * t2 + tt, oovv))
    # form_T1 :202-212
    n1 = (_es("ie,ae->ia", t1, Fae) - _es("ma,mi->ia", t1, Fmi)
          + _es("me,imae->ia", Fme, 2 * t2 - t2.transpose(1, 0, 2, 3))
          + _es("me,imae->ia", t1, 2 * oovv) - _es("me,maie->ia", t1, ovov)
          - _es("mnae,mnie->ia", t2, 2 * ooov - ooov.transpose(1, 0, 2, 3))
          + _es("imef,mafe->ia", t2, 2 * ovvv) - _es("imef,maef->ia", t2, ovvv))
    # form_T2 :213-241
    Fb = Fae - 0.5 * _es("mb,me->be", t1, Fme)
    Fj = Fmi + 0.5 * _es("je,me->mj", t1, Fme)
    n2 = (oovv
          + _es("ijae,be->ijab", t2, Fb) + _es("ijeb,ae->ijab", t2, Fb)
          - _es("imab,mj->ijab", t2, Fj) - _es("mjab,mi->ijab", t2, Fj)
          + _es("mnab,mnij->ijab", t2 + tt, Wmnij)
          + _es("ijef,abef->ijab", t2 + tt, Wabef)
          + _es("imae,mbej->ijab", t2 - t2.transpose(1, 0, 2, 3), WmBeJ)
          - _es("imea,mjeb->ijab", tt, oovv)
          + _es("imae,mbej->ijab", t2, WmBeJ + WmBEj)
          + _es("mibe,maej->ijab", t2, WmBEj) - _es("imeb,maje->ijab", tt, ovov)
          + _es("mjae,mbei->ijab", t2, WmBEj) - _es("jmea,mbie->ijab", tt, ovov)
          + _es("jmbe,maei->ijab", t2 - t2.transpose(1, 0, 2, 3), WmBeJ)
          - _es("jmeb,miea->ijab", tt, oovv)
          + _es("jmbe,maei->ijab", t2, WmBeJ + WmBEj)
          + _es("ei,jabe->ijab", t1.T, ovvv) + _es("ej,ibae->ijab", t1.T, ovvv)
          - _es("ma,mjib->ijab", t1, ooov) - _es("mb,jima->ijab", t1, ooov))
    return n1, n2


def rccsd_step(t1, t2, I, Dia, Dijab):
    n1, n2 = rccsd_update(t1, t2, I)
    return n1 / Dia, n2 / Dijab


# ---------------------------------------------------------------------------
# CoupledCluster/RCCSD.jl (15 blocks) -- second RCCSD restatement
# ---------------------------------------------------------------------------

def make_rccsd15_integrals(tei, Cao, Cav):
    """RCCSD.jl:117-142: 15 blocks, each from its own transform and permutation.

    ``tei(C1, C2, C3, C4)`` returns the chemist tensor (e.g. ``tei_transform``
    on dense AO integrals or ``mo_from_factors`` on factors).
    """
    o, v = Cao, Cav
    P = lambda X: np.transpose(X, (0, 2, 1, 3))            # permutedims(.,[1,3,2,4])
    b = dict(
        vvvv=P(tei(v, v, v, v)), ovvv=P(tei(o, v, v, v)), vovv=P(tei(v, v, o, v)),
        vvov=P(tei(v, o, v, v)), vvvo=P(tei(v, v, v, o)), oovv=P(tei(o, v, o, v)),
        ovvo=P(tei(o, v, v, o)), vovo=P(tei(v, v, o, o)), ovov=P(tei(o, o, v, v)),
        voov=P(tei(v, o, o, v)), ooov=P(tei(o, o, o, v)), oovo=P(tei(o, v, o, o)),
        ovoo=P(tei(o, o, v, o)), vooo=P(tei(v, o, o, o)), oooo=P(tei(o, o, o, o)))
    b["vvov"] = _pd(b["vvov"], (4, 1, 2, 3))                # :137
    b["vvvo"] = _pd(b["vvvo"], (3, 1, 2, 4))                # :138
    b["vovv"] = _pd(b["vovv"], (2, 1, 3, 4))                # :139
    b["vooo"] = _pd(b["vooo"], (2, 1, 3, 4))                # :140
    return b


def rccsd15_update(t1, t2, b):
    """RCCSD.jl cciter (:150-172) with form_* (:187-289), argument wiring kept."""
    oovv, vovv, ovvv, ooov, oooo = b["oovv"], b["vovv"], b["ovvv"], b["ooov"], b["oooo"]
    oovo, vvvv, ovvo, voov, vovo = b["oovo"], b["vvvv"], b["ovvo"], b["voov"], b["vovo"]
    ovov, vvvo, vvov, ovoo, vooo = b["ovov"], b["vvvo"], b["vvov"], b["ovoo"], b["vooo"]
    tt = _es("ma,nf->mnaf", t1, t1)
    # form_Fae!(Fae, maef=ovvv, amef=vovv, mnef=oovv)
    maef, amef, mnef = ovvv, vovv, oovv
    Fae = (_es("mf,maef->ae", t1, 2 * amef - maef)
           - _es("mnaf,mnef->ae", t2 + 0.5 * tt, 2 * mnef - mnef.transpose(1, 0, 2, 3)))
    # form_Fmi!(Fmi, mnie=ooov, mnef=oovv)
    mnie = ooov
    Fmi = (_es("ne,mnie->mi", t1, 2 * mnie - mnie.transpose(1, 0, 2, 3))
           + _es("inef,mnef->mi", t2 + 0.5 * tt, 2 * mnef - mnef.transpose(0, 1, 3, 2)))
    Fme = _es("nf,mnef->me", t1, 2 * mnef - mnef.transpose(1, 0, 2, 3))
    # form_Wabef!(Wabef, abef=vvvv, amef=vovv, mbef=ovvv, mnef=oovv)
    Wabef = (vvvv - _es("mb,maef->abef", t1, vovv) - _es("ma,mbef->abef", t1, ovvv)
             + 0.5 * _es("mnab,mnef->abef", t2 + tt, oovv))
    # form_Wmnij!(Wmnij, mnij=oooo, mnie=ooov, mnej=oovo, mnef=oovv)
    Wmnij = (oooo + _es("je,mnie->mnij", t1, ooov) + _es("ie,mnej->mnij", t1, oovo)
             + 0.5 * _es("ijef,mnef->mnij", t2 + tt, oovv))
    # form_WmBeJ!(WmBeJ, mbej=ovvo, mbef=ovvv, mnej=oovo, mnef=oovv)
    WmBeJ = (ovvo + _es("jf,mbef->mbej", t1, ovvv) - _es("nb,mnej->mbej", t1, oovo)
             - 0.5 * _es("jnfb,mnef->mbej", t2 + 2 * tt, oovv)
             + 0.5 * _es("njfb,mnef->mbej", t2, 2 * oovv - oovv.transpose(1, 0, 2, 3)))
    # form_WmBEj!(WmBEj, bmej=vovo, mbfe=ovvv, nmej=oovo, nmef=oovv)
    WmBEj = (-_es("bmej->mbej", vovo) - _es("jf,mbfe->mbej", t1, ovvv)
             + _es("nb,nmej->mbej", t1, oovo)
             + _es("jnfb,nmef->mbej", 0.5 * t2 + tt, oovv))
    # form_T1(Fae,Fmi,Fme, amie=voov, maie=ovov, mnie=ooov, amef=vovv)
    n1 = (_es("ie,ae->ia", t1, Fae) - _es("ma,mi->ia", t1, Fmi)
          + _es("me,imae->ia", Fme, 2 * t2 - t2.transpose(1, 0, 2, 3))
          + _es("me,amie->ia", t1, 2 * voov) - _es("me,maie->ia", t1, ovov)
          - _es("mnae,mnie->ia", t2, 2 * ooov - ooov.transpose(1, 0, 2, 3))
          + _es("imef,maef->ia", t2, 2 * vovv - vovv.transpose(0, 1, 3, 2)))
    # form_T2(..., ijab=oovv, mbej=ovvo, amej=vovo, abej=vvvo, abie=vvov, mbij=ovoo, amij=vooo)
    Fb = Fae - 0.5 * _es("mb,me->be", t1, Fme)
    Fj = Fmi + 0.5 * _es("je,me->mj", t1, Fme)
    n2 = (oovv
          + _es("ijae,be->ijab", t2, Fb) + _es("ijeb,ae->ijab", t2, Fb)
          - _es("imab,mj->ijab", t2, Fj) - _es("mjab,mi->ijab", t2, Fj)
          + _es("mnab,mnij->ijab", t2 + tt, Wmnij)
          + _es("ijef,abef->ijab", t2 + tt, Wabef)
          + _es("imae,mbej->ijab", t2 - t2.transpose(1, 0, 2, 3), WmBeJ)
          - _es("imea,mbej->ijab", tt, ovvo)
          + _es("imae,mbej->ijab", t2, WmBeJ + WmBEj)
          + _es("mibe,maej->ijab", t2, WmBEj) - _es("imeb,amej->ijab", tt, vovo)
          + _es("mjae,mbei->ijab", t2, WmBEj) - _es("jmea,bmei->ijab", tt, vovo)
          + _es("jmbe,maei->ijab", t2 - t2.transpose(1, 0, 2, 3), WmBeJ)
          - _es("jmeb,maei->ijab", tt, ovvo)
          + _es("jmbe,maei->ijab", t2, WmBeJ + WmBEj)
          + _es("ei,eabj->ijab", t1.T, vvvo) + _es("ej,eabi->ijab", t1.T, vvov)
          - _es("ma,mbij->ijab", t1, ovoo) - _es("mb,maij->ijab", t1, vooo))
    return n1, n2


def solve_rccsd(I, eps, o, *, maxit=200, e_conv=1e-13, r_conv=1e-12, diis=True, max_vecs=5,
                update=rccsd_update, history=None):
    """Fixed point of mRCCSD.jl's update from the MP2 guess (:66).

    mRCCSD.jl runs 40 undamped iterations; here the same map is iterated to
    convergence, optionally with the DIIS of mRCCD.jl applied to [t1; t2].
    """
    v = eps.shape[0] - o
    Dia = form_Dia(eps, o, v)
    Dijab = form_Dijab(eps, o, v)
    oovv = I["oovv"]
    t1 = np.zeros((o, v))
    t2 = oovv / Dijab
    d = DIIS(max_vecs) if diis else None
    E_old = ccsd_energy(oovv, t1, t2)
    n1s = o * v
    for it in range(maxit):
        n1, n2 = update(t1, t2, I)
        nt1, nt2 = n1 / Dia, n2 / Dijab
        e1, e2 = nt1 - t1, nt2 - t2
        E_new = ccsd_energy(oovv, nt1, nt2)
        rn = float(np.sqrt(np.sum(e1 ** 2) + np.sum(e2 ** 2)))
        if history is not None:
            history.append((E_new, rn))
        if d is not None:
            d.push(np.concatenate([nt1.ravel(), nt2.ravel()]),
                   np.concatenate([e1.ravel(), e2.ravel()]))
            x, _ = d.extrapolate()
            t1, t2 = x[:n1s].reshape(o, v), x[n1s:].reshape(o, o, v, v)
        else:
            t1, t2 = nt1, nt2
        dE = E_new - E_old
        E_old = E_new
        if abs(dE) < e_conv and rn < r_conv:
            break
    return ccsd_energy(oovv, t1, t2), t1, t2, it + 1


# ---------------------------------------------------------------------------
# Density fitting: Backend/DF.jl:52-58, MollerPlesset/DF-RMP2.jl:1-46,
# CoupledCluster/DF-RCCD.jl:11-271
# ---------------------------------------------------------------------------

def metric_inv_sqrt(J):
    """``Jpq^(-1/2)`` of DF.jl:50 for a symmetric positive definite metric (eigen route)."""
    w, V = np.linalg.eigh(J)
    return (V / np.sqrt(w)) @ V.T


def make_bmunu(pqP, Jpqh):
    """DF.jl:52-58: ``b[p,q,Q] = sum_P pqP[p,q,P] Jpqh[P,Q]``."""
    return np.tensordot(pqP, Jpqh, axes=(2, 0))


def df_rmp2(b, Co, Cv, eps):
    """DF-RMP2.jl:17-43: half transforms to b_ia^Q, then the (i,j) pair loop with
    bAB[a,b] = sum_Q b[i,a,Q] b[j,b,Q], ibja = bAB[b,a]."""
    o = Co.shape[1]
    biv = _es("mi,mnQ->inQ", Co, b)                                   # :17-19
    bia = _es("na,inQ->iaQ", Cv, biv)                                 # :20-22
    ev = eps[o:o + Cv.shape[1]]
    E = 0.0
    for i in range(o):
        for j in range(o):
            bAB = bia[i] @ bia[j].T                                   # :33-35
            bBA = bAB.T                                               # :36
            D = eps[i] + eps[j] - ev[:, None] - ev[None, :]
            E += float(np.sum(bAB * (2.0 * bAB - bBA) / D))           # :37-42
    return E


def df_rccd_integrals(b, Cao, Cav):
    """DF-RCCD.jl:78-89: bov[i,a,Q], bvo[a,i,Q], boo[i,j,Q], bvv[a,b,Q]."""
    h = lambda C1, C2: _es("mp,mnQ,nq->pqQ", C1, b, C2)  # noqa: E731
    return dict(bov=h(Cao, Cav), bvo=h(Cav, Cao), boo=h(Cao, Cao), bvv=h(Cav, Cav))


def df_ccenergy(t, bov):
    """DF-RCCD.jl:91-111: ``sum bAB[a,b](2t[i,j,a,b] - t[j,i,a,b])``."""
    iajb = _es("iaQ,jbQ->ijab", bov, bov)
    return float(np.sum(iajb * (2.0 * t - t.transpose(1, 0, 2, 3))))


def df_rccd_sweep(t, b, D, literal=False):
    """One ``cciter`` (DF-RCCD.jl:55-76): the intermediates of :137-271 and
    ``form_T2`` (:172-219), every integral contracted from the factors.

    ``literal=True`` keeps DF-RCCD.jl:250 as written: the last WmBeJ term is
    -1/2 sum (me|nf) t[n,j,f,b].  The default uses (ne|mf) there, the
    <nm|ef> of RCCD.jl:400-401 (the two differ unless (me|nf) = (ne|mf)).
    """
    bov, bvo, boo, bvv = b["bov"], b["bvo"], b["boo"], b["bvv"]
    menf = _es("meQ,nfQ->menf", bov, bov)                                    # (me|nf)
    tx = t.transpose(1, 0, 2, 3)
    Fae = -_es("menf,mnaf->ae", menf, 2 * t - tx)                              # :148-150
    Fmi = _es("menf,inef->mi", menf, 2 * t - t.transpose(0, 1, 3, 2))          # :166-168
    Wmnij = _es("miQ,njQ->mnij", boo, boo) + _es("ijef,menf->mnij", t, menf) / 2   # :233-235
    exch = menf if literal else menf.transpose(2, 1, 0, 3)                      # (me|nf) | (ne|mf)
    WmBeJ = (_es("meQ,bjQ->mbej", bov, bvo) + _es("menf,njfb->mbej", menf, 2 * t - tx) / 2
             - _es("menf,njfb->mbej", exch, t) / 2)                             # :248-252
    WmBEj = (-_es("mjQ,beQ->mbej", boo, bvv)
             + _es("jnfb,nemf->mbej", t, menf) / 2)                             # :268-271
    R = (_es("iaQ,jbQ->ijab", bov, bov) + _es("ijae,be->ijab", t, Fae) + _es("jibe,ae->ijab", t, Fae)
         - _es("imab,mj->ijab", t, Fmi) - _es("mjab,mi->ijab", t, Fmi)
         + _es("mnab,mnij->ijab", t, Wmnij)
         + _es("ijef,mnab,menf->ijab", t, t, menf) / 2
         + _es("ijef,aeQ,bfQ->ijab", t, bvv, bvv)                                # :207 ladder
         + 2 * _es("imae,mbej->ijab", t, WmBeJ) - _es("miae,mbej->ijab", t, WmBeJ)
         + _es("imae,mbej->ijab", t, WmBEj) + _es("mibe,maej->ijab", t, WmBEj)
         + _es("mjae,mbei->ijab", t, WmBEj)
         + 2 * _es("jmbe,maei->ijab", t, WmBeJ) - _es("mjbe,maei->ijab", t, WmBeJ)
         + _es("jmbe,maei->ijab", t, WmBEj))                                    # :198-216
    return R / D                                                               # :217


def df_rccd(bmunu, Cao, Cav, eps, maxit=40, literal=False):
    """DF-RCCD.jl:11-53: T2_init! (:113-135), ``maxit`` sweeps with no DIIS and no
    convergence test, energy of the last amplitudes.  Returns (E, T2, energies)."""
    o, v = Cao.shape[1], Cav.shape[1]
    b = df_rccd_integrals(bmunu, Cao, Cav)
    D = form_Dijab(eps, o, v)
    t = _es("iaQ,jbQ->ijab", b["bov"], b["bov"]) / D
    hist = []
    for _ in range(maxit):
        t = df_rccd_sweep(t, b, D, literal)
        hist.append(df_ccenergy(t, b["bov"]))
    return df_ccenergy(t, b["bov"]), t, hist


# ---------------------------------------------------------------------------
# CoupledCluster/PerturbativeTriples.jl:35-138 and AutoRCCSD.jl:294-300
# ---------------------------------------------------------------------------

def pt_integrals(tei, Cao, Cav):
    """AutoRCCSD.jl:239-240,295-297: the phys blocks OOOV, OOVV, OVVV permuted into
    Vvvvo[a,b,c,i] = <ib|ca> = (ic|ba), Vvooo[a,i,j,k] = <ji|ka> = (jk|ia),
    Vvovo[a,i,b,j] = <ij|ab> = (ia|jb).  ``tei(C1,C2,C3,C4)`` returns chemist
    (pq|rs) over the four orbital blocks."""
    OOOV = np.transpose(tei(Cao, Cao, Cao, Cav), (0, 2, 1, 3))    # <ij|ka>
    OOVV = np.transpose(tei(Cao, Cav, Cao, Cav), (0, 2, 1, 3))    # <ij|ab>
    OVVV = np.transpose(tei(Cao, Cav, Cav, Cav), (0, 2, 1, 3))    # <ia|bc>
    Vvvvo = _pd(OVVV, (4, 2, 3, 1))
    Vvooo = _pd(OOOV, (4, 2, 1, 3))
    Vvovo = _pd(OOVV, (3, 1, 4, 2))
    return Vvvvo, Vvooo, Vvovo


def _pt_W(T2, Vvvvo, Vvooo, i, j, k):
    """PerturbativeTriples.jl:105-111: W[a,b,c] for one occupied triple."""
    T = lambda p, q: T2[p, q]                                  # noqa: E731  [c,d] blocks
    Ti = lambda p: T2[p]                                       # noqa: E731  [l,a,b]
    Vo = lambda p, q: Vvooo[:, p, q, :]                        # noqa: E731  [c,l]
    Vv = lambda p: Vvvvo[:, :, :, p]                           # noqa: E731  [b,d,a]
    return (_es("bda,cd->abc", Vv(i), T(k, j)) - _es("cl,lab->abc", Vo(k, j), Ti(i))
            + _es("cda,bd->abc", Vv(i), T(j, k)) - _es("bl,lac->abc", Vo(j, k), Ti(i))
            + _es("adc,bd->abc", Vv(k), T(j, i)) - _es("bl,lca->abc", Vo(j, i), Ti(k))
            + _es("bdc,ad->abc", Vv(k), T(i, j)) - _es("al,lcb->abc", Vo(i, j), Ti(k))
            + _es("cdb,ad->abc", Vv(j), T(i, k)) - _es("al,lbc->abc", Vo(i, k), Ti(j))
            + _es("adb,cd->abc", Vv(j), T(k, i)) - _es("cl,lba->abc", Vo(k, i), Ti(j)))


def _pt_summand(W, V):
    """PerturbativeTriples.jl:118-125 for every (a,b,c): the bracket E before the
    division by D; symmetric under any permutation of (a,b,c)."""
    P = lambda A, s: np.einsum(s + "->abc", A)                 # noqa: E731
    Wc = W + P(W, "bca") + P(W, "cab")                         # W[abc]+W[bca]+W[cab]
    Wa = P(W, "acb") + P(W, "bac") + P(W, "cba")
    Y = V + P(V, "bca") + P(V, "cab")
    Z = P(V, "acb") + P(V, "bac") + P(V, "cba")
    X = sum(P(W, s) * P(V, s) for s in ("abc", "acb", "bac", "bca", "cab", "cba"))
    return (Y - 2 * Z) * Wc + (Z - 2 * Y) * Wa + 3 * X


def compute_pT(T1, T2, Vvvvo, Vvooo, Vvovo, fo, fv):
    """PerturbativeTriples.jl:35-138: i>=j>=k, W and V per triple (:105-113), the
    a>=b>=c sum with weights (2-d_ij-d_jk)/(D (1+d_ab+d_bc)) (:116-130)."""
    o, v = T1.shape
    a, b, c = np.meshgrid(np.arange(v), np.arange(v), np.arange(v), indexing="ij")
    mask = (a >= b) & (b >= c)
    wabc = np.where(mask, 1.0 / (1.0 + (a == b) + (b == c)), 0.0)
    Et = 0.0
    for i in range(o):
        for j in range(i + 1):
            for k in range(j + 1):
                W = _pt_W(T2, Vvvvo, Vvooo, i, j, k)
                V = (W + _es("bc,a->abc", Vvovo[:, j, :, k], T1[i]) + _es("ac,b->abc", Vvovo[:, i, :, k], T1[j])
                     + _es("ab,c->abc", Vvovo[:, i, :, j], T1[k]))                       # :112
                D = fo[i] + fo[j] + fo[k] - fv[:, None, None] - fv[None, :, None] - fv[None, None, :]
                E = _pt_summand(W, V)
                Et += (2 - (i == j) - (j == k)) * float(np.sum(E * wabc / D))
    return Et


def pT_spin_orbital(T1, T2, g_mo, eps, o):
    """Independent check of ``compute_pT``: the spin-orbital (T) of Raghavachari et al.
    (Chem. Phys. Lett. 157, 479 (1989)) in the form
    D t_c = P(i/jk)P(a/bc)[sum_e t_jk^ae <ei||bc> - sum_m t_im^bc <ma||jk>],
    D t_d = P(i/jk)P(a/bc) t_i^a <jk||bc>,  E = (1/36) sum t_c D (t_c + t_d),
    with P(i/jk) f = f(ijk) - f(jik) - f(kji).  Spin orbitals are built from the
    closed-shell T1, T2 (pair-symmetric), chemist MO integrals g_mo over all
    orbitals and canonical orbital energies eps."""
    n = g_mo.shape[0]
    ns = 2 * n
    spin = np.arange(ns) % 2
    sp = np.arange(ns) // 2
    # <PQ|RS> = (pr|qs) delta(sP,sR) delta(sQ,sS), antisymmetrised
    phys = np.transpose(g_mo, (0, 2, 1, 3))[np.ix_(sp, sp, sp, sp)]
    phys = phys * (spin[:, None, None, None] == spin[None, None, :, None]) \
        * (spin[None, :, None, None] == spin[None, None, None, :])
    A = phys - phys.transpose(0, 1, 3, 2)
    occ = np.sort(np.where(sp < o)[0])
    vir = np.sort(np.where(sp >= o)[0])
    so, sv = sp[occ], sp[vir] - o
    ao, av = spin[occ], spin[vir]
    t1 = T1[np.ix_(so, sv)] * (ao[:, None] == av[None, :])
    t2 = (T2[np.ix_(so, so, sv, sv)] * (ao[:, None, None, None] == av[None, None, :, None])
          * (ao[None, :, None, None] == av[None, None, None, :])
          - T2[np.ix_(so, so, sv, sv)].transpose(0, 1, 3, 2)
          * (ao[:, None, None, None] == av[None, None, None, :]) * (ao[None, :, None, None] == av[None, None, :, None]))
    e_o, e_v = eps[so], eps[sv + o]
    D = (e_o[:, None, None, None, None, None] + e_o[None, :, None, None, None, None]
         + e_o[None, None, :, None, None, None] - e_v[None, None, None, :, None, None]
         - e_v[None, None, None, None, :, None] - e_v[None, None, None, None, None, :])
    vovv = A[np.ix_(vir, occ, vir, vir)]      # <ei||bc>
    ovoo = A[np.ix_(occ, vir, occ, occ)]      # <ma||jk>
    oovv = A[np.ix_(occ, occ, vir, vir)]      # <jk||bc>

    def P(x):
        y = x - x.transpose(1, 0, 2, 3, 4, 5) - x.transpose(2, 1, 0, 3, 4, 5)
        return y - y.transpose(0, 1, 2, 4, 3, 5) - y.transpose(0, 1, 2, 5, 4, 3)

    tc = P(_es("jkae,eibc->ijkabc", t2, vovv) - _es("imbc,majk->ijkabc", t2, ovoo)) / D
    td = P(_es("ia,jkbc->ijkabc", t1, oovv)) / D
    return float(np.sum(tc * D * (tc + td)) / 36.0)


# ---------------------------------------------------------------------------
# MollerPlesset/UMP2.jl:12-79 and IntegralTransformation.get_fock (:119-141)
# ---------------------------------------------------------------------------

def ump2(gao, Cao, Cav, Cbo, Cbv, epsa, epsb):
    """UMP2.jl:12-79: alpha-alpha over i > j, a > b of <ij||ab>^2 / D (:30-44), alpha-beta over
    all (i,J,a,B) of (ia|JB)^2 / D (:46-59), beta-beta as alpha-alpha (:61-77)."""
    na, nb = Cao.shape[1], Cbo.shape[1]
    E = 0.0
    for Co, Cv, eps, no in ((Cao, Cav, epsa, na), (Cbo, Cbv, epsb, nb)):
        K = tei_transform(gao, Co, Cv, Co, Cv)                      # (ia|jb)
        A = K - K.transpose(0, 3, 2, 1)                              # <ij||ab> at [i,a,j,b]
        eo, ev = eps[:no], eps[no:]
        D = eo[:, None, None, None] + eo[None, None, :, None] - ev[None, :, None, None] - ev[None, None, None, :]
        i, a, j, b = np.meshgrid(*(np.arange(s) for s in K.shape), indexing="ij")
        E += float(np.sum(np.where((i > j) & (a > b), A * A / D, 0.0)))
    K = tei_transform(gao, Cao, Cav, Cbo, Cbv)                       # (ia|JB)
    eoa, eva, eob, evb = epsa[:na], epsa[na:], epsb[:nb], epsb[nb:]
    D = eoa[:, None, None, None] + eob[None, None, :, None] - eva[None, :, None, None] - evb[None, None, None, :]
    return E + float(np.sum(K * K / D))


def ump2_spin_orbital(gao, Ca, Cb, epsa, epsb, na, nb):
    """Independent check: 1/4 sum_{IJAB} |<IJ||AB>|^2 / D over spin orbitals built from
    the UHF alpha and beta orbitals."""
    n = Ca.shape[0]
    C = np.zeros((2 * n, 2 * n))
    C[:n, :n], C[n:, n:] = Ca, Cb                                     # AO (alpha, beta) x MO (alpha, beta)
    spin_ao = np.r_[np.zeros(n), np.ones(n)]
    g2 = np.zeros((2 * n,) * 4)
    for s1 in (0, 1):
        for s2 in (0, 1):
            g2[s1 * n:(s1 + 1) * n, s1 * n:(s1 + 1) * n, s2 * n:(s2 + 1) * n, s2 * n:(s2 + 1) * n] = gao
    gm = tei_transform(g2, C, C, C, C)                                # (pq|rs) over spin orbitals
    phys = gm.transpose(0, 2, 1, 3)
    A = phys - phys.transpose(0, 1, 3, 2)
    eps = np.r_[epsa, epsb]
    occ = np.r_[np.arange(na), n + np.arange(nb)]
    vir = np.r_[np.arange(na, n), n + np.arange(nb, n)]
    Ao = A[np.ix_(occ, occ, vir, vir)]
    eo, ev = eps[occ], eps[vir]
    D = eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]
    del spin_ao
    return float(np.sum(Ao * Ao / D) / 4.0)


def get_fock(gao, hao, C, Co):
    """IntegralTransformation.jl:119-141: f = C^T h C + 2 C^T J[Co Co^T] C - C^T K[Co Co^T] C."""
    f = C.T @ hao @ C
    f += 2 * _es("mp,nq,lk,sk,mnls->pq", C, C, Co, Co, gao)
    f -= _es("mp,nq,lk,sk,mlns->pq", C, C, Co, Co, gao)
    return f


# ---------------------------------------------------------------------------
# CoupledCluster/AutoRCCSD.jl:39-300: CCSD with the full Fock matrix
# ---------------------------------------------------------------------------

def auto_rccsd_integrals(tei, Cao, Cav):
    """AutoRCCSD.jl:239-240: phys OOOO, OOOV, OOVV, OVOV, OVVV, VVVV."""
    ph = lambda a, b, c, d: np.transpose(tei(a, c, b, d), (0, 2, 1, 3))  # noqa: E731  <pq|rs>
    o, v = Cao, Cav
    return dict(oooo=ph(o, o, o, o), ooov=ph(o, o, o, v), oovv=ph(o, o, v, v), ovov=ph(o, v, o, v),
                ovvv=ph(o, v, v, v), vvvv=ph(v, v, v, v))


def auto_rccsd_energy(T1, T2, fov, oovv):
    """AutoRCCSD.jl:39-51."""
    B = -_es("lc,kd->lckd", T1, T1) - _es("lkcd->lckd", T2) + 2 * _es("klcd->lckd", T2)
    return float(2 * np.sum(fov * T1) + _es("lckd,klcd->", B, oovv) + 2 * _es("lc,kd,lkcd->", T1, T1, oovv))


# AutoRCCSD.jl:76-165, one row per term: (coefficient, einsum inputs, operand names);
# F? are the off-diagonal Fock blocks, V.... the phys integral blocks.
_AUTO_T1 = [
    (1, "ia", "Fov"), (-1, "ik,ka", "Foo T1"), (1, "ca,ic", "Fvv T1"), (-1, "kc,ic,ka", "Fov T1 T1"),
    (2, "kc,ikac", "Fov T2"), (-1, "kc,kiac", "Fov T2"), (-1, "kc,icka", "T1 ovov"), (2, "kc,kica", "T1 oovv"),
    (-1, "kicd,kadc", "T2 ovvv"), (2, "ikcd,kadc", "T2 ovvv"), (-2, "klac,klic", "T2 ooov"),
    (1, "lkac,klic", "T2 ooov"), (-2, "kc,la,lkic", "T1 T1 ooov"), (-1, "kc,id,kadc", "T1 T1 ovvv"),
    (2, "kc,id,kacd", "T1 T1 ovvv"), (1, "kc,la,klic", "T1 T1 ooov"), (-2, "kc,ilad,lkcd", "T1 T2 oovv"),
    (-2, "kc,liad,klcd", "T1 T2 oovv"), (1, "kc,liad,lkcd", "T1 T2 oovv"), (-2, "ic,lkad,lkcd", "T1 T2 oovv"),
    (1, "ic,lkad,klcd", "T1 T2 oovv"), (-2, "la,ikdc,klcd", "T1 T2 oovv"), (1, "la,ikcd,klcd", "T1 T2 oovv"),
    (1, "kc,id,la,lkcd", "T1 T1 T1 oovv"), (-2, "kc,id,la,klcd", "T1 T1 T1 oovv"),
    (4, "kc,ilad,klcd", "T1 T2 oovv"),
]
_AUTO_T2 = [
    (1, "ijab", "oovv"), (1, "ic,jd,cdab", "T1 T1 vvvv"), (1, "ijcd,cdab", "T2 vvvv"),
    (1, "ka,lb,ijkl", "T1 T1 oooo"), (1, "klab,ijkl", "T2 oooo"), (-1, "ic,jd,ka,kbcd", "T1 T1 T1 ovvv"),
    (-1, "ic,jd,kb,kadc", "T1 T1 T1 ovvv"), (1, "ic,ka,lb,lkjc", "T1 T1 T1 ooov"),
    (1, "jc,ka,lb,klic", "T1 T1 T1 ooov"), (1, "klac,ijdb,klcd", "T2 T2 oovv"),
    (-2, "ikac,ljbd,klcd", "T2 T2 oovv"), (-2, "lkac,ijdb,klcd", "T2 T2 oovv"),
    (1, "kiac,ljdb,lkcd", "T2 T2 oovv"), (1, "ikac,ljbd,lkcd", "T2 T2 oovv"),
    (-2, "ikac,jlbd,lkcd", "T2 T2 oovv"), (1, "kiac,ljbd,klcd", "T2 T2 oovv"),
    (-2, "kiac,jlbd,klcd", "T2 T2 oovv"), (1, "ijac,lkbd,klcd", "T2 T2 oovv"),
    (-2, "ijac,klbd,klcd", "T2 T2 oovv"), (1, "kjac,ildb,lkcd", "T2 T2 oovv"),
    (4, "ikac,jlbd,klcd", "T2 T2 oovv"), (1, "ijdc,lkab,klcd", "T2 T2 oovv"),
    (1, "ic,jd,ka,lb,klcd", "T1 T1 T1 T1 oovv"), (1, "ic,jd,lkab,lkcd", "T1 T1 T2 oovv"),
    (1, "ka,lb,ijdc,lkcd", "T1 T1 T2 oovv"),
]
_AUTO_P = [
    (-1, "ik,kjab", "Foo T2"), (1, "ca,ijcb", "Fvv T2"), (-1, "kb,jika", "T1 ooov"), (1, "jc,icab", "T1 ovvv"),
    (-1, "kc,ic,kjab", "Fov T1 T2"), (-1, "kc,ka,ijcb", "Fov T1 T2"), (-1, "kiac,kjcb", "T2 oovv"),
    (-1, "ic,ka,kjcb", "T1 T1 oovv"), (-1, "ic,kb,jcka", "T1 T1 ovov"), (2, "ikac,kjcb", "T2 oovv"),
    (-1, "ikac,jckb", "T2 ovov"), (-1, "kjac,ickb", "T2 ovov"), (-2, "lb,ikac,lkjc", "T1 T2 ooov"),
    (1, "lb,kiac,lkjc", "T1 T2 ooov"), (-1, "jc,ikdb,kacd", "T1 T2 ovvv"), (-1, "jc,kiad,kbdc", "T1 T2 ovvv"),
    (-1, "jc,ikad,kbcd", "T1 T2 ovvv"), (1, "jc,lkab,lkic", "T1 T2 ooov"), (1, "lb,ikac,kljc", "T1 T2 ooov"),
    (-1, "ka,ijdc,kbdc", "T1 T2 ovvv"), (1, "ka,ilcb,lkjc", "T1 T2 ooov"), (2, "jc,ikad,kbdc", "T1 T2 ovvv"),
    (-1, "kc,ijad,kbdc", "T1 T2 ovvv"), (2, "kc,ijad,kbcd", "T1 T2 ovvv"), (1, "kc,ilab,kljc", "T1 T2 ooov"),
    (-2, "kc,ilab,lkjc", "T1 T2 ooov"), (1, "jkcd,ilab,klcd", "T2 T2 oovv"),
    (-2, "kc,jd,ilab,klcd", "T1 T1 T2 oovv"), (1, "kc,jd,ilab,lkcd", "T1 T1 T2 oovv"),
    (-2, "kc,la,ijdb,klcd", "T1 T1 T2 oovv"), (1, "kc,la,ijdb,lkcd", "T1 T1 T2 oovv"),
    (1, "ic,ka,ljbd,klcd", "T1 T1 T2 oovv"), (-2, "ic,ka,jlbd,klcd", "T1 T1 T2 oovv"),
    (1, "ic,ka,ljdb,lkcd", "T1 T1 T2 oovv"), (1, "ic,lb,kjad,klcd", "T1 T1 T2 oovv"),
    (-2, "ikdc,ljab,klcd", "T2 T2 oovv"),
]


def _auto_sum(terms, out, ops):
    acc = 0.0
    for coef, spec, names in terms:
        acc = acc + coef * _es(spec + "->" + out, *[ops[nm] for nm in names.split()])
    return acc


def auto_rccsd_update(T1, T2, F, V, d, D):
    """AutoRCCSD.jl:67-177: new T1, T2 and the residues r1, r2 (:170-174).
    F = (fock_OO, fock_OV, fock_VV) with the diagonal removed; V = auto_rccsd_integrals."""
    ops = dict(T1=T1, T2=T2, Foo=F[0], Fov=F[1], Fvv=F[2], **V)
    n1 = _auto_sum(_AUTO_T1, "ia", ops)
    P = _auto_sum(_AUTO_P, "ijab", ops)
    n2 = _auto_sum(_AUTO_T2, "ijab", ops) + P + P.transpose(1, 0, 3, 2)
    t1, t2 = n1 / d, n2 / D
    r1 = np.sqrt(np.sum((t1 - T1) ** 2)) / T1.size
    r2 = np.sqrt(np.sum((t2 - T2) ** 2)) / T2.size
    return t1, t2, r1, r2


def solve_auto_rccsd(fock, V, o, *, e_conv=1e-10, max_rms=1e-10, maxit=200):
    """AutoRCCSD.jl:193-292 (fcn = 0): Fock diagonal in the resolvents, MP2-like guess
    T1 = f_ov/d, T2 = <ij|ab>/D, plain updates until |dE| <= e_conv and max(r1,r2) <= max_rms."""
    fd = np.diag(fock)
    f = fock - np.diag(fd)
    F = (f[:o, :o], f[:o, o:], f[o:, o:])
    eo, ev = fd[:o], fd[o:]
    d = eo[:, None] - ev[None, :]
    D = eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]
    T1, T2 = F[1] / d, V["oovv"] / D
    E = auto_rccsd_energy(T1, T2, F[1], V["oovv"])
    dE, rms, it = 1.0, 1.0, 0
    while abs(dE) > e_conv or rms > max_rms:
        if it >= maxit:
            break
        T1, T2, r1, r2 = auto_rccsd_update(T1, T2, F, V, d, D)
        rms = max(r1, r2)
        E_old, E = E, auto_rccsd_energy(T1, T2, F[1], V["oovv"])
        dE = E - E_old
        it += 1
    return E, T1, T2, it


def ccsd_spin_orbital(fock, g_mo, o, *, e_conv=1e-13, r_conv=1e-12, maxit=400):
    """Independent check of the Fock-driven CCSD: the spin-orbital CCSD equations of
    Stanton, Gauss, Watts and Bartlett (J. Chem. Phys. 94, 4334 (1991)) with the full
    Fock matrix (any reference), from a closed-shell MO Fock matrix and chemist MO
    integrals; Jacobi iterations with DIIS.  Returns the correlation energy
    sum f_ia t_ia + 1/4 <ij||ab> t_ijab + 1/2 <ij||ab> t_ia t_jb."""
    n = fock.shape[0]
    sp = np.repeat(np.arange(n), 2)
    spin = np.tile([0, 1], n)
    same = spin[:, None] == spin[None, :]
    f = fock[np.ix_(sp, sp)] * same
    phys = np.transpose(g_mo, (0, 2, 1, 3))[np.ix_(sp, sp, sp, sp)]
    phys = phys * same[:, None, :, None] * same[None, :, None, :]
    A = phys - phys.transpose(0, 1, 3, 2)
    O_, V_ = slice(0, 2 * o), slice(2 * o, 2 * n)
    fd = np.diag(f)
    Dia = fd[O_, None] - fd[None, V_]
    Dijab = fd[O_, None, None, None] + fd[None, O_, None, None] - fd[None, None, V_, None] - fd[None, None, None, V_]
    g = lambda a, b, c, d: A[a, b, c, d]  # noqa: E731
    oo, ov, vv = f[O_, O_], f[O_, V_], f[V_, V_]
    t1 = ov / Dia
    t2 = g(O_, O_, V_, V_) / Dijab
    no, nv = 2 * o, 2 * (n - o)
    Io, Iv = np.eye(no), np.eye(nv)

    def energy(t1, t2):
        return float(np.sum(ov * t1) + 0.25 * np.sum(g(O_, O_, V_, V_) * t2)
                     + 0.5 * _es("ijab,ia,jb->", g(O_, O_, V_, V_), t1, t1))

    hist = []
    E = energy(t1, t2)
    for _ in range(maxit):
        taut = t2 + 0.5 * (_es("ia,jb->ijab", t1, t1) - _es("ib,ja->ijab", t1, t1))
        tau = t2 + _es("ia,jb->ijab", t1, t1) - _es("ib,ja->ijab", t1, t1)
        Fae = (1 - Iv) * vv - 0.5 * _es("me,ma->ae", ov, t1) + _es("mf,mafe->ae", t1, g(O_, V_, V_, V_)) \
            - 0.5 * _es("mnaf,mnef->ae", taut, g(O_, O_, V_, V_))
        Fmi = (1 - Io) * oo + 0.5 * _es("ie,me->mi", t1, ov) + _es("ne,mnie->mi", t1, g(O_, O_, O_, V_)) \
            + 0.5 * _es("inef,mnef->mi", taut, g(O_, O_, V_, V_))
        Fme = ov + _es("nf,mnef->me", t1, g(O_, O_, V_, V_))
        Wmnij = g(O_, O_, O_, O_) + _es("je,mnie->mnij", t1, g(O_, O_, O_, V_)) \
            - _es("ie,mnje->mnij", t1, g(O_, O_, O_, V_)) + 0.25 * _es("ijef,mnef->mnij", tau, g(O_, O_, V_, V_))
        Wabef = g(V_, V_, V_, V_) - _es("mb,amef->abef", t1, g(V_, O_, V_, V_)) \
            + _es("ma,bmef->abef", t1, g(V_, O_, V_, V_)) + 0.25 * _es("mnab,mnef->abef", tau, g(O_, O_, V_, V_))
        Wmbej = g(O_, V_, V_, O_) + _es("jf,mbef->mbej", t1, g(O_, V_, V_, V_)) \
            - _es("nb,mnej->mbej", t1, g(O_, O_, V_, O_)) \
            - _es("jnfb,mnef->mbej", 0.5 * t2 + _es("jf,nb->jnfb", t1, t1), g(O_, O_, V_, V_))
        r1 = ov + _es("ie,ae->ia", t1, Fae) - _es("ma,mi->ia", t1, Fmi) + _es("imae,me->ia", t2, Fme) \
            - _es("nf,naif->ia", t1, g(O_, V_, O_, V_)) - 0.5 * _es("imef,maef->ia", t2, g(O_, V_, V_, V_)) \
            - 0.5 * _es("mnae,nmei->ia", t2, g(O_, O_, V_, O_))
        Fb = Fae - 0.5 * _es("mb,me->be", t1, Fme)
        Fj = Fmi + 0.5 * _es("je,me->mj", t1, Fme)
        x = _es("ijae,be->ijab", t2, Fb)
        r2 = g(O_, O_, V_, V_) + x - x.transpose(0, 1, 3, 2)
        x = _es("imab,mj->ijab", t2, Fj)
        r2 -= x - x.transpose(1, 0, 2, 3)
        r2 += 0.5 * _es("mnab,mnij->ijab", tau, Wmnij) + 0.5 * _es("ijef,abef->ijab", tau, Wabef)
        x = _es("imae,mbej->ijab", t2, Wmbej) - _es("ie,ma,mbej->ijab", t1, t1, g(O_, V_, V_, O_))
        r2 += x - x.transpose(1, 0, 2, 3) - x.transpose(0, 1, 3, 2) + x.transpose(1, 0, 3, 2)
        x = _es("ie,abej->ijab", t1, g(V_, V_, V_, O_))
        r2 += x - x.transpose(1, 0, 2, 3)
        x = _es("ma,mbij->ijab", t1, g(O_, V_, O_, O_))
        r2 -= x - x.transpose(0, 1, 3, 2)
        n1, n2 = r1 / Dia, r2 / Dijab
        e = np.concatenate([(n1 - t1).ravel(), (n2 - t2).ravel()])
        hist.append((np.concatenate([n1.ravel(), n2.ravel()]), e))
        hist = hist[-8:]
        m = len(hist)
        B = -np.ones((m + 1, m + 1))
        B[m, m] = 0
        for a_ in range(m):
            for b_ in range(m):
                B[a_, b_] = hist[a_][1] @ hist[b_][1]
        rhs = np.zeros(m + 1)
        rhs[m] = -1
        c = np.linalg.solve(B, rhs)
        x = sum(c[k] * hist[k][0] for k in range(m))
        t1, t2 = x[:t1.size].reshape(t1.shape), x[t1.size:].reshape(t2.shape)
        E_old, E = E, energy(t1, t2)
        if abs(E - E_old) < e_conv and np.linalg.norm(e) < r_conv:
            break
    return E
