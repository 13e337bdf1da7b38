"""Perturbative triples (T): PerturbativeTriples.compute_pT (PerturbativeTriples.jl:35-138)
and the call after CCSD (AutoRCCSD.jl:294-300).

CPU: the oracle's literal restatement against an independent spin-orbital (T)
(Raghavachari et al.) built from the same closed-shell amplitudes and integrals,
with and without T1, on random pair-symmetric amplitudes and on converged RCCSD
amplitudes of psi4's water in STO-3G.  The reference's tests hold no (T) value.
GPU: jues_compute_pT (host arrays, the reference's interface) and jues_rccsd_pT
(integrals built on the device) against the oracle, at virtual sizes below,
at and across the 8-wide energy tile and the 128-wide GEMM tile.
"""
import numpy as np
import pytest

import jues_oracle as O
import sto3g


def _system(n, o, naux, seed, scale=0.3, gap=3.0):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((n, n, naux)) * scale
    B = (B + B.transpose(1, 0, 2)) / 2
    C = np.linalg.qr(rng.standard_normal((n, n)))[0]
    eps = np.sort(rng.standard_normal(n))
    eps[o:] += gap
    v = n - o
    T1 = rng.standard_normal((o, v)) * 0.1
    T2 = rng.standard_normal((o, o, v, v)) * 0.1
    T2 = (T2 + T2.transpose(1, 0, 3, 2)) / 2
    return dict(B=np.asfortranarray(B), C=np.asfortranarray(C), eps=eps, o=o, v=v, T1=T1, T2=T2)


def _ints(s):
    g = np.einsum("mnP,lsP->mnls", s["B"], s["B"])
    Co, Cv = s["C"][:, :s["o"]], s["C"][:, s["o"]:]
    tei = lambda a, b, c, d: O.tei_transform(g, a, b, c, d)  # noqa: E731
    return g, O.pt_integrals(tei, Co, Cv)


def _oracle(s):
    _, (Vvvvo, Vvooo, Vvovo) = _ints(s)
    o = s["o"]
    return O.compute_pT(s["T1"], s["T2"], Vvvvo, Vvooo, Vvovo, s["eps"][:o], s["eps"][o:])


@pytest.fixture(scope="module")
def water():
    w = sto3g.system(sto3g.zmat_water())
    o = w["nocc"]
    Co, Cv = w["C"][:, :o], w["C"][:, o:]
    I = O.make_rccsd15_integrals(lambda a, b, c, d: O.tei_transform(w["g"], a, b, c, d), Co, Cv)
    E, T1, T2, _ = O.solve_rccsd(I, w["eps"], o, update=O.rccsd15_update, e_conv=1e-14, r_conv=1e-13, maxit=200)
    tei = lambda a, b, c, d: O.tei_transform(w["g"], a, b, c, d)  # noqa: E731
    V = O.pt_integrals(tei, Co, Cv)
    Et = O.compute_pT(T1, T2, *V, w["eps"][:o], w["eps"][o:])
    return dict(w=w, T1=T1, T2=T2, V=V, Et=Et, Ecc=E)


# ---------------------------------------------------------------- oracle (CPU)


@pytest.mark.parametrize("with_t1", [True, False])
def test_oracle_pT_matches_spin_orbital(with_t1):
    s = _system(9, 3, 20, seed=3)
    if not with_t1:
        s["T1"] = np.zeros_like(s["T1"])
    g, (Vvvvo, Vvooo, Vvovo) = _ints(s)
    o = s["o"]
    E1 = O.compute_pT(s["T1"], s["T2"], Vvvvo, Vvooo, Vvovo, s["eps"][:o], s["eps"][o:])
    E2 = O.pT_spin_orbital(s["T1"], s["T2"], O.tei_transform(g, s["C"], s["C"], s["C"], s["C"]), s["eps"], o)
    assert abs(E1 - E2) < 1e-12 * abs(E2)


def test_oracle_pT_water(water):
    w = water["w"]
    o = w["nocc"]
    gmo = O.tei_transform(w["g"], w["C"], w["C"], w["C"], w["C"])
    E2 = O.pT_spin_orbital(water["T1"], water["T2"], gmo, w["eps"], o)
    assert abs(water["Et"] - E2) < 1e-14
    assert -1e-3 < water["Et"] < 0.0


# ---------------------------------------------------------------- GPU


@pytest.mark.gpu
@pytest.mark.parametrize("n,o,seed", [(9, 3, 3), (13, 4, 4), (12, 3, 5), (41, 4, 6)])
def test_gpu_compute_pT_arrays(n, o, seed):
    """v = 6, 9, 9 (one partial energy tile) and 37 (five tiles, ragged)."""
    from jues import CoupledCluster as CC
    s = _system(n, o, 2 * n, seed)
    _, (Vvvvo, Vvooo, Vvovo) = _ints(s)
    ref = O.compute_pT(s["T1"], s["T2"], Vvvvo, Vvooo, Vvovo, s["eps"][:o], s["eps"][o:])
    kw = dict(T1=s["T1"], T2=s["T2"], Vvvvo=Vvvvo, Vvooo=Vvooo, Vvovo=Vvovo, fo=s["eps"][:o], fv=s["eps"][o:])
    assert abs(CC.PerturbativeTriples.compute_pT(**kw) - ref) < 1e-12 * abs(ref)
    assert abs(CC.PerturbativeTriples.compute_pT(prec="fp32_mixed", **kw) - ref) < 2e-5 * abs(ref)


@pytest.mark.gpu
@pytest.mark.parametrize("dense", [True, False])
def test_gpu_rccsd_pT_from_ao(dense):
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn, FactoredERI
    s = _system(41, 5, 60, seed=7)
    g, _ = _ints(s)
    ref = _oracle(s)
    eri = np.asfortranarray(g) if dense else FactoredERI(s["B"])
    w = Wfn.from_orbitals(s["C"], s["eps"], s["o"], eri)
    assert abs(CC.PerturbativeTriples.rccsd_pT(w, s["T1"], s["T2"]) - ref) < 1e-12 * abs(ref)


@pytest.mark.gpu
def test_gpu_pT_water_after_rccsd(water):
    """psi4 water: RCCSD (RCCSD.jl equations) on the GPU, then (T) on the GPU."""
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn
    w = water["w"]
    wf = Wfn.from_orbitals(w["C"], w["eps"], w["nocc"], np.asfortranarray(w["g"]))
    E, T1, T2 = CC.RCCSD.do_rccsd(wf, return_T=True, cc_e_conv=1e-13, cc_max_rms=1e-12, maxit=200)
    assert abs(E - water["Ecc"]) < 1e-10
    Et = CC.PerturbativeTriples.rccsd_pT(wf, T1, T2)
    assert abs(Et - water["Et"]) < 1e-12
    o = w["nocc"]
    Et2 = CC.PerturbativeTriples.compute_pT(T1=T1, T2=T2, Vvvvo=water["V"][0], Vvooo=water["V"][1],
                                            Vvovo=water["V"][2], fo=w["eps"][:o], fv=w["eps"][o:])
    assert abs(Et2 - water["Et"]) < 1e-12


@pytest.mark.gpu
def test_gpu_pT_bad_shapes():
    from jues import CoupledCluster as CC
    s = _system(9, 3, 20, seed=3)
    _, (Vvvvo, Vvooo, Vvovo) = _ints(s)
    with pytest.raises(ValueError):
        CC.PerturbativeTriples.compute_pT(T1=s["T1"], T2=s["T2"], Vvvvo=Vvvvo[:, :, :, :2], Vvooo=Vvooo,
                                          Vvovo=Vvovo, fo=s["eps"][:3], fv=s["eps"][3:])
