"""Open-shell MP2 (MollerPlesset/UMP2.jl:12-79) and the MO Fock matrix
(Backend/IntegralTransformation.jl:119-141).

Known answer: UMP2 of psi4's UHF water cation in STO-3G (charge 1, doublet,
"O; H 1 1.1; H 1 1.1 2 104.0", PK integrals) is -0.03588729625230
(test/TestMollerPlesset.jl:36), replayed from oracle/sto3g.py's integrals and UHF.
test/TestMollerPlesset.jl:37 also asserts do_rmp2 == do_ump2 on a closed shell.
"""
import numpy as np
import pytest

import jues_oracle as O
import sto3g

E_UMP2_CATION = -0.03588729625230


@pytest.fixture(scope="module")
def cation():
    return sto3g.system_uhf(sto3g.zmat_water(), 5, 4)


@pytest.fixture(scope="module")
def water():
    return sto3g.system(sto3g.zmat_water())


def _split(w):
    na, nb = w["na"], w["nb"]
    return w["Ca"][:, :na], w["Ca"][:, na:], w["Cb"][:, :nb], w["Cb"][:, nb:]


def _random_uhf(n=14, na=5, nb=3, naux=40, seed=9):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((n, n, naux)) * 0.2
    B = (B + B.transpose(1, 0, 2)) / 2
    Ca = np.linalg.qr(rng.standard_normal((n, n)))[0]
    Cb = np.linalg.qr(rng.standard_normal((n, n)))[0]
    ea = np.sort(rng.standard_normal(n)); ea[na:] += 2.5
    eb = np.sort(rng.standard_normal(n)); eb[nb:] += 2.5
    return dict(B=np.asfortranarray(B), Ca=Ca, Cb=Cb, epsa=ea, epsb=eb, na=na, nb=nb,
                g=np.einsum("mnP,lsP->mnls", B, B))


# ---------------------------------------------------------------- oracle (CPU)


def test_oracle_ump2_cation_known_answer(cation):
    w = cation
    E = O.ump2(w["g"], *_split(w), w["epsa"], w["epsb"])
    assert abs(E - E_UMP2_CATION) < 1e-12
    E2 = O.ump2_spin_orbital(w["g"], w["Ca"], w["Cb"], w["epsa"], w["epsb"], w["na"], w["nb"])
    assert abs(E - E2) < 1e-14


def _ump2_from_get_eri_blocks(blk, epsa, epsb, na, nb):
    """UMP2 from get_eri's physicist blocks: 1/4 sum <IJ||AB>^2/D over "OOVV" and "oovv",
    plus sum <Ij|Ab>^2/D over "OoVv" (upper case alpha, lower case beta)."""
    def den(e1o, e2o, e1v, e2v):
        return e1o[:, None, None, None] + e2o[None, :, None, None] - e1v[None, None, :, None] - e2v[None, None, None, :]
    ea_o, ea_v, eb_o, eb_v = epsa[:na], epsa[na:], epsb[:nb], epsb[nb:]
    E = 0.0
    for s, eo, ev in (("OOVV", ea_o, ea_v), ("oovv", eb_o, eb_v)):
        V = blk(s)
        A = V - V.transpose(0, 1, 3, 2)
        E += 0.25 * float(np.sum(A * A / den(eo, eo, ev, ev)))
    V = blk("OoVv")
    return E + float(np.sum(V * V / den(ea_o, eb_o, ea_v, eb_v)))


def test_oracle_get_eri_spin_blocks_cation(cation):
    """get_eri's lower-case letters are the beta orbitals (IntegralTransformation.jl:52-68):
    the alpha, beta and mixed-spin blocks of the UHF cation give the UMP2 known answer, and
    taking alpha orbitals for the lower-case letters does not."""
    w = cation
    Cao, Cav, Cbo, Cbv = _split(w)
    blk = lambda s: O.get_eri(w["g"], Cao, Cav, s, Cbo=Cbo, Cbv=Cbv)  # noqa: E731
    assert blk("OoVv").shape == (w["na"], w["nb"], Cav.shape[1], Cbv.shape[1])
    E = _ump2_from_get_eri_blocks(blk, w["epsa"], w["epsb"], w["na"], w["nb"])
    assert abs(E - E_UMP2_CATION) < 1e-12
    # alpha orbitals standing in for the beta ones give a different number
    wrong = lambda s: O.get_eri(w["g"], Cao, Cav, s.upper())  # noqa: E731
    Ew = _ump2_from_get_eri_blocks(wrong, w["epsa"], w["epsa"], w["na"], w["na"])
    assert abs(Ew - E_UMP2_CATION) > 1e-4
    # fcn drops the first occupied orbitals of either spin (:53,57)
    f = O.get_eri(w["g"], Cao, Cav, "OoVv", fcn=1, Cbo=Cbo, Cbv=Cbv)
    assert np.allclose(f, blk("OoVv")[1:, 1:])


def test_oracle_ump2_closed_shell_is_rmp2(water):
    w = water
    o = w["nocc"]
    Co, Cv = w["C"][:, :o], w["C"][:, o:]
    E_u = O.ump2(w["g"], Co, Cv, Co, Cv, w["eps"], w["eps"])
    E_r = O.rmp2(O.get_eri(w["g"], Co, Cv, "OOVV"), w["eps"], o)
    assert abs(E_u - E_r) < 1e-14


def test_oracle_get_fock_canonical(water, cation):
    """Canonical SCF orbitals diagonalise their own Fock matrix with the orbital energies."""
    w = water
    o = w["nocc"]
    f = O.get_fock(w["g"], w["H"], w["C"], w["C"][:, :o])
    assert np.abs(f - np.diag(w["eps"])).max() < 1e-10
    # the UHF alpha Fock matrix needs both spins' densities: get_fock's closed-shell
    # form 2J[Da] - K[Da] is not it, so it must differ from diag(epsa)
    c = cation
    fa = O.get_fock(c["g"], c["H"], c["Ca"], c["Ca"][:, :c["na"]])
    assert np.abs(fa - np.diag(c["epsa"])).max() > 1e-3


# ---------------------------------------------------------------- GPU


@pytest.mark.gpu
def test_gpu_ump2_cation_known_answer(cation):
    from jues import MollerPlesset
    from jues.Wavefunction import Wfn
    w = cation
    wf = Wfn.from_uhf(w["Ca"], w["Cb"], w["epsa"], w["epsb"], w["na"], w["nb"], np.asfortranarray(w["g"]))
    assert abs(MollerPlesset.do_ump2(wf) - E_UMP2_CATION) < 1e-12
    assert abs(MollerPlesset.do_ump2(wf, prec="fp32_mixed") - E_UMP2_CATION) < 1e-6 * abs(E_UMP2_CATION)


@pytest.mark.gpu
def test_gpu_get_eri_spin_blocks_cation(cation):
    """Mixed-spin get_eri blocks of the UHF cation against the oracle; the device blocks
    reproduce the UMP2 known answer."""
    from jues import IntegralTransformation as IT
    from jues.Wavefunction import Wfn
    w = cation
    Cao, Cav, Cbo, Cbv = _split(w)
    wf = Wfn.from_uhf(w["Ca"], w["Cb"], w["epsa"], w["epsb"], w["na"], w["nb"], np.asfortranarray(w["g"]))
    for s in ("OoVv", "oOvV", "oovv", "OOVV", "OvOv", "ovOV", "vVoO"):
        for note in ("phys", "chem"):
            ref = O.get_eri(w["g"], Cao, Cav, s, notation=note, Cbo=Cbo, Cbv=Cbv)
            got = IT.get_eri(wf, s, notation=note)
            assert got.shape == ref.shape, (s, note)
            assert np.abs(got - ref).max() < 1e-13, (s, note)
    ref = O.get_eri(w["g"], Cao, Cav, "OoVv", fcn=1, Cbo=Cbo, Cbv=Cbv)
    assert np.abs(IT.get_eri(wf, "OoVv", fcn=1) - ref).max() < 1e-13
    E = _ump2_from_get_eri_blocks(lambda s: IT.get_eri(wf, s), w["epsa"], w["epsb"], w["na"], w["nb"])
    assert abs(E - E_UMP2_CATION) < 1e-12


@pytest.mark.gpu
def test_gpu_get_eri_spin_blocks_random_factors():
    """Beta blocks of a different size than the alpha ones, through the factor route."""
    from jues import IntegralTransformation as IT
    from jues.Wavefunction import Wfn, FactoredERI
    s = _random_uhf()
    na, nb = s["na"], s["nb"]
    Cao, Cav, Cbo, Cbv = s["Ca"][:, :na], s["Ca"][:, na:], s["Cb"][:, :nb], s["Cb"][:, nb:]
    wf = Wfn.from_uhf(s["Ca"], s["Cb"], s["epsa"], s["epsb"], na, nb, FactoredERI(s["B"]))
    for blk in ("OoVv", "ooVV", "vOoV", "oooo"):
        ref = O.get_eri(s["g"], Cao, Cav, blk, Cbo=Cbo, Cbv=Cbv)
        for prec, tol in (("fp64", 1e-12), ("fp32_mixed", 1e-5)):
            got = IT.get_eri(wf, blk, prec=prec)
            assert got.shape == ref.shape
            assert np.abs(got - ref).max() < tol * np.abs(ref).max(), (blk, prec)


@pytest.mark.gpu
@pytest.mark.parametrize("dense", [True, False])
def test_gpu_ump2_random(dense):
    from jues import MollerPlesset
    from jues.Wavefunction import Wfn, FactoredERI
    s = _random_uhf()
    na, nb = s["na"], s["nb"]
    ref = O.ump2(s["g"], s["Ca"][:, :na], s["Ca"][:, na:], s["Cb"][:, :nb], s["Cb"][:, nb:], s["epsa"], s["epsb"])
    eri = np.asfortranarray(s["g"]) if dense else FactoredERI(s["B"])
    wf = Wfn.from_uhf(s["Ca"], s["Cb"], s["epsa"], s["epsb"], na, nb, eri)
    assert abs(MollerPlesset.do_ump2(wf) - ref) < 1e-12 * abs(ref)


@pytest.mark.gpu
def test_gpu_ump2_closed_shell_is_rmp2(cfg1):
    from jues import MollerPlesset
    from jues.Wavefunction import Wfn
    w = Wfn.from_synthetic(cfg1, dense=False)
    assert abs(MollerPlesset.do_ump2(w) - MollerPlesset.do_rmp2(w)) < 1e-10 * abs(MollerPlesset.do_rmp2(w))


@pytest.mark.gpu
@pytest.mark.parametrize("dense", [True, False])
def test_gpu_get_fock(dense):
    from jues import IntegralTransformation as IT
    from jues.Wavefunction import Wfn, FactoredERI
    s = _random_uhf()
    rng = np.random.default_rng(4)
    n = s["Ca"].shape[0]
    h = rng.standard_normal((n, n))
    h = (h + h.T) / 2
    eri = np.asfortranarray(s["g"]) if dense else FactoredERI(s["B"])
    wf = Wfn.from_uhf(s["Ca"], s["Cb"], s["epsa"], s["epsb"], s["na"], s["nb"], eri, hao=h)
    for spin, C, no in (("alpha", s["Ca"], s["na"]), ("b", s["Cb"], s["nb"])):
        ref = O.get_fock(s["g"], h, C, C[:, :no])
        got = IT.get_fock(wf, spin=spin)
        assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    with pytest.raises(ValueError):
        IT.get_fock(wf, spin="gamma")


@pytest.mark.gpu
def test_gpu_get_fock_canonical_water(water):
    from jues import IntegralTransformation as IT
    from jues.Wavefunction import Wfn
    w = water
    wf = Wfn.from_orbitals(w["C"], w["eps"], w["nocc"], np.asfortranarray(w["g"]), hao=w["H"])
    assert np.abs(IT.get_fock(wf) - np.diag(w["eps"])).max() < 1e-10
