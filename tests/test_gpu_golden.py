"""The library against the committed golden fixtures (tests/golden/make_golden.py).

cfg1 (n=24, o=5, v=19): RMP2, the first plain-Jacobi energies, converged RCCD
and RCCSD (both equation sets) amplitudes, and the (ia|jb) checksum.
cfg2 (n=120, o=20, v=100, naux=360) at coupling 0.02: RCCD and mRCCSD energies,
amplitude norms and sampled entries, in FP64 and FP32-mixed; reached two ways,
with the factors scaled by sqrt(0.02) (the drop-in call) and with the solver's
``coupling`` option on the unscaled factors (every residual term carries one
integral, so the fixed points coincide).
"""
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
TIGHT = dict(cc_e_conv=1e-12, cc_max_rms=1e-11, maxit=300)


@pytest.fixture(scope="module")
def g1():
    return dict(np.load(GOLD / "cfg1.npz"))


@pytest.fixture(scope="module")
def g2():
    return dict(np.load(GOLD / "cfg2_coupling0.02.npz"))


@pytest.fixture(scope="module")
def cfg2():
    from jues import synthetic
    return synthetic.workload("cfg2_rccsd_n120")


def _wfn_scaled(s, lam):
    from jues.Wavefunction import Wfn, FactoredERI
    return Wfn.from_orbitals(s.C, s.eps, s.o, FactoredERI(np.asfortranarray(s.B * np.sqrt(lam))))


def _wfn_factors(s):
    from jues.Wavefunction import Wfn, FactoredERI
    return Wfn.from_orbitals(s.C, s.eps, s.o, FactoredERI(s.B))


# ---------------------------------------------------------------- cfg1


@pytest.mark.parametrize("dense", [True, False])
def test_cfg1_mp2_and_checksum(cfg1, g1, dense):
    from jues import MollerPlesset
    from jues.IntegralTransformation import get_eri
    from jues.Wavefunction import Wfn
    w = Wfn.from_synthetic(cfg1, dense=dense)
    assert abs(MollerPlesset.do_rmp2(w) - g1["E_mp2"]) < 1e-10
    K = get_eri(w, "OVOV", notation="chem")
    assert abs(K.sum() - g1["iajb_sum"]) < 1e-10 * max(1.0, abs(g1["iajb_sum"]))
    assert abs((K * K).sum() - g1["iajb_sumsq"]) < 1e-10 * g1["iajb_sumsq"]


def test_cfg1_jacobi_energies(cfg1, g1):
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn
    hist = []
    CC.mRCCD.do_rccd(Wfn.from_synthetic(cfg1), maxit=3, diis=False, guess="zero", check_convergence=False,
                     callback=hist.append)
    got = np.array([h["energy"] for h in hist])
    assert np.max(np.abs(got - g1["jacobi_E"])) < 1e-11


@pytest.mark.parametrize("prec", ["fp64", "fp32_mixed"])
def test_cfg1_rccd(cfg1, g1, prec):
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn
    E, T2 = CC.mRCCD.do_rccd(Wfn.from_synthetic(cfg1), prec=prec, return_T2=True, **TIGHT)
    assert abs(E - g1["E_rccd"]) < 1e-10
    assert np.max(np.abs(T2 - g1["T2_rccd"])) < 1e-9


@pytest.mark.parametrize("eqs", ["mRCCSD", "RCCSD"])
def test_cfg1_rccsd(cfg1, g1, eqs):
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn
    key = eqs.lower()
    mod = CC.mRCCSD if eqs == "mRCCSD" else CC.RCCSD
    E, T1, T2 = mod.do_rccsd(Wfn.from_synthetic(cfg1, dense=False), return_T=True, **TIGHT)
    assert abs(E - g1[f"E_{key}"]) < 1e-10
    assert np.max(np.abs(T1 - g1[f"T1_{key}"])) < 1e-9
    assert np.max(np.abs(T2 - g1[f"T2_{key}"])) < 1e-9


# ---------------------------------------------------------------- cfg2 (coupling 0.02)


def test_cfg2_mp2(cfg2, g2):
    from jues import MollerPlesset
    E = MollerPlesset.do_rmp2(_wfn_scaled(cfg2, float(g2["coupling"])))
    assert abs(E - g2["E_mp2"]) < 1e-10 * max(1.0, abs(g2["E_mp2"]))


def _check_t2(T2, g2, key):
    assert abs(np.linalg.norm(T2) - g2[f"T2_{key}_norm"]) < 1e-9 * max(1.0, g2[f"T2_{key}_norm"])
    got = T2.ravel(order="F")[g2["sample_idx"]]
    assert np.max(np.abs(got - g2[f"T2_{key}_sample"])) < 1e-9


@pytest.mark.parametrize("via", ["scaled", "coupling"])
@pytest.mark.parametrize("prec", ["fp64", "fp32_mixed"])
def test_cfg2_rccd(cfg2, g2, prec, via):
    from jues import CoupledCluster as CC
    lam = float(g2["coupling"])
    if via == "scaled":
        w, kw = _wfn_scaled(cfg2, lam), {}
    else:
        w, kw = _wfn_factors(cfg2), dict(coupling=lam)
    E, T2 = CC.mRCCD.do_rccd(w, prec=prec, return_T2=True, diis_max=8, **kw, **TIGHT)
    assert abs(E - g2["E_rccd"]) < 1e-10
    _check_t2(T2, g2, "rccd")


@pytest.mark.parametrize("via", ["scaled", "coupling"])
@pytest.mark.parametrize("prec", ["fp64", "fp32_mixed"])
def test_cfg2_mrccsd(cfg2, g2, prec, via):
    from jues import CoupledCluster as CC
    lam = float(g2["coupling"])
    if via == "scaled":
        w, kw = _wfn_scaled(cfg2, lam), {}
    else:
        w, kw = _wfn_factors(cfg2), dict(coupling=lam)
    E, T1, T2 = CC.mRCCSD.do_rccsd(w, prec=prec, return_T=True, diis_max=8, **kw, **TIGHT)
    assert abs(E - g2["E_mrccsd"]) < 1e-10
    assert np.max(np.abs(T1 - g2["T1_mrccsd"])) < 1e-9
    _check_t2(T2, g2, "mrccsd")
