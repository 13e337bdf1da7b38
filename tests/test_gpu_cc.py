"""RCCD / RCCSD on the GPU against the CPU oracle (converged FP64 fixed points)."""
import numpy as np
import pytest

import jues_oracle as O

pytestmark = pytest.mark.gpu


def _sys_small():
    from jues import synthetic
    return synthetic.make_system(12, 3, 9, 36, 7, name="tiny")


@pytest.fixture(scope="module")
def ref_rccd(cfg1):
    s = cfg1
    I = O.rccd_integrals_from_factors(s.B, s.Co, s.Cv)
    E, T, it = O.solve_rccd(I, s.eps, s.o, maxit=300)
    return E, T


@pytest.mark.parametrize("dense", [True, False])
def test_rccd_fp64_cfg1(cfg1, ref_rccd, dense):
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn
    Eref, Tref = ref_rccd
    w = Wfn.from_synthetic(cfg1, dense=dense)
    E, T2 = CC.mRCCD.do_rccd(w, return_T2=True, cc_e_conv=1e-12, cc_max_rms=1e-11, maxit=300)
    assert abs(E - Eref) < 1e-10
    assert np.max(np.abs(T2 - Tref)) < 1e-9
    assert np.max(np.abs(T2 - T2.transpose(1, 0, 3, 2))) < 1e-12


def test_rccd_jacobi_first_step_is_mp2(cfg1):
    from jues import CoupledCluster as CC, MollerPlesset
    from jues.Wavefunction import Wfn
    w = Wfn.from_synthetic(cfg1)
    hist = []
    CC.mRCCD.do_rccd(w, maxit=1, diis=False, guess="zero", check_convergence=False, callback=hist.append)
    assert abs(hist[0]["energy"] - MollerPlesset.do_rmp2(w)) < 1e-10


def test_rccd_guesses_agree(cfg1, ref_rccd):
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn
    w = Wfn.from_synthetic(cfg1, dense=False)
    for g in ("zero", "mp2", "ovov"):
        E = CC.RCCD.do_rccd(w, guess=g, cc_e_conv=1e-12, cc_max_rms=1e-11, maxit=300)
        assert abs(E - ref_rccd[0]) < 1e-10, g


def test_rccd_jacobi_weak_coupling():
    """Plain Jacobi (RCCD.jl's iteration) where it converges: scaled-down integrals."""
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn, FactoredERI
    s = _sys_small()
    B = 0.3 * s.B
    I = O.rccd_integrals_from_factors(B, s.Co, s.Cv)
    Eref, Tref, _ = O.solve_rccd(I, s.eps, s.o, diis=False, guess="ovov", maxit=500)
    w = Wfn.from_orbitals(s.C, s.eps, s.o, FactoredERI(np.asfortranarray(B)))
    E, T = CC.RCCD.do_rccd(w, diis=False, return_T2=True, cc_e_conv=1e-13, cc_max_rms=1e-12, maxit=500)
    assert abs(E - Eref) < 1e-10
    assert np.max(np.abs(T - Tref)) < 1e-9


def test_rccd_fp32_mixed(cfg1, ref_rccd):
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn
    w = Wfn.from_synthetic(cfg1, dense=False)
    hist = []
    E = CC.mRCCD.do_rccd(w, prec="fp32_mixed", cc_e_conv=1e-12, cc_max_rms=1e-11, maxit=300,
                         callback=hist.append)
    assert any(h["stage"] == 1 for h in hist) and hist[-1]["stage"] == 0
    assert abs(E - ref_rccd[0]) < 1e-10


@pytest.mark.parametrize("eqs", ["mRCCSD", "RCCSD"])
def test_rccsd_fp64_cfg1(cfg1, eqs):
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn
    s = cfg1
    if eqs == "mRCCSD":
        I = O.rccsd_integrals_from_factors(s.B, s.Co, s.Cv)
        upd = O.rccsd_update
    else:
        I = O.make_rccsd15_integrals(lambda a, b, c, d: O.mo_from_factors(s.B, a, b, c, d), s.Co, s.Cv)
        I["oovv"] = I["oovv"]
        upd = O.rccsd15_update
    Eref, T1r, T2r, _ = O.solve_rccsd(I, s.eps, s.o, update=upd, maxit=300)
    w = Wfn.from_synthetic(s, dense=False)
    mod = CC.mRCCSD if eqs == "mRCCSD" else CC.RCCSD
    E, T1, T2 = mod.do_rccsd(w, return_T=True, cc_e_conv=1e-12, cc_max_rms=1e-11, maxit=300)
    assert abs(E - Eref) < 1e-10
    assert np.max(np.abs(T1 - T1r)) < 1e-9
    assert np.max(np.abs(T2 - T2r)) < 1e-9


def test_rccsd_fp32_mixed(cfg1):
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn
    s = cfg1
    I = O.rccsd_integrals_from_factors(s.B, s.Co, s.Cv)
    Eref, _, _, _ = O.solve_rccsd(I, s.eps, s.o, maxit=300)
    w = Wfn.from_synthetic(s, dense=False)
    E = CC.mRCCSD.do_rccsd(w, prec="fp32_mixed", cc_e_conv=1e-12, cc_max_rms=1e-11, maxit=300)
    assert abs(E - Eref) < 1e-10


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32_mixed", 1e-5)])
def test_rccsd_steps_factored_batched_short_products(prec, tol):
    """Two Jacobi steps of factored RCCSD at n=96, o=32, v=64, naux=800 against the oracle's
    mRCCSD.jl update (:118-241).  The batched W2[i,f,P] = sum_(em) T[i,(em),f] LovT[(em),P] product
    (M = o = 32, K = ov = 2048, 64 batches of 800 columns) takes the skinny-M tile through the
    batched short-product route of the GEMM dispatch; the cfg1 and cfg2 systems are too small for it."""
    from jues import synthetic
    from jues.CoupledCluster import Solver
    from jues.Wavefunction import Wfn, FactoredERI
    o, v = 32, 64
    s = synthetic.make_system(o + v, o, v, 800, 11, name="w2")
    B = np.asfortranarray(0.3 * s.B)
    I = O.rccsd_integrals_from_factors(B, s.Co, s.Cv)
    Dia, Dijab = O.form_Dia(s.eps, o, v), O.form_Dijab(s.eps, o, v)
    t1, t2 = np.zeros((o, v)), I["oovv"] / Dijab
    sv = Solver(Wfn.from_orbitals(s.C, s.eps, o, FactoredERI(B)), method="rccsd", prec=prec, guess="mp2", diis=False)
    for _ in range(2):
        n1, n2 = O.rccsd_update(t1, t2, I)
        t1, t2 = n1 / Dia, n2 / Dijab
        sv.iterate(1, prec=prec)
        _, T1, T2 = sv.result()
        assert np.max(np.abs(T1 - t1)) < tol * np.max(np.abs(t1))
        assert np.max(np.abs(T2 - t2)) < tol * np.max(np.abs(t2))
