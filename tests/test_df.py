"""Density fitting (Backend/DF.jl, MollerPlesset/DF-RMP2.jl, CoupledCluster/DF-RCCD.jl).

The reference's own DF known answer (DF-RMP2 -0.04913505451294127,
test/TestMollerPlesset.jl:39) is for psi4's DF-SCF water in STO-3G with the
def2-SVP-RI auxiliary basis (d and f functions, spherical) and def2-universal-JKFIT
orbitals; neither basis nor psi4 is available, so these inputs are synthetic:
a seeded symmetric positive definite metric J and pqP = b0 . J^(1/2), so that
make_bmunu must give back b0 (a closed form independent of the code under test).

CPU: the oracle's DF restatements against the closed forms and the non-DF
restatements.  GPU: jues.DF / do_df_rmp2 / DFRCCD.do_df_rccd against the oracle.
"""
import pathlib

import numpy as np
import pytest

import jues_oracle as O

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
LAM = 0.2  # cfg1 coupling at which 40 plain Jacobi sweeps stay bounded (not yet converged)


def _metric(naux, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((naux, naux)))
    w = np.linspace(0.3, 3.0, naux)
    J = (Q * w) @ Q.T
    Jh = (Q * np.sqrt(w)) @ Q.T
    return (J + J.T) / 2, (Jh + Jh.T) / 2


def df_inputs(b0, seed=11):
    """(pqP, Jpqh) with pqP . Jpqh = b0 in exact arithmetic."""
    naux = b0.shape[2]
    J, Jh = _metric(naux, seed)
    pqP = np.asfortranarray(np.tensordot(b0, Jh, axes=(2, 0)))
    return pqP, np.asfortranarray(O.metric_inv_sqrt(J))


@pytest.fixture(scope="module")
def cfg1_df(cfg1):
    b0 = cfg1.B * np.sqrt(LAM)
    pqP, Jpqh = df_inputs(b0)
    return dict(s=cfg1, b0=b0, pqP=pqP, Jpqh=Jpqh, Co=cfg1.C[:, :cfg1.o], Cv=cfg1.C[:, cfg1.o:])


@pytest.fixture(scope="module")
def cfg1_df_ref(cfg1_df):
    d = cfg1_df
    E, T2, hist = O.df_rccd(O.make_bmunu(d["pqP"], d["Jpqh"]), d["Co"], d["Cv"], d["s"].eps, maxit=40)
    return dict(E=E, T2=T2, hist=hist)


# ---------------------------------------------------------------- oracle (CPU)


def test_metric_inv_sqrt_and_bmunu(cfg1_df):
    d = cfg1_df
    J, _ = _metric(d["b0"].shape[2], 11)
    Jh = O.metric_inv_sqrt(J)
    assert np.abs(Jh @ J @ Jh - np.eye(J.shape[0])).max() < 1e-12
    b = O.make_bmunu(d["pqP"], d["Jpqh"])
    assert np.abs(b - d["b0"]).max() < 1e-12 * np.abs(d["b0"]).max()
    # (mu nu|lam sig) = sum_PQ (mu nu|P) J^-1[P,Q] (Q|lam sig), the DF approximation itself
    n = b.shape[0]
    lhs = b.reshape(n * n, -1) @ b.reshape(n * n, -1).T
    t = d["pqP"].reshape(n * n, -1)
    rhs = t @ np.linalg.solve(J, t.T)
    assert np.abs(lhs - rhs).max() < 1e-12 * np.abs(rhs).max()


def test_oracle_df_rmp2_is_rmp2_of_the_fitted_integrals(cfg1_df):
    d = cfg1_df
    b = d["b0"]
    g = np.einsum("mnP,lsP->mnls", b, b)
    E_ref = O.rmp2(O.get_eri(g, d["Co"], d["Cv"], "OOVV"), d["s"].eps, d["s"].o)
    assert abs(O.df_rmp2(b, d["Co"], d["Cv"], d["s"].eps) - E_ref) < 1e-12 * abs(E_ref)


def test_oracle_df_rccd_sweep_is_rccd_step(cfg1_df):
    """DF-RCCD.jl's intermediates built from b reproduce RCCD.jl's update on the
    same integrals; the term of DF-RCCD.jl:250 as written does not."""
    d = cfg1_df
    s = d["s"]
    rng = np.random.default_rng(5)
    I = O.rccd_integrals_from_factors(d["b0"], d["Co"], d["Cv"])
    b = O.df_rccd_integrals(d["b0"], d["Co"], d["Cv"])
    D = O.form_Dijab(s.eps, s.o, s.v)
    t = rng.standard_normal((s.o, s.o, s.v, s.v)) * 0.01
    t = (t + t.transpose(1, 0, 3, 2)) / 2
    ref = O.rccd_step(t, I, D)
    assert np.abs(O.df_rccd_sweep(t, b, D) - ref).max() < 1e-12 * np.abs(ref).max()
    assert np.abs(O.df_rccd_sweep(t, b, D, literal=True) - ref).max() > 1e-6 * np.abs(ref).max()


def test_oracle_df_rccd_guess_and_fixed_point(cfg1_df, cfg1_df_ref):
    d = cfg1_df
    s = d["s"]
    E0, _, _ = O.df_rccd(d["b0"], d["Co"], d["Cv"], s.eps, maxit=0)
    assert abs(E0 - O.df_rmp2(d["b0"], d["Co"], d["Cv"], s.eps)) < 1e-12 * abs(E0)   # T2_init! = MP2
    I = O.rccd_integrals_from_factors(d["b0"], d["Co"], d["Cv"])
    E_fix, _, _ = O.solve_rccd(I, s.eps, s.o, guess="mp2", e_conv=1e-14, r_conv=1e-13, maxit=300)
    assert abs(cfg1_df_ref["E"] - E_fix) < 1e-6                                      # 40 sweeps: close
    assert abs(cfg1_df_ref["hist"][-1] - cfg1_df_ref["hist"][-2]) > 1e-10             # not yet converged


# ---------------------------------------------------------------- GPU


def _wfn(d, df=True):
    from jues.Wavefunction import Wfn, DFIntegrals, FactoredERI
    eri = DFIntegrals(d["pqP"], d["Jpqh"]) if df else FactoredERI(np.asfortranarray(d["b0"]))
    return Wfn.from_orbitals(d["s"].C, d["s"].eps, d["s"].o, eri)


@pytest.mark.gpu
def test_gpu_make_bmunu(cfg1_df):
    from jues import DF
    d = cfg1_df
    b = DF.make_bμν(d["pqP"], d["Jpqh"])
    ref = O.make_bmunu(d["pqP"], d["Jpqh"])
    assert np.abs(b - ref).max() < 1e-13 * np.abs(ref).max()


@pytest.mark.gpu
def test_gpu_df_rmp2(cfg1_df):
    from jues import MollerPlesset
    d = cfg1_df
    ref = O.df_rmp2(O.make_bmunu(d["pqP"], d["Jpqh"]), d["Co"], d["Cv"], d["s"].eps)
    assert abs(MollerPlesset.do_df_rmp2(_wfn(d)) - ref) < 1e-12 * abs(ref)
    assert abs(MollerPlesset.do_df_rmp2(_wfn(d), prec="fp32_mixed") - ref) < 1e-6 * abs(ref)


@pytest.mark.gpu
@pytest.mark.parametrize("ladder", ["recompute", "stored"])
def test_gpu_df_rccd_40_sweeps(cfg1_df, cfg1_df_ref, ladder):
    from jues import CoupledCluster as CC
    hist = []
    E, T2 = CC.DFRCCD.do_df_rccd(_wfn(cfg1_df), return_T2=True, ladder=ladder, callback=hist.append)
    assert len(hist) == 40
    assert abs(E - cfg1_df_ref["E"]) < 1e-12
    assert np.abs(np.array([h["energy"] for h in hist]) - cfg1_df_ref["hist"]).max() < 1e-12
    assert np.abs(T2 - cfg1_df_ref["T2"]).max() < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("fp32_mixed", 1e-6), ("fp16x3_mixed", 2e-5)])
def test_gpu_df_rccd_fast_forms(cfg1_df, cfg1_df_ref, prec, tol):
    """The same 40 sweeps with FP32-MFMA or split-FP16 contractions: relative energy error at
    the contraction arithmetic's level, accumulated over the sweeps."""
    from jues import CoupledCluster as CC
    hist = []
    E = CC.DFRCCD.do_df_rccd(_wfn(cfg1_df), prec=prec, callback=hist.append)
    assert len(hist) == 40 and all(h["stage"] != 0 for h in hist)
    assert abs(E - cfg1_df_ref["E"]) < tol * abs(cfg1_df_ref["E"])


@pytest.mark.gpu
def test_gpu_df_rccd_maxit_and_factored_input(cfg1_df):
    from jues import CoupledCluster as CC
    d = cfg1_df
    E5, _, _ = O.df_rccd(d["b0"], d["Co"], d["Cv"], d["s"].eps, maxit=5)
    assert abs(CC.DFRCCD.do_df_rccd(_wfn(d, df=False), maxit=5) - E5) < 1e-12


@pytest.mark.gpu
def test_gpu_df_errors(cfg1_df, cfg1):
    from jues import CoupledCluster as CC, MollerPlesset, JuesError
    from jues.Wavefunction import Wfn
    dense = Wfn.from_synthetic(cfg1, dense=True)
    with pytest.raises(TypeError):
        MollerPlesset.do_df_rmp2(dense)
    with pytest.raises(TypeError):
        CC.DFRCCD.do_df_rccd(dense)
    with pytest.raises(ValueError):
        CC.DFRCCD.do_df_rccd(_wfn(cfg1_df), dfbname="cc-pvdz-ri")
    from jues import _lib
    ctx = _lib.context(0)
    with pytest.raises(JuesError):
        ctx.get_ao_factors(3, 3)


@pytest.mark.gpu
def test_gpu_df_rmp2_cfg2_golden():
    """cfg2 (n=120, naux=360) at coupling 0.02 through (pqP, Jpqh): the golden MP2 energy."""
    from jues import MollerPlesset, synthetic
    from jues.Wavefunction import Wfn, DFIntegrals
    g2 = np.load(GOLD / "cfg2_coupling0.02.npz")
    s = synthetic.workload("cfg2_rccsd_n120")
    pqP, Jpqh = df_inputs(s.B * np.sqrt(float(g2["coupling"])), seed=12)
    E = MollerPlesset.do_df_rmp2(Wfn.from_orbitals(s.C, s.eps, s.o, DFIntegrals(pqP, Jpqh)))
    assert abs(E - float(g2["E_mp2"])) < 1e-10 * abs(float(g2["E_mp2"]))
