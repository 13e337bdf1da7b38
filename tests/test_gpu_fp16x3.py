"""Split-FP16 fast stage (JUES_FP16X3_MIXED).

A converged result is always the FP64 fixed point (the FP64 refinement decides
it), so those checks alone would not show that the split ring and ladder
products compute the right thing.  The per-iteration tests therefore take the
same plain-Jacobi steps from the MP2 guess in FP64, FP32 and split-FP16 and
compare the amplitudes: a wrong or missing term shows at O(1), while split
rounding (2^-21 relative per product) stays near the FP32 level.
"""
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
TIGHT = dict(cc_e_conv=1e-12, cc_max_rms=1e-11, maxit=300)


def _scaled_cfg2(lam=0.02):
    from jues import synthetic
    from jues.Wavefunction import Wfn, FactoredERI
    s = synthetic.workload("cfg2_rccsd_n120")
    return Wfn.from_orbitals(s.C, s.eps, s.o, FactoredERI(np.asfortranarray(s.B * np.sqrt(lam))))


def _steps(w, method, nsteps, prec):
    from jues.CoupledCluster import Solver
    s = Solver(w, method=method, prec=prec, guess="mp2", diis=False)
    info, _ = s.iterate(nsteps, prec)
    return info, s.result()


@pytest.mark.parametrize("which", ["cfg1", "cfg2"])
@pytest.mark.parametrize("method,nsteps", [("rccd", 1), ("rccd", 3), ("rccsd", 2)])
def test_split_steps_track_fp64(cfg1, which, method, nsteps):
    from jues.Wavefunction import Wfn
    w = Wfn.from_synthetic(cfg1, dense=False) if which == "cfg1" else _scaled_cfg2()
    i64, r64 = _steps(w, method, nsteps, "fp64")
    i32, r32 = _steps(w, method, nsteps, "fp32_mixed")
    i16, r16 = _steps(w, method, nsteps, "fp16x3")
    assert (i64["stage"], i32["stage"], i16["stage"]) == (0, 1, 2)
    T64, T32, T16 = r64[-1], r32[-1], r16[-1]
    scale = np.max(np.abs(T64))
    e32 = np.max(np.abs(T32 - T64)) / scale
    e16 = np.max(np.abs(T16 - T64)) / scale
    assert e16 < 2e-5, (e16, e32)
    assert e16 < 20 * e32 + 1e-12, (e16, e32)
    assert abs(i16["energy"] - i64["energy"]) < 2e-5 * abs(i64["energy"])
    if method == "rccsd":
        t1s = np.max(np.abs(r64[1]))
        assert t1s > 0 and np.max(np.abs(r16[1] - r64[1])) / t1s < 2e-5


def test_cfg1_rccd_converges_to_fp64(cfg1):
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn
    g = np.load(GOLD / "cfg1.npz")
    hist = []
    E, T2 = CC.mRCCD.do_rccd(Wfn.from_synthetic(cfg1, dense=False), prec="fp16x3", return_T2=True,
                             callback=hist.append, **TIGHT)
    stages = [h["stage"] for h in hist]
    assert 2 in stages and stages[-1] == 0
    assert abs(E - g["E_rccd"]) < 1e-10
    assert np.max(np.abs(T2 - g["T2_rccd"])) < 1e-9


@pytest.mark.parametrize("eqs", ["mRCCSD", "RCCSD"])
def test_cfg1_rccsd_converges_to_fp64(cfg1, eqs):
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn
    g = np.load(GOLD / "cfg1.npz")
    key = eqs.lower()
    mod = CC.mRCCSD if eqs == "mRCCSD" else CC.RCCSD
    E, T1, T2 = mod.do_rccsd(Wfn.from_synthetic(cfg1, dense=False), prec="fp16x3", return_T=True, **TIGHT)
    assert abs(E - g[f"E_{key}"]) < 1e-10
    assert np.max(np.abs(T1 - g[f"T1_{key}"])) < 1e-9
    assert np.max(np.abs(T2 - g[f"T2_{key}"])) < 1e-9


@pytest.mark.parametrize("method", ["rccd", "mrccsd"])
def test_cfg2_coupling_golden(method):
    from jues import CoupledCluster as CC, synthetic
    from jues.Wavefunction import Wfn, FactoredERI
    g = np.load(GOLD / "cfg2_coupling0.02.npz")
    s = synthetic.workload("cfg2_rccsd_n120")
    w = Wfn.from_orbitals(s.C, s.eps, s.o, FactoredERI(s.B))
    kw = dict(prec="fp16x3", coupling=float(g["coupling"]), diis_max=8, **TIGHT)
    if method == "rccd":
        E = CC.mRCCD.do_rccd(w, **kw)
        assert abs(E - g["E_rccd"]) < 1e-10
    else:
        E, T1, _ = CC.mRCCSD.do_rccsd(w, return_T=True, **kw)
        assert abs(E - g["E_mrccsd"]) < 1e-10
        assert np.max(np.abs(T1 - g["T1_mrccsd"])) < 1e-9


def test_water_known_answers_fp16x3():
    import sto3g
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn
    w = sto3g.system(sto3g.zmat_water())
    wf = Wfn.from_orbitals(w["C"], w["eps"], w["nocc"], np.asfortranarray(w["g"]))
    assert abs(CC.RCCD.do_rccd(wf, prec="fp16x3", **TIGHT) - (-0.07015050066089029)) < 1e-10
    assert abs(CC.mRCCSD.do_rccsd(wf, prec="fp16x3", **TIGHT) - (-0.070680102078571)) < 1e-10
