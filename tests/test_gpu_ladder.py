"""Recomputed particle-particle ladder (jues_cc_opts.ladder = RECOMPUTE): <ab|ef> rebuilt
from the 3-index factors every iteration in slabs of f, never stored.  It must give the
stored ladder's results: single Jacobi steps agree to rounding, converged energies and
amplitudes match the golden fixtures.  JUES_LADDER_SLAB_BYTES forces many slabs so that the
slab bookkeeping (row offsets of the packed pairs, the first minus slab) is exercised.
"""
import os
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
TIGHT = dict(cc_e_conv=1e-12, cc_max_rms=1e-11, maxit=300)


@pytest.fixture()
def small_slabs(monkeypatch):
    # cfg1 (v = 19): 20 KB slabs hold 1-2 values of f in FP64, 3 in FP32
    monkeypatch.setenv("JUES_LADDER_SLAB_BYTES", "20000")
    yield


def _wfn(s):
    from jues.Wavefunction import Wfn, FactoredERI
    return Wfn.from_orbitals(s.C, s.eps, s.o, FactoredERI(s.B))


def _steps(w, method, nsteps, prec, ladder):
    from jues.CoupledCluster import Solver
    s = Solver(w, method=method, prec=prec, guess="mp2", diis=False, ladder=ladder)
    s.iterate(nsteps, prec)
    return s.result()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32_mixed", 2e-6), ("fp16x3", 2e-5)])
@pytest.mark.parametrize("method,nsteps", [("rccd", 2), ("rccsd", 2)])
def test_recompute_steps_match_stored(cfg1, small_slabs, prec, tol, method, nsteps):
    w = _wfn(cfg1)
    ref = _steps(w, method, nsteps, "fp64", "stored")
    got = _steps(w, method, nsteps, prec, "recompute")
    scale = np.max(np.abs(ref[-1]))
    assert np.max(np.abs(got[-1] - ref[-1])) < tol * scale
    assert abs(got[0] - ref[0]) < tol * abs(ref[0])


@pytest.mark.parametrize("prec", ["fp64", "fp32_mixed", "fp16x3"])
def test_recompute_converges_to_golden(cfg1, small_slabs, prec):
    from jues import CoupledCluster as CC
    g = np.load(GOLD / "cfg1.npz")
    E, T2 = CC.mRCCD.do_rccd(_wfn(cfg1), prec=prec, ladder="recompute", return_T2=True, **TIGHT)
    assert abs(E - g["E_rccd"]) < 1e-10
    assert np.max(np.abs(T2 - g["T2_rccd"])) < 1e-9
    E, T1, T2 = CC.mRCCSD.do_rccsd(_wfn(cfg1), prec=prec, ladder="recompute", return_T=True, **TIGHT)
    assert abs(E - g["E_mrccsd"]) < 1e-10
    assert np.max(np.abs(T1 - g["T1_mrccsd"])) < 1e-9


def test_recompute_cfg2_coupling_golden():
    """Default slab size (one slab at v = 100) at cfg2 scale."""
    from jues import CoupledCluster as CC, synthetic
    g = np.load(GOLD / "cfg2_coupling0.02.npz")
    s = synthetic.workload("cfg2_rccsd_n120")
    E = CC.mRCCD.do_rccd(_wfn(s), prec="fp16x3", ladder="recompute", coupling=float(g["coupling"]), diis_max=8,
                         **TIGHT)
    assert abs(E - g["E_rccd"]) < 1e-10


def test_recompute_needs_factors(cfg1):
    from jues import CoupledCluster as CC, _lib
    from jues.Wavefunction import Wfn
    with pytest.raises(_lib.JuesError, match="factored"):
        CC.mRCCD.do_rccd(Wfn.from_synthetic(cfg1, dense=True), ladder="recompute")
