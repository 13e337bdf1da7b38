"""Defect-correction refinement of RCCD (jues_cc_opts.refine = DEFECT, jues_cc_refine) at
n=24, o=5, v=19.

An FP64 iteration at t0 keeps its numerator R0; the iterations after it solve
D t = R0 + [R(t) - R(t0)] with the bracket formed from d = t - t0 (RCCD.jl:242-256 residual).
* With an FP64 inner stage the bracket is exact, so anchor, one correction step, anchor
  reproduces three FP64 Jacobi steps (RCCD.jl:55-74 update) to rounding.
* Near the fixed point a correction step in FP32 or split-FP16 errs by the inner type's
  rounding times |t - t0| instead of times |t|: after the same three steps its distance from
  the FP64 result is below that of a plain fast step in the middle by about |t - t0| / |t|
  (at full coupling this system needs ~20 DIIS iterations before that ratio is small).
* Full solves with refine = DEFECT reach the oracle's FP64 fixed point with a handful of FP64
  iterations (this strongly coupled system converges slowly; the o=60, v=340 test holds the
  count to five), with stored and recomputed ladders and with dense or factored integrals.
"""
import numpy as np
import pytest

import jues_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def e_ccd(cfg1):
    s = cfg1
    I = O.rccd_integrals_from_factors(s.B, s.Co, s.Cv)
    E, _, _ = O.solve_rccd(I, s.eps, s.o, maxit=300)
    return E


def _wfn(cfg1, dense=False):
    from jues.Wavefunction import Wfn
    return Wfn.from_synthetic(cfg1, dense=dense)


def _solver(w, prec):
    from jues.CoupledCluster import Solver
    return Solver(w, method="rccd", prec=prec, guess="mp2", diis=True, diis_max=6)


def _three_steps(w, prefix, middle, setup_prec):
    """prefix FP64 DIIS iterations, then three Jacobi steps: FP64, `middle`, FP64."""
    sv = _solver(w, setup_prec)
    if prefix:
        sv.iterate(prefix, prec="fp64")
    for p in ("fp64", middle, "fp64"):
        sv.reset_diis()
        sv.iterate(1, prec=p)
    return sv.result()


def _anchor_step_anchor(w, prefix, inner):
    sv = _solver(w, inner)
    if prefix:
        sv.iterate(prefix, prec="fp64")
    info, conv = sv.refine(inner, max_anchors=2, inner_maxit=1)
    assert info["stage"] == 0 and info["n_fp64"] == 2 and not conv
    return sv.result()


@pytest.mark.parametrize("dense", [False, True])
@pytest.mark.parametrize("prefix", [0, 6])
def test_fp64_inner_reproduces_jacobi_steps(cfg1, dense, prefix):
    w = _wfn(cfg1, dense)
    E_ref, T_ref = _three_steps(w, prefix, "fp64", "fp64")
    E_dc, T_dc = _anchor_step_anchor(w, prefix, "fp64")
    scale = np.max(np.abs(T_ref))
    assert np.max(np.abs(T_dc - T_ref)) < 1e-12 * scale
    assert abs(E_dc - E_ref) < 1e-12 * abs(E_ref)


@pytest.mark.parametrize("inner", ["fp32_mixed", "fp16x3"])
def test_fast_correction_error_scales_with_d(cfg1, inner):
    w = _wfn(cfg1)
    prefix = 20
    # |t - t0| / |t| of the correction step: t0 after the prefix, t one Jacobi step later
    sv = _solver(w, inner)
    sv.iterate(prefix, prec="fp64")
    T0 = sv.result()[1]
    sv.reset_diis()
    sv.iterate(1, prec="fp64")
    rel_d = np.max(np.abs(sv.result()[1] - T0)) / np.max(np.abs(T0))
    _, T_ref = _three_steps(w, prefix, "fp64", inner)
    _, T_plain = _three_steps(w, prefix, inner, inner)
    _, T_dc = _anchor_step_anchor(w, prefix, inner)
    e_plain = np.max(np.abs(T_plain - T_ref))
    e_dc = np.max(np.abs(T_dc - T_ref))
    print(f"{inner}: |d|/|t| {rel_d:.2e}; plain fast step {e_plain:.3e}, correction step {e_dc:.3e}")
    assert e_plain > 0 and rel_d < 1e-2
    assert e_dc < 30 * rel_d * e_plain


@pytest.mark.parametrize("prec", ["fp32_mixed", "fp16x3"])
@pytest.mark.parametrize("ladder,dense", [("stored", False), ("recompute", False), ("stored", True)])
def test_defect_solve_reaches_fp64_fixed_point(cfg1, e_ccd, prec, ladder, dense):
    from jues import CoupledCluster as CC
    hist = []
    E = CC.mRCCD.do_rccd(_wfn(cfg1, dense), prec=prec, ladder=ladder, refine="defect", cc_e_conv=1e-11,
                         cc_max_rms=1e-11, maxit=300, callback=hist.append)
    n64 = sum(h["stage"] == 0 for h in hist)
    anchors = [f"{h['rnorm']:.1e}" for h in hist if h["stage"] == 0]
    print(f"{prec} {ladder} dense={dense}: |E - E_oracle| {abs(E - e_ccd):.2e}, {n64} FP64 of {len(hist)} "
          f"iterations; FP64 steps {anchors}")
    assert abs(E - e_ccd) < 1e-10
    assert hist[-1]["stage"] == 0 and 1 <= n64 <= 6


def test_rccsd_falls_back_to_restart(cfg1):
    """RCCSD has no correction form: refine = DEFECT runs the restart refinement (same result),
    and the stepwise entry point refuses it."""
    from jues import CoupledCluster as CC, _lib
    from jues.CoupledCluster import Solver
    w = _wfn(cfg1)
    kw = dict(prec="fp32_mixed", cc_e_conv=1e-12, cc_max_rms=1e-11, maxit=300)
    E_r = CC.mRCCSD.do_rccsd(w, **kw)
    E_d = CC.mRCCSD.do_rccsd(w, refine="defect", **kw)
    assert E_r == E_d
    sv = Solver(w, method="rccsd", prec="fp32_mixed")
    with pytest.raises(_lib.JuesError, match="RCCD"):
        sv.refine("fp32_mixed")


@pytest.mark.parametrize("prec", ["fp32_mixed", "fp16x3"])
def test_defect_solve_with_released_fp64_blocks(cfg1, e_ccd, prec, monkeypatch):
    """With the FP64 integral blocks released during every fast stage (JUES_THIN_FP64_INTS=1, what
    low device memory triggers at o=80, v=420), each FP64 anchor rebuilds them and each inner solve
    converts its FP32 copies again: the solve still reaches the oracle's fixed point, with the same
    energy as when nothing is released."""
    from jues import CoupledCluster as CC
    kw = dict(prec=prec, refine="defect", cc_e_conv=1e-11, cc_max_rms=1e-11, maxit=300)
    E_keep = CC.mRCCD.do_rccd(_wfn(cfg1), **kw)
    monkeypatch.setenv("JUES_THIN_FP64_INTS", "1")
    E_thin = CC.mRCCD.do_rccd(_wfn(cfg1), **kw)
    assert abs(E_thin - e_ccd) < 1e-10
    assert E_thin == E_keep


def test_refine_bad_arguments(cfg1):
    """A refine value other than RESTART / DEFECT, an unknown inner precision and zero anchors
    return JUES_BAD_ARG before any device work; the state stays usable."""
    from jues import _lib
    from jues import CoupledCluster as CC
    from jues.CoupledCluster import Solver
    w = _wfn(cfg1)
    with pytest.raises(_lib.JuesError) as e:
        CC.mRCCD.do_rccd(w, prec="fp32_mixed", refine=7)
    assert e.value.status == _lib.JUES_BAD_ARG
    sv = _solver(w, "fp32_mixed")
    for args in ((9, 2, 5), (_lib.FP32_MIXED, 0, 5), (_lib.FP32_MIXED, 2, -1)):
        info, conv, n64 = _lib.IterInfo(), _lib.C.c_int(0), _lib.C.c_int(0)
        st = sv.ctx.lib.jues_cc_refine(sv.ctx.h, *args, _lib.C.byref(info), _lib.C.byref(conv), _lib.C.byref(n64))
        assert st == _lib.JUES_BAD_ARG, args
    info, conv = sv.refine("fp32_mixed", max_anchors=1, inner_maxit=0)
    assert info["n_fp64"] == 1 and np.isfinite(info["energy"])


@pytest.mark.parametrize("extra", [dict(fcn=1), dict(coupling=0.8, continuation=2)])
def test_defect_solve_with_frozen_core_and_continuation(cfg1, extra):
    """Frozen core (get_eri's fcn) and coupling continuation (the defect correction runs at the last
    coupling only): the same fixed point as the plain FP64 solve with the same options."""
    from jues import CoupledCluster as CC
    kw = dict(cc_e_conv=1e-12, cc_max_rms=1e-11, maxit=300, **extra)
    E64 = CC.mRCCD.do_rccd(_wfn(cfg1), prec="fp64", **kw)
    for prec in ("fp32_mixed", "fp16x3"):
        E = CC.mRCCD.do_rccd(_wfn(cfg1), prec=prec, refine="defect", **kw)
        assert abs(E - E64) < 1e-10, (prec, extra, E, E64)
