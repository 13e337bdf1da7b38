"""Documented failure statuses and bitwise reproducibility, forced from the host.

Each failure below ends in the status include/jues.h documents for it, never in a
GPU fault: a zero denominator gives JUES_NONFINITE (IEEE infinities in the update
and energy kernels, caught by the host's finiteness check); an iteration limit with
the convergence check on gives JUES_NOT_CONVERGED (AutoRCCSD.jl:273-276 reports the
same condition); a device allocation larger than the GPU gives JUES_OOM before any
host data is read; wrong sizes or letters give JUES_BAD_ARG (the reference's
``error``, IntegralTransformation.jl:41-43).  The reductions have a fixed order,
so two identical solves in one process give identical bits.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _status(fn, *a, **k):
    from jues import _lib
    with pytest.raises(_lib.JuesError) as e:
        fn(*a, **k)
    return e.value.status


def _zero_gap_eps(s):
    """Orbital energies with e_0 = e_o, so D[0,0,0,0] = 2 e_0 - 2 e_o = 0."""
    eps = s.eps.copy()
    eps[s.o] = eps[0]
    return eps


def test_nonfinite_from_zero_denominator(cfg1):
    from jues import _lib, MollerPlesset
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn, FactoredERI
    s = cfg1
    w = Wfn.from_orbitals(s.C, _zero_gap_eps(s), s.o, FactoredERI(s.B))
    assert _status(MollerPlesset.do_rmp2, w) == _lib.JUES_NONFINITE
    assert _status(MollerPlesset.rmp2_grad, w, want_dg=False) == _lib.JUES_NONFINITE
    for prec in ("fp64", "fp32_mixed", "fp16x3"):
        assert _status(CC.mRCCD.do_rccd, w, guess="zero", prec=prec) == _lib.JUES_NONFINITE, prec
    assert _status(CC.mRCCSD.do_rccsd, w, guess="zero") == _lib.JUES_NONFINITE
    # the context stays usable after the error
    w_ok = Wfn.from_orbitals(s.C, s.eps, s.o, FactoredERI(s.B))
    assert np.isfinite(MollerPlesset.do_rmp2(w_ok))


def test_not_converged_at_iteration_limit(cfg1):
    from jues import _lib
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn
    w = Wfn.from_synthetic(cfg1, dense=False)
    assert _status(CC.mRCCD.do_rccd, w, maxit=1) == _lib.JUES_NOT_CONVERGED
    assert _status(CC.mRCCSD.do_rccsd, w, maxit=2) == _lib.JUES_NOT_CONVERGED
    # the refinement limit of a fast form counts as well
    assert _status(CC.mRCCD.do_rccd, w, prec="fp32_mixed", refine_maxit=1) == _lib.JUES_NOT_CONVERGED
    # without the check the last iterate's energy comes back
    E = CC.mRCCD.do_rccd(w, maxit=1, check_convergence=False)
    assert np.isfinite(E)


def test_full_strength_cfg2_returns_a_status():
    """The seeded cfg2 integrals at full strength have no convergent CCD iteration (E_MP2 is
    about -1.7e4 Eh): the solver must return a status, not a number."""
    from jues import _lib, synthetic
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn
    s = synthetic.workload("cfg2_rccsd_n120")
    w = Wfn.from_synthetic(s, dense=False)
    st = _status(CC.mRCCD.do_rccd, w, maxit=40)
    assert st in (_lib.JUES_NOT_CONVERGED, _lib.JUES_NONFINITE)


def test_oom_from_oversized_factors():
    """A factor array far larger than device memory: the allocation fails first (the small
    host buffer is never read) and the context keeps working."""
    from jues import _lib
    ctx = _lib.Context(0)
    try:
        B = np.zeros(16)
        st = ctx.lib.jues_set_ao_factors(ctx.h, _lib.ptr(B), 20000, 100000)   # 3.2e14 bytes
        assert st == _lib.JUES_OOM
        assert b"hipMalloc" in ctx.lib.jues_last_error(ctx.h)
        B = np.asfortranarray(np.ones((4, 4, 3)))
        assert ctx.lib.jues_set_ao_factors(ctx.h, _lib.ptr(B), 4, 3) == _lib.JUES_OK
    finally:
        ctx.close()


def test_bad_arguments(cfg1):
    from jues import _lib, IntegralTransformation as IT
    from jues.Wavefunction import Wfn
    s = cfg1
    w = Wfn.from_synthetic(s, dense=False)
    assert _status(IT.get_eri, w, "OOXV") == _lib.JUES_BAD_ARG
    assert _status(IT.get_eri, w, "OOVVV") == _lib.JUES_BAD_ARG
    IT.get_eri(w, "OOVV")                        # uploads the factors
    ctx = _lib.context(0)
    Co, Cv = _lib.f64(s.Co), _lib.f64(s.Cv)
    out = np.zeros((s.o, s.o, s.v, s.v), order="F")
    call = lambda prec, n: ctx.lib.jues_get_eri(  # noqa: E731
        ctx.h, prec, b"OOVV", b"phys", _lib.ptr(Co), s.o, _lib.ptr(Cv), s.v, _lib.ptr(Co), s.o, _lib.ptr(Cv), s.v,
        n, 0, _lib.ptr(out))
    assert call(_lib.FP64, s.n - 1) == _lib.JUES_BAD_ARG      # n must match the AO integrals
    assert call(7, s.n) == _lib.JUES_BAD_ARG                  # unknown precision
    assert call(_lib.FP16X3_MIXED, s.n) == _lib.JUES_BAD_ARG  # not a transform precision
    assert call(_lib.FP64, s.n) == _lib.JUES_OK
    # a beta block named without beta orbitals
    st = ctx.lib.jues_get_eri(ctx.h, _lib.FP64, b"OoVv", b"phys", _lib.ptr(Co), s.o, _lib.ptr(Cv), s.v, None, 0,
                              None, 0, s.n, 0, _lib.ptr(out))
    assert st == _lib.JUES_BAD_ARG
    E = np.zeros(1)
    eps = _lib.f64(s.eps)
    assert ctx.lib.jues_rmp2(ctx.h, 0, _lib.ptr(Co), s.o, _lib.ptr(Cv), s.v, s.n + 1, _lib.ptr(eps),
                             _lib.ptr(E)) == _lib.JUES_BAD_ARG


@pytest.mark.parametrize("prec", ["fp64", "fp32_mixed", "fp16x3"])
def test_identical_solves_are_bitwise_identical(cfg1, prec):
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn
    w = Wfn.from_synthetic(cfg1, dense=False)
    runs = [CC.mRCCD.do_rccd(w, prec=prec, return_T2=True, cc_e_conv=1e-12, cc_max_rms=1e-11, maxit=300)
            for _ in range(2)]
    (E0, T0), (E1, T1) = runs
    assert E0 == E1
    assert np.array_equal(T0, T1)


def test_identical_rccsd_and_gradient_bitwise(cfg1):
    from jues import CoupledCluster as CC, MollerPlesset
    from jues.Wavefunction import Wfn
    w = Wfn.from_synthetic(cfg1, dense=False)
    a = CC.mRCCSD.do_rccsd(w, return_T=True, cc_e_conv=1e-12, cc_max_rms=1e-11, maxit=300)
    b = CC.mRCCSD.do_rccsd(w, return_T=True, cc_e_conv=1e-12, cc_max_rms=1e-11, maxit=300)
    assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    g1 = MollerPlesset.rmp2_grad(w, want_dg=False)
    g2 = MollerPlesset.rmp2_grad(w, want_dg=False)
    assert g1[0] == g2[0] and np.array_equal(g1[1], g2[1]) and np.array_equal(g1[3], g2[3])
