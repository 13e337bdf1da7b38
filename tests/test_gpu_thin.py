"""Released FP64 integral blocks during a fast stage (JUES_THIN_FP64_INTS=1 forces what low
device memory triggers at o=80, v=420): each fast form still converges to the FP64 fixed
point, and switching stages (FP32 -> split-FP16 -> FP64, as the benchmark does) rebuilds
what the next stage reads."""
import os

import numpy as np
import pytest

import jues_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture
def thin():
    os.environ["JUES_THIN_FP64_INTS"] = "1"
    yield
    os.environ.pop("JUES_THIN_FP64_INTS", None)


@pytest.fixture(scope="module")
def refs(cfg1):
    s = cfg1
    I = O.rccd_integrals_from_factors(s.B, s.Co, s.Cv)
    E_ccd, _, _ = O.solve_rccd(I, s.eps, s.o, maxit=300)
    Ic = O.rccsd_integrals_from_factors(s.B, s.Co, s.Cv)
    E_ccsd, _, _, _ = O.solve_rccsd(Ic, s.eps, s.o, maxit=300)
    return E_ccd, E_ccsd


@pytest.mark.parametrize("prec", ["fp32_mixed", "fp16x3"])
@pytest.mark.parametrize("ladder", ["stored", "recompute"])
def test_thin_fast_forms_converge(cfg1, refs, thin, prec, ladder):
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn
    w = Wfn.from_synthetic(cfg1, dense=False)
    kw = dict(prec=prec, ladder=ladder, cc_e_conv=1e-12, cc_max_rms=1e-11, maxit=300)
    assert abs(CC.mRCCD.do_rccd(w, **kw) - refs[0]) < 1e-10
    assert abs(CC.mRCCSD.do_rccsd(w, **kw) - refs[1]) < 1e-10


def test_thin_stage_switches(cfg1, thin):
    """The benchmark's sequence on one resident state: FP32 steps, split-FP16 steps, FP64
    steps, each stage from the amplitudes the previous left; the result equals the same
    sequence without releasing anything."""
    from jues.CoupledCluster import Solver
    from jues.Wavefunction import Wfn
    w = Wfn.from_synthetic(cfg1, dense=False)

    def run():
        sv = Solver(w, method="rccsd", prec="fp16x3", guess="zero", diis=True, diis_max=6)
        for prec, k in (("fp32_mixed", 4), ("fp16x3", 4), ("fp64", 4), ("fp32_mixed", 2), ("fp64", 2)):
            sv.reset_diis()
            sv.iterate(k, prec=prec)
        return sv.result()

    E1, T1a, T2a = run()
    os.environ.pop("JUES_THIN_FP64_INTS", None)
    E0, T1b, T2b = run()
    assert E1 == E0 and np.array_equal(T1a, T1b) and np.array_equal(T2a, T2b)
