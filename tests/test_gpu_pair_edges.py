"""Packed-pair edge cases: one occupied or one virtual orbital leaves no strict pairs
(i < j or a < b), so the antisymmetric parts of both ladders vanish and their products are
skipped; o = 2, v = 7 gives odd pair counts.  RCCD in all three forms and RCCSD in FP64 go
to the oracle's converged fixed points (weakly coupled factors so that plain DIIS converges).
"""
import numpy as np
import pytest

import jues_oracle as O

pytestmark = pytest.mark.gpu

SYSTEMS = [(6, 1, 5), (6, 5, 1), (9, 2, 7)]


def _system(n, o, v):
    from jues import synthetic
    s = synthetic.make_system(n, o, v, 3 * n, 40 + n + o, name="edge")
    return s, np.asfortranarray(0.3 * s.B)


@pytest.mark.parametrize("n,o,v", SYSTEMS)
@pytest.mark.parametrize("prec", ["fp64", "fp32_mixed", "fp16x3_mixed"])
def test_rccd_pair_edges(n, o, v, prec):
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn, FactoredERI
    s, B = _system(n, o, v)
    I = O.rccd_integrals_from_factors(B, s.Co, s.Cv)
    Eref, Tref, _ = O.solve_rccd(I, s.eps, o, maxit=300)
    w = Wfn.from_orbitals(s.C, s.eps, o, FactoredERI(B))
    E, T = CC.mRCCD.do_rccd(w, prec=prec, return_T2=True, cc_e_conv=1e-12, cc_max_rms=1e-11, maxit=300)
    assert abs(E - Eref) < 1e-10
    assert np.max(np.abs(T - Tref)) < 1e-9


@pytest.mark.parametrize("n,o,v", SYSTEMS)
def test_rccsd_pair_edges(n, o, v):
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn, FactoredERI
    s, B = _system(n, o, v)
    I = O.rccsd_integrals_from_factors(B, s.Co, s.Cv)
    Eref, T1ref, T2ref, _ = O.solve_rccsd(I, s.eps, o)
    w = Wfn.from_orbitals(s.C, s.eps, o, FactoredERI(B))
    E, T1, T2 = CC.mRCCSD.do_rccsd(w, return_T=True, cc_e_conv=1e-12, cc_max_rms=1e-11, maxit=300)
    assert abs(E - Eref) < 1e-10
    assert np.max(np.abs(T1 - T1ref)) < 1e-9
    assert np.max(np.abs(T2 - T2ref)) < 1e-9
