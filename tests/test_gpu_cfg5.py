"""RCCSD and the MP2 adjoint at o=80, v=420 (n=500, naux=1500), the largest workload: the
pp-ladder integrals are rebuilt from the factors every iteration (DF-RCCD.jl:198-215) and
the o v^3 block (ia|bc) is never stored; the MP2 adjoint takes the factor route.

RCCSD from t = 0 (t1 = t2 = 0; canonical orbitals, so the T1 residual starts at zero): the
first Jacobi step gives t2 = lambda <ij|ab>/D and t1 = 0, whose energy is lambda^2 E_MP2
with E_MP2 from the oracle's factor route (tests/golden/cfg5.npz), in every form
(mRCCSD.jl:111-117 energy).  The second step, the first with every contraction active,
differs between the forms by contraction rounding only.  The integrals are scaled by
lambda = 0.005 (every residual term carries one integral), as for o=60, v=340.
"""
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
COUPLING = 0.005
_second = {}


@pytest.fixture(scope="module")
def cfg5():
    from jues import synthetic, _lib
    from jues.Wavefunction import Wfn
    s = synthetic.workload("cfg5_rccsd_n500")
    yield s, Wfn.from_synthetic(s, dense=False)
    _lib.context(0).release()


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD / "cfg5.npz")


def _two_steps(w, prec):
    from jues.CoupledCluster import Solver
    sv = Solver(w, method="rccsd", prec=prec, guess="zero", diis=False, coupling=COUPLING, ladder="recompute")
    e = [sv.iterate(1, prec=prec)[0] for _ in range(2)]
    print(f"cfg5 RCCSD {prec}: E1 {e[0]['energy']:.12f} E2 {e[1]['energy']:.12f} "
          f"({e[0]['ms'] / 1e3:.2f} / {e[1]['ms'] / 1e3:.2f} s)", flush=True)
    return [x["energy"] for x in e]


@pytest.mark.timeout(600)
@pytest.mark.parametrize("prec", ["fp64", "fp32_mixed", "fp16x3"])
def test_cfg5_rccsd_first_steps(cfg5, gold, prec):
    E1, E2 = _two_steps(cfg5[1], prec)
    ref = COUPLING ** 2 * float(gold["E_mp2"])
    assert abs(E1 - ref) < 1e-12 * abs(ref)
    _second[prec] = E2
    if prec != "fp64":
        E2_64 = _second["fp64"]
        print(f"cfg5 RCCSD {prec}: second-step |dE| vs FP64 {abs(E2 - E2_64):.3e} ({abs(E2 - E2_64) / abs(E2_64):.2e})")
        assert abs(E2 - E2_64) < 1e-5 * abs(E2_64)


@pytest.mark.timeout(300)
def test_cfg5_mp2_adjoint_factor_route(cfg5, gold):
    """dE/dC at n = 500 through the factor route: the energy matches the fixture and two seeded
    entries match the five-point stencil of the device energy (exact for the quartic E along
    one entry of C) at h = 0.1."""
    from jues import MollerPlesset
    from jues.Wavefunction import Wfn, FactoredERI
    s, _ = cfg5
    wf = lambda C: Wfn.from_orbitals(C, s.eps, s.o, FactoredERI(s.B))  # noqa: E731
    E, dC, dg, de = MollerPlesset.rmp2_grad(wf(s.C), want_dg=False)
    assert dg is None
    assert abs(E - float(gold["E_mp2"])) < 1e-12 * abs(float(gold["E_mp2"]))
    assert abs(de.sum()) < 1e-11 * np.abs(de).sum()
    h = 0.1
    for mu, p in np.random.default_rng(6).integers(0, s.n, (2, 2)):
        def E_at(step):
            C = s.C.copy()
            C[mu, p] += step
            return MollerPlesset.do_rmp2(wf(C))
        fd5 = (-E_at(2 * h) + 8 * E_at(h) - 8 * E_at(-h) + E_at(-2 * h)) / (12 * h)
        print(f"cfg5 dE/dC[{mu},{p}]: analytic {dC[mu, p]:.10f} five-point {fd5:.10f}")
        assert abs(fd5 - dC[mu, p]) < 1e-6 + 1e-10 * abs(fd5)
