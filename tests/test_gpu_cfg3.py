"""cfg3 (n=256, o=40, v=216, naux=768): RMP2 energy and the MP2 adjoint dE/dC
through the factor route against the committed fixture (tests/golden/cfg3.npz):
64 seeded entries, analytic (oracle, factor route) and by central differences.

The synthetic integrals make E_MP2 about -6.4e5 Eh, so the central differences
(h = 1e-4) carry about 1e-5 of rounding noise; the comparison with them is
relative 1e-7 plus that floor, the comparison with the oracle's analytic
entries is relative 1e-10.
"""
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = pathlib.Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def cfg3():
    from jues import synthetic
    return synthetic.workload("cfg3_mp2grad_n256")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD / "cfg3.npz")


def _wfn(s):
    from jues.Wavefunction import Wfn, FactoredERI
    return Wfn.from_orbitals(s.C, s.eps, s.o, FactoredERI(s.B))


def test_cfg3_mp2_energy(cfg3, gold):
    from jues import MollerPlesset
    E = MollerPlesset.do_rmp2(_wfn(cfg3))
    assert abs(E - gold["E_mp2"]) < 1e-12 * abs(gold["E_mp2"])


def test_cfg3_mp2_gradient(cfg3, gold):
    from jues import MollerPlesset
    E, dC, dg = MollerPlesset.rmp2_grad(_wfn(cfg3), want_dg=False)
    assert dg is None
    assert abs(E - gold["E_mp2"]) < 1e-12 * abs(gold["E_mp2"])
    idx = gold["idx"]
    got = dC[idx[:, 0], idx[:, 1]]
    assert np.max(np.abs(got - gold["grad"])) < 1e-10 * np.max(np.abs(gold["grad"]))
    assert np.all(np.abs(got - gold["fd"]) < 1e-7 * np.abs(gold["fd"]) + 5e-5)


def test_cfg3_mp2_gradient_fp32_mixed(cfg3, gold):
    from jues import MollerPlesset
    E, dC, _ = MollerPlesset.rmp2_grad(_wfn(cfg3), want_dg=False, prec="fp32_mixed")
    idx = gold["idx"]
    got = dC[idx[:, 0], idx[:, 1]]
    assert abs(E - gold["E_mp2"]) < 1e-5 * abs(gold["E_mp2"])
    assert np.max(np.abs(got - gold["grad"])) < 1e-4 * np.max(np.abs(gold["grad"]))
