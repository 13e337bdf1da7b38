"""MP2 adjoint (dE/dC, dE/dg, dE/deps) on the GPU against the oracle and finite differences."""
import numpy as np
import pytest

import jues_oracle as O
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _small(n=10, o=3, naux=16, k=5, scale=0.6):
    from jues import synthetic
    s = synthetic.make_system(n, o, n - o, naux, k, name="g")
    s.B = np.asfortranarray(s.B * scale)
    return s


@pytest.mark.parametrize("dense", [True, False])
def test_grad_matches_oracle(dense):
    from jues import MollerPlesset
    from jues.Wavefunction import Wfn, FactoredERI
    s = _small()
    g = s.ao_eri()
    E0, dC0, dg0 = O.rmp2_grad(g, s.C, s.eps, s.o)
    de0 = O.rmp2_grad_eps(O.tei_transform(g, s.Co, s.Cv, s.Co, s.Cv), s.eps, s.o)
    w = Wfn.from_orbitals(s.C, s.eps, s.o, g if dense else FactoredERI(s.B))
    E, dC, dg, de = MollerPlesset.rmp2_grad(w, want_dg=dense)
    assert abs(E - E0) < 1e-12
    assert rel_err(dC, dC0) < 1e-11
    assert rel_err(de, de0) < 1e-11
    if dense:
        assert rel_err(dg, dg0) < 1e-11


def test_grad_nonsymmetric_tensor():
    """Every g entry independent: a random non-symmetric tensor still matches the oracle."""
    from jues import MollerPlesset
    from jues.Wavefunction import Wfn
    s = _small(n=8, o=3)
    rng = np.random.default_rng(9)
    g = np.asfortranarray(s.ao_eri() + 0.01 * rng.normal(size=(8, 8, 8, 8)))
    E0, dC0, dg0 = O.rmp2_grad(g, s.C, s.eps, s.o)
    de0 = O.rmp2_grad_eps(O.tei_transform(g, s.Co, s.Cv, s.Co, s.Cv), s.eps, s.o)
    E, dC, dg, de = MollerPlesset.rmp2_grad(Wfn.from_orbitals(s.C, s.eps, s.o, g))
    assert abs(E - E0) < 1e-12 and rel_err(dC, dC0) < 1e-11 and rel_err(dg, dg0) < 1e-11
    assert rel_err(de, de0) < 1e-11


def test_grad_central_differences_and_direction():
    from jues import MollerPlesset
    from jues.Wavefunction import Wfn
    s = _small(n=12, o=4, naux=20)
    g = s.ao_eri()
    w = Wfn.from_orbitals(s.C, s.eps, s.o, g)
    E, dC, dg, de = MollerPlesset.rmp2_grad(w)
    h = 1e-5
    rng = np.random.default_rng(64)
    for _ in range(16):
        mu, p = rng.integers(0, s.n, 2)
        Cp, Cm = s.C.copy(), s.C.copy()
        Cp[mu, p] += h
        Cm[mu, p] -= h
        fd = (MollerPlesset.do_rmp2(Wfn.from_orbitals(Cp, s.eps, s.o, g))
              - MollerPlesset.do_rmp2(Wfn.from_orbitals(Cm, s.eps, s.o, g))) / (2 * h)
        assert abs(fd - dC[mu, p]) < 1e-7, (mu, p, fd, dC[mu, p])
    X = np.asfortranarray(rng.normal(size=g.shape))
    fd = (MollerPlesset.do_rmp2(Wfn.from_orbitals(s.C, s.eps, s.o, np.asfortranarray(g + h * X)))
          - MollerPlesset.do_rmp2(Wfn.from_orbitals(s.C, s.eps, s.o, np.asfortranarray(g - h * X)))) / (2 * h)
    assert abs(fd - np.sum(dg * X)) < 1e-7 * max(1.0, abs(fd))
    for p in range(s.n):
        ep, em = s.eps.copy(), s.eps.copy()
        ep[p] += h
        em[p] -= h
        fd = (MollerPlesset.do_rmp2(Wfn.from_orbitals(s.C, ep, s.o, g))
              - MollerPlesset.do_rmp2(Wfn.from_orbitals(s.C, em, s.o, g))) / (2 * h)
        assert abs(fd - de[p]) < 1e-7 * max(1.0, abs(fd)), (p, fd, de[p])
    assert abs(de.sum()) < 1e-12 * np.abs(de).max()


def test_grad_fp32_mixed_vs_fp64():
    from jues import MollerPlesset
    from jues.Wavefunction import Wfn
    s = _small(n=16, o=4, naux=24)
    w = Wfn.from_orbitals(s.C, s.eps, s.o, s.ao_eri())
    E, dC, dg, de = MollerPlesset.rmp2_grad(w)
    E32, dC32, dg32, de32 = MollerPlesset.rmp2_grad(w, prec="fp32_mixed")
    assert abs(E32 - E) < 1e-5 * abs(E)
    assert rel_err(dC32, dC) < 1e-4 and rel_err(dg32, dg) < 1e-4 and rel_err(de32, de) < 1e-4


def test_grad_factor_route_nonsymmetric_factors():
    """Factor route with B_P not symmetric in (mu, nu): matches the dense restatement of the
    same integrals (the route never forms an n^2 o v half-transform)."""
    from jues import MollerPlesset
    from jues.Wavefunction import Wfn, FactoredERI
    s = _small(n=14, o=4, naux=22)
    rng = np.random.default_rng(21)
    B = np.asfortranarray(s.B + 0.05 * rng.normal(size=s.B.shape))
    g = np.einsum("mnP,lsP->mnls", B, B)
    E0, dC0, _ = O.rmp2_grad(g, s.C, s.eps, s.o)
    E1, dC1 = O.rmp2_grad_factors(B, s.C, s.eps, s.o)
    assert abs(E1 - E0) < 1e-12 * abs(E0) and rel_err(dC1, dC0) < 1e-11
    for prec, tol in (("fp64", 1e-11), ("fp32_mixed", 1e-4)):
        E, dC, dg, de = MollerPlesset.rmp2_grad(Wfn.from_orbitals(s.C, s.eps, s.o, FactoredERI(B)), prec=prec)
        assert dg is None
        assert abs(E - E0) < max(tol, 1e-12) * abs(E0) and rel_err(dC, dC0) < tol, prec
