"""cfg3 (n=256, o=40, v=216, naux=768): RMP2 energy and the MP2 adjoint on the device
against the committed fixture (tests/golden/cfg3.npz: E_MP2 and 64 seeded entries of
dE/dC, analytic from the oracle's factor route, central differences at the configured
h = 1e-5, and the five-point stencil at h = 1e-1).

How close the analytic gradient can be shown to be to finite differences is set by
the energy's size: the synthetic integrals make E_MP2 about -6.4e5 Eh, so each energy
carries about eps |E| ~ 1e-10 Eh of rounding and a central difference at h carries
about that divided by h (1e-5 at h = 1e-5), which makes an absolute 1e-7 agreement with
h = 1e-5 differences unreachable in FP64.  Along one entry of C, though, E is a
polynomial of degree 4 in the step (K is quadratic in C and E quadratic in K), which
the five-point stencil differentiates exactly at any step: at h = 1e-1 its only error is
rounding (~1e-8), and the device gradient is held to 1e-7 of it.  The Taylor remainder
E(C + hX) - E(C) - h dE.X along a random direction falls as h^2 for h = 1e-2 ... 1e-5.
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


def _wfn(s, C=None, eps=None, B=None):
    from jues.Wavefunction import Wfn, FactoredERI
    return Wfn.from_orbitals(s.C if C is None else C, s.eps if eps is None else eps, s.o,
                             FactoredERI(s.B if B is None else B))


@pytest.fixture(scope="module")
def grad64(cfg3):
    from jues import MollerPlesset
    return MollerPlesset.rmp2_grad(_wfn(cfg3), want_dg=False)


def test_cfg3_mp2_energy(cfg3, gold):
    from jues import MollerPlesset
    E = MollerPlesset.do_rmp2(_wfn(cfg3))
    assert abs(E - gold["E_mp2"]) < 1e-12 * abs(gold["E_mp2"])


def test_cfg3_mp2_gradient(grad64, gold):
    E, dC, dg, de = grad64
    assert dg is None
    assert abs(E - gold["E_mp2"]) < 1e-12 * abs(gold["E_mp2"])
    idx = gold["idx"]
    got = dC[idx[:, 0], idx[:, 1]]
    # the oracle's analytic factor-route entries
    assert np.max(np.abs(got - gold["grad"])) < 1e-10 * np.max(np.abs(gold["grad"]))
    # exact-for-the-quartic five-point differences (rounding ~1e-8)
    assert np.max(np.abs(got - gold["fd5"])) < 1e-7
    # the configured central differences, h = 1e-5: agreement to their rounding floor
    floor = 200 * np.finfo(float).eps * abs(E) / gold["fd_h"]
    assert np.all(np.abs(got - gold["fd"]) < 1e-7 * np.abs(gold["fd"]) + floor)


def test_cfg3_taylor_remainder(cfg3, grad64):
    """E(C + hX) - E(C) - h <dE/dC, X> is O(h^2): each tenfold smaller h divides it by ~100."""
    from jues import MollerPlesset
    E0, dC, _, _ = grad64
    # a direction of the size of C's own entries (~n^-1/2), so h = 1e-2 is a 1% step
    X = np.random.default_rng(17).standard_normal(cfg3.C.shape) / np.sqrt(cfg3.n)
    slope = float(np.sum(dC * X))
    hs = (1e-2, 1e-3, 1e-4, 1e-5)
    r = [abs(MollerPlesset.do_rmp2(_wfn(cfg3, C=cfg3.C + h * X)) - E0 - h * slope) for h in hs]
    print("taylor remainders", dict(zip(hs, r)))
    for a, b in zip(r, r[1:]):
        assert 95 < a / b < 105, r
    # an error d in the slope adds d |slope| h to the remainder, so at the last decade (where
    # the h^2 term is smallest) a ratio within 1% of 100 bounds d by about 1e-5
    assert 99 < r[-2] / r[-1] < 101, r
    assert abs(r[-2] / r[-1] - 100) * r[-1] / 100 < 1e-5 * abs(slope) * 1e-5


def test_cfg3_eps_gradient_and_invariants(cfg3, grad64):
    """dE/deps: a uniform shift of eps leaves D, so E, unchanged (sum dE/de = 0); E is
    homogeneous of degree -1 in eps (sum e dE/de = -E); and a central difference."""
    from jues import MollerPlesset
    E, _, _, de = grad64
    assert abs(de.sum()) < 1e-11 * np.abs(de).sum()
    assert abs(np.dot(cfg3.eps, de) + E) < 1e-11 * abs(E)
    h = 1e-4
    for p in (0, cfg3.o - 1, cfg3.o, cfg3.n - 1):
        ep, em = cfg3.eps.copy(), cfg3.eps.copy()
        ep[p] += h
        em[p] -= h
        fd = (MollerPlesset.do_rmp2(_wfn(cfg3, eps=ep)) - MollerPlesset.do_rmp2(_wfn(cfg3, eps=em))) / (2 * h)
        assert abs(fd - de[p]) < 1e-6 * abs(fd), (p, fd, de[p])


def test_cfg3_mp2_gradient_fp32_mixed(cfg3, grad64, gold):
    from jues import MollerPlesset
    E, dC, _, de = grad64
    E32, dC32, _, de32 = MollerPlesset.rmp2_grad(_wfn(cfg3), want_dg=False, prec="fp32_mixed")
    rel_E = abs(E32 - E) / abs(E)
    rel_C = np.max(np.abs(dC32 - dC)) / np.max(np.abs(dC))
    rel_e = np.max(np.abs(de32 - de)) / np.max(np.abs(de))
    print(f"FP32-mixed vs FP64 at n=256: |dE|/|E| {rel_E:.2e}, dE/dC {rel_C:.2e}, dE/deps {rel_e:.2e}")
    assert rel_E < 1e-5 and rel_C < 1e-4 and rel_e < 1e-4


def test_cfg3_dense_dEdg_direction(cfg3):
    """dE/dg at n = 256 (dense AO tensor, 34 GB) along X = u u^T (u an n x n matrix): E is
    quadratic in g, so with D(h) = (E(g + hX) - E(g)) / h the value 2 D(h/2) - D(h) is the
    exact directional derivative; E(g + hX) comes from the factors [B, sqrt(h) u]."""
    from jues import MollerPlesset, _lib
    from jues.Wavefunction import Wfn
    s = cfg3
    n = s.n
    Bm = s.B.reshape(n * n, s.naux, order="F")
    G2 = np.empty((n * n, n * n))
    for r0 in range(0, n * n, 4096):                 # g as an n^2 x n^2 matrix, in row blocks
        G2[r0:r0 + 4096] = Bm[r0:r0 + 4096] @ Bm.T
    g = G2.T.reshape((n, n, n, n), order="F")
    ctx = _lib.context(0)
    try:
        E, dC, dg, _ = MollerPlesset.rmp2_grad(Wfn.from_orbitals(s.C, s.eps, s.o, g), want_dg=True)
        del g, G2
        u = np.random.default_rng(23).normal(0.0, 0.1, (n, n))
        uv = u.reshape(-1, order="F")
        dirn = float(uv @ (dg.reshape(n * n, n * n, order="F") @ uv))
        del dg
        ctx.release()

        def E_of(h):
            Bh = np.asfortranarray(np.concatenate([s.B, np.sqrt(h) * u[:, :, None]], axis=2))
            return MollerPlesset.do_rmp2(_wfn(s, B=Bh))

        E0 = MollerPlesset.do_rmp2(_wfn(s))
        assert abs(E0 - E) < 1e-12 * abs(E)
        h = 1.0
        D1, D2 = (E_of(h) - E0) / h, (E_of(h / 2) - E0) / (h / 2)
        fd = 2 * D2 - D1
        print(f"dE/dg . X: analytic {dirn:.10e} exact-quadratic difference {fd:.10e}")
        assert abs(dirn - fd) < 1e-9 * abs(fd) + 1e-8
    finally:
        ctx.release()
