"""Pins the CPU oracle (no GPU): closed forms, brute force, invariants, two
independent restatements of each method, finite differences and the golden
fixtures.  Every check is chosen so that a plausible slip (a dropped term, a
swapped index, a transposed operand) fails it."""
import pathlib

import numpy as np
import pytest

import jues_oracle as O
from jues import synthetic

GOLD = pathlib.Path(__file__).resolve().parent / "golden"


def tiny(n=6, o=2, naux=10, k=11, scale=1.0):
    s = synthetic.make_system(n, o, n - o, naux, k, name="tiny")
    s.B = s.B * scale
    return s


# ---------------------------------------------------------------- transform
def test_transform_matches_bruteforce_nonsymmetric():
    rng = np.random.default_rng(1)
    n = 4
    g = rng.normal(size=(n, n, n, n))
    Cs = [rng.normal(size=(n, d)) for d in (2, 3, 1, 4)]
    assert np.allclose(O.tei_transform(g, *Cs), O.tei_transform_bruteforce(g, *Cs), atol=1e-13)


def test_transform_factor_route_and_invariants():
    s = tiny(n=8, o=3, naux=12)
    g = s.ao_eri()
    for Cs in [(s.Co, s.Cv, s.Co, s.Cv), (s.Cv, s.Co, s.Co, s.Cv), (s.C,) * 4]:
        assert np.allclose(O.tei_transform(g, *Cs), O.mo_from_factors(s.B, *Cs), atol=1e-13)
    I = np.eye(s.n)
    assert np.allclose(O.tei_transform(g, I, I, I, I), g, atol=1e-14)
    full = O.tei_transform(g, s.C, s.C, s.C, s.C)
    assert abs(np.linalg.norm(full) - np.linalg.norm(g)) < 1e-12 * np.linalg.norm(g)
    # 8-fold symmetry survives a common orthogonal C
    for p in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)]:
        assert np.allclose(full, full.transpose(p), atol=1e-13)


def test_get_eri_notation_and_fcn():
    s = tiny(n=8, o=3, naux=12)
    g = s.ao_eri()
    chem = O.get_eri(g, s.Co, s.Cv, "OVOV", notation="chem")
    phys = O.get_eri(g, s.Co, s.Cv, "OOVV")
    assert np.allclose(phys, chem.transpose(0, 2, 1, 3))
    fz = O.get_eri(g, s.Co, s.Cv, "OOVV", fcn=1)
    assert np.allclose(fz, phys[1:, 1:])
    with pytest.raises(ValueError):
        O.get_eri(g, s.Co, s.Cv, "OOV")


# ---------------------------------------------------------------- MP2
def test_rmp2_loops_and_vectorised_agree():
    s = tiny()
    oovv = O.get_eri(s.ao_eri(), s.Co, s.Cv, "OOVV")
    assert abs(O.rmp2(oovv, s.eps, s.o) - O.rmp2_loops(oovv, s.eps, s.o)) < 1e-13


def test_rmp2_rank1_closed_form_and_homogeneity():
    rng = np.random.default_rng(3)
    n, o = 6, 2
    b = rng.normal(size=(n, n))
    b = 0.5 * (b + b.T)
    B = b[:, :, None]
    s = tiny(n=n, o=o)
    Co, Cv = s.Co, s.Cv
    bia = Co.T @ b @ Cv  # (o, v)
    D = O.form_Dijab(s.eps, o, n - o)
    E_closed = sum(bia[i, a] * bia[j, c] * (2 * bia[i, a] * bia[j, c] - bia[i, c] * bia[j, a]) / D[i, j, a, c]
                   for i in range(o) for j in range(o) for a in range(n - o) for c in range(n - o))
    g = np.einsum("mnP,lsP->mnls", B, B)
    E = O.rmp2(O.get_eri(g, Co, Cv, "OOVV"), s.eps, o)
    assert abs(E - E_closed) < 1e-13 * max(1, abs(E))
    lam = 1.7
    E2 = O.rmp2(O.get_eri(lam ** 2 * g, Co, Cv, "OOVV"), s.eps, o)
    assert abs(E2 - lam ** 4 * E) < 1e-12 * abs(E2)


def test_rmp2_gradient_central_differences():
    s = tiny(n=6, o=2, naux=8, scale=0.5)
    g = s.ao_eri()
    E, dC, dg = O.rmp2_grad(g, s.C, s.eps, s.o)
    assert abs(E - O.rmp2_energy_from_C(g, s.C, s.eps, s.o)) < 1e-14
    h = 1e-5
    rng = np.random.default_rng(0)
    for _ in range(12):
        mu, p = rng.integers(0, s.n, 2)
        Cp, Cm = s.C.copy(), s.C.copy()
        Cp[mu, p] += h
        Cm[mu, p] -= h
        fd = (O.rmp2_energy_from_C(g, Cp, s.eps, s.o) - O.rmp2_energy_from_C(g, Cm, s.eps, s.o)) / (2 * h)
        assert abs(fd - dC[mu, p]) < 1e-8 * max(1, abs(fd))
    # dE/dg along a random direction (E is quadratic in g: central difference is exact up to rounding)
    X = rng.normal(size=g.shape)
    Co, Cv = s.Co, s.Cv
    Eg = lambda gg: O.rmp2(O.get_eri(gg, Co, Cv, "OOVV"), s.eps, s.o)  # noqa: E731
    fd = (Eg(g + h * X) - Eg(g - h * X)) / (2 * h)
    assert abs(fd - np.sum(dg * X)) < 1e-7 * max(1, abs(fd))


def test_rmp2_factor_route_gradient():
    """The factor-route energy and dE/dC (used where the n^4 tensor cannot exist)
    equal the dense restatement, and its dE/dC passes central differences on a
    non-symmetric factor tensor; a sign slip in the exchange term would fail both."""
    s = tiny(n=9, o=3, naux=14, scale=0.5)
    g = s.ao_eri()
    E0, dC0, _ = O.rmp2_grad(g, s.C, s.eps, s.o)
    E1, dC1 = O.rmp2_grad_factors(s.B, s.C, s.eps, s.o)
    assert abs(E1 - E0) < 1e-13 and np.max(np.abs(dC1 - dC0)) < 1e-12 * np.max(np.abs(dC0))
    assert abs(O.rmp2_factors(s.B, s.C, s.eps, s.o) - E0) < 1e-13
    rng = np.random.default_rng(4)
    Bn = np.asfortranarray(s.B + 0.05 * rng.normal(size=s.B.shape))   # (mu nu) symmetry broken
    E2, dC2 = O.rmp2_grad_factors(Bn, s.C, s.eps, s.o)
    h = 1e-5
    for _ in range(10):
        mu, p = rng.integers(0, s.n, 2)
        Cp, Cm = s.C.copy(), s.C.copy()
        Cp[mu, p] += h
        Cm[mu, p] -= h
        fd = (O.rmp2_factors(Bn, Cp, s.eps, s.o) - O.rmp2_factors(Bn, Cm, s.eps, s.o)) / (2 * h)
        assert abs(fd - dC2[mu, p]) < 1e-8 * max(1, abs(fd))


def test_rmp2_factors_slabs():
    """The slab-by-slab factor-route energy (used for the o=60/v=340 and o=80/v=420 fixtures)
    equals the dense restatement on a system where direct and exchange terms differ."""
    s = tiny(n=11, o=4, naux=15, scale=0.5)
    E0 = O.rmp2(O.get_eri(s.ao_eri(), s.Co, s.Cv, "OOVV"), s.eps, s.o)
    assert abs(O.rmp2_factors_slabs(s.B, s.C, s.eps, s.o) - E0) < 1e-13 * abs(E0)
    assert abs(O.rmp2_factors(s.B, s.C, s.eps, s.o) - E0) < 1e-13 * abs(E0)
    K = O.get_eri(s.ao_eri(), s.Co, s.Cv, "OOVV")
    assert np.abs(K - K.transpose(0, 1, 3, 2)).max() > 1e-3 * np.abs(K).max()


def test_rmp2_grad_eps():
    """dE/deps: central differences on every orbital energy (with a K that has no pair
    symmetry, so both slots of each orbital are exercised), plus two exact invariants
    of D = e_i + e_j - e_a - e_b: a uniform shift leaves E unchanged (sum_p dE/de_p = 0)
    and E is homogeneous of degree -1 in eps (sum_p e_p dE/de_p = -E)."""
    s = tiny(n=8, o=3, naux=12, scale=0.5)
    o, v = s.o, s.v
    rng = np.random.default_rng(11)
    K = O.tei_transform(s.ao_eri(), s.Co, s.Cv, s.Co, s.Cv) + 0.02 * rng.normal(size=(o, v, o, v))
    E_of = lambda e: O.rmp2(K.transpose(0, 2, 1, 3), e, o)  # noqa: E731
    de = O.rmp2_grad_eps(K, s.eps, o)
    assert de.shape == (s.n,)
    h = 1e-6
    for p in range(s.n):
        ep, em = s.eps.copy(), s.eps.copy()
        ep[p] += h
        em[p] -= h
        fd = (E_of(ep) - E_of(em)) / (2 * h)
        assert abs(fd - de[p]) < 1e-8 * max(1.0, abs(fd)), p
    E = E_of(s.eps)
    assert abs(de.sum()) < 1e-13 * np.abs(de).max()
    assert abs(np.dot(s.eps, de) + E) < 1e-13 * abs(E)
    # a dropped slot (only the (ia) position of each orbital) fails the FD check
    D = O.form_Dijab(s.eps, o, v).transpose(0, 2, 1, 3)
    W = K * (2 * K - K.transpose(0, 3, 2, 1)) / D ** 2
    one_slot = -W.sum(axis=(1, 2, 3))
    assert np.abs(one_slot - de[:o]).max() > 1e-3 * np.abs(de[:o]).max()


def test_golden_cfg3_consistent():
    """cfg3 fixture: the stored analytic entries agree with the five-point differences (exact
    for the quartic E along one entry of C, rounding ~1e-8) to 1e-7, and with the central
    differences at h = 1e-5 to their rounding floor (E ~ 6.4e5 Eh)."""
    path = GOLD / "cfg3.npz"
    if not path.exists():
        pytest.skip("cfg3 fixture not generated")
    z = np.load(path)
    assert z["grad"].shape == (64,) and z["idx"].shape == (64, 2)
    assert float(z["fd_h"]) == 1e-5
    assert np.max(np.abs(z["grad"] - z["fd5"])) < 1e-7
    floor = 200 * np.finfo(float).eps * abs(float(z["E_mp2"])) / float(z["fd_h"])
    assert np.all(np.abs(z["grad"] - z["fd"]) < 1e-7 * np.abs(z["fd"]) + floor)


# ---------------------------------------------------------------- RCCD
@pytest.fixture(scope="module")
def sys1():
    s = synthetic.workload("cfg1_rccd_n24")
    return s, O.make_rccd_integrals(s.ao_eri(), s.Co, s.Cv)


def test_rccd_two_restatements_agree(sys1):
    s, I = sys1
    rng = np.random.default_rng(2)
    t = rng.normal(size=(s.o, s.o, s.v, s.v)) * 0.02
    t = 0.5 * (t + t.transpose(1, 0, 3, 2))
    R1 = O.rccd_residual_tensor(t, I)
    R2 = O.mrccd_residual_tensor(t, I)
    assert np.max(np.abs(R1 - R2)) < 1e-13
    # R keeps the pair symmetry t[ijab] = t[jiba]
    assert np.max(np.abs(R1 - R1.transpose(1, 0, 3, 2))) < 1e-13


def test_rccd_first_jacobi_step_is_mp2(sys1):
    s, I = sys1
    D = O.form_Dijab(s.eps, s.o, s.v)
    t1 = O.rccd_step(np.zeros_like(I["oovv"]), I, D)
    assert abs(O.ccd_energy(t1, I["oovv"]) - O.rmp2(I["oovv"], s.eps, s.o)) < 1e-12


def test_rccd_converged_residual_vanishes_in_second_restatement():
    gold = np.load(GOLD / "cfg1.npz")
    s = synthetic.workload("cfg1_rccd_n24")
    I = O.make_rccd_integrals(s.ao_eri(), s.Co, s.Cv)
    T = gold["T2_rccd"]
    D = O.form_Dijab(s.eps, s.o, s.v)
    assert np.max(np.abs(O.mrccd_residual_tensor(T, I) - D * T)) < 1e-10
    assert abs(O.ccd_energy(T, I["oovv"]) - float(gold["E_rccd"])) < 1e-12


def test_rccd_fixed_point_independent_of_guess_and_diis():
    s = tiny(n=10, o=3, naux=20, scale=0.4)
    I = O.rccd_integrals_from_factors(s.B, s.Co, s.Cv)
    Es = []
    for guess, diis in [("zero", True), ("mp2", True), ("ovov", True), ("ovov", False)]:
        E, T, _ = O.solve_rccd(I, s.eps, s.o, guess=guess, diis=diis, maxit=500)
        Es.append(E)
    assert max(Es) - min(Es) < 1e-11


def test_diis_one_vector_is_jacobi():
    s = tiny(n=10, o=3, naux=20, scale=0.4)
    I = O.rccd_integrals_from_factors(s.B, s.Co, s.Cv)
    h1, h2 = [], []
    O.solve_rccd(I, s.eps, s.o, diis=True, max_vecs=1, maxit=6, e_conv=0, r_conv=0, history=h1)
    O.solve_rccd(I, s.eps, s.o, diis=False, maxit=6, e_conv=0, r_conv=0, history=h2)
    assert np.allclose(np.array(h1), np.array(h2), rtol=1e-13, atol=1e-15)


def test_rccd_small_coupling_limit():
    s = tiny(n=10, o=3, naux=20)
    I = O.rccd_integrals_from_factors(s.B, s.Co, s.Cv)
    Emp2 = O.rmp2(I["oovv"], s.eps, s.o)
    for lam in (1e-2, 1e-3):
        Il = {k: lam * x for k, x in I.items()}
        E, _, _ = O.solve_rccd(Il, s.eps, s.o, maxit=200)
        assert abs(E / (lam ** 2 * Emp2) - 1) < 30 * lam


# ---------------------------------------------------------------- RCCSD
@pytest.fixture(scope="module")
def ccsd_blocks():
    s = tiny(n=10, o=3, naux=20, scale=0.6)
    g = s.ao_eri()
    Ic = O.make_rccsd_integrals(g, s.Co, s.Cv)
    b = O.make_rccsd15_integrals(lambda a, b_, c, d: O.tei_transform(g, a, b_, c, d), s.Co, s.Cv)
    return s, Ic, b


def test_rccsd_restatements_differ_only_in_fae_t1_term(ccsd_blocks):
    """mRCCSD.jl:140 contracts t1 with <me|af>, RCCSD.jl:189 with <ma|ef>; nothing else differs."""
    s, Ic, b = ccsd_blocks
    rng = np.random.default_rng(4)
    t1 = rng.normal(size=(s.o, s.v)) * 0.05
    t2 = rng.normal(size=(s.o, s.o, s.v, s.v)) * 0.05
    t2 = 0.5 * (t2 + t2.transpose(1, 0, 3, 2))
    n1a, n2a = O.rccsd_update(t1, t2, Ic)
    n1b, n2b = O.rccsd15_update(t1, t2, b)
    ovvv = Ic["ovvv"]  # <ma|fe>
    dF = np.einsum("mf,maef->ae", t1, ovvv) - np.einsum("mf,meaf->ae", t1, ovvv)  # (<ma|ef> - <me|af>) t1[mf]
    assert np.max(np.abs((n1b - n1a) - (-np.einsum("ie,ae->ia", t1, dF)))) < 1e-13
    d2 = -(np.einsum("ijae,be->ijab", t2, dF) + np.einsum("ijeb,ae->ijab", t2, dF))
    assert np.max(np.abs((n2b - n2a) - d2)) < 1e-13
    n1c, n2c = O.rccsd_update(0 * t1, t2, Ic)
    n1d, n2d = O.rccsd15_update(0 * t1, t2, b)
    assert np.max(np.abs(n1c - n1d)) < 1e-14 and np.max(np.abs(n2c - n2d)) < 1e-14


def test_rccsd_reduces_to_rccd_without_ov_couplings(ccsd_blocks):
    s, Ic, _ = ccsd_blocks
    I0 = dict(Ic)
    I0["ooov"] = 0 * I0["ooov"]
    I0["ovvv"] = 0 * I0["ovvv"]
    E, T1, T2, _ = O.solve_rccsd(I0, s.eps, s.o, maxit=400)
    Ir = O.make_rccd_integrals(s.ao_eri(), s.Co, s.Cv)
    Ed, Td, _ = O.solve_rccd(Ir, s.eps, s.o, maxit=400)
    assert np.max(np.abs(T1)) == 0.0
    assert abs(E - Ed) < 1e-11 and np.max(np.abs(T2 - Td)) < 1e-10


def test_rccsd_guess_energy_is_mp2(ccsd_blocks):
    s, Ic, _ = ccsd_blocks
    D = O.form_Dijab(s.eps, s.o, s.v)
    E0 = O.ccsd_energy(Ic["oovv"], np.zeros((s.o, s.v)), Ic["oovv"] / D)
    assert abs(E0 - O.rmp2(Ic["oovv"], s.eps, s.o)) < 1e-13


# ---------------------------------------------------------------- golden fixtures
def test_golden_cfg1_reproduced():
    gold = np.load(GOLD / "cfg1.npz")
    s = synthetic.workload("cfg1_rccd_n24")
    g = s.ao_eri()
    K = O.tei_transform(g, s.Co, s.Cv, s.Co, s.Cv)
    assert abs(K.sum() - float(gold["iajb_sum"])) < 1e-10
    oovv = O.get_eri(g, s.Co, s.Cv, "OOVV")
    assert abs(O.rmp2(oovv, s.eps, s.o) - float(gold["E_mp2"])) < 1e-12
    I = O.make_rccd_integrals(g, s.Co, s.Cv)
    h = []
    O.solve_rccd(I, s.eps, s.o, guess="zero", diis=False, maxit=3, history=h)
    assert np.allclose([x[0] for x in h], gold["jacobi_E"], rtol=1e-13)
    assert abs(float(gold["E_mp2"]) - float(gold["jacobi_E"][0])) < 1e-12
