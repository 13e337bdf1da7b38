"""Fock-driven closed-shell CCSD (CoupledCluster/AutoRCCSD.jl:39-300).

The oracle restates AutoRCCSD.jl's update term by term.  It is pinned:
* on canonical orbitals one update equals RCCSD.jl's restatement, and the solve
  reproduces psi4's CCSD for water, -0.070680102078571 (test/TestCoupledCluster.jl:44);
* occupied-occupied and virtual-virtual rotations (non-canonical Fock matrix, off-diagonal
  elements up to ~2 Eh) leave the converged energy unchanged;
* on a non-Hartree-Fock reference (an occupied-virtual rotation, |f_ia| up to 0.5) it
  agrees with an independent spin-orbital CCSD (Stanton-Gauss form, full Fock matrix).
The GPU tests run the same systems through jues_rccsd_fock.
"""
import numpy as np
import pytest
from scipy.linalg import expm

import jues_oracle as O
import sto3g

E_CCSD_PSI4 = -0.070680102078571


@pytest.fixture(scope="module")
def water():
    return sto3g.system(sto3g.zmat_water())


def _rotated(w, kind, scale, seed):
    """C rotated by exp(K): kind "oovv" mixes occupied and virtual orbitals among themselves,
    "ov" mixes occupied with virtual (a non-Hartree-Fock reference)."""
    C, o = w["C"], w["nocc"]
    n = C.shape[1]
    rng = np.random.default_rng(seed)
    K = np.zeros((n, n))
    if kind == "oovv":
        K[:o, :o] = rng.standard_normal((o, o))
        K[o:, o:] = rng.standard_normal((n - o, n - o))
    else:
        K[:o, o:] = rng.standard_normal((o, n - o))
    K = scale * (K - K.T)
    return C @ expm(K)


def _oracle_solve(w, C):
    o = w["nocc"]
    tei = lambda a, b, c, d: O.tei_transform(w["g"], a, b, c, d)  # noqa: E731
    f = O.get_fock(w["g"], w["H"], C, C[:, :o])
    V = O.auto_rccsd_integrals(tei, C[:, :o], C[:, o:])
    E, T1, T2, _ = O.solve_auto_rccsd(f, V, o, e_conv=1e-13, max_rms=1e-15, maxit=800)
    return E, T1, T2, f, V


@pytest.fixture(scope="module")
def non_hf(water):
    C = _rotated(water, "ov", 0.05, 2)
    E, T1, T2, f, V = _oracle_solve(water, C)
    return dict(C=C, E=E, T1=T1, T2=T2, f=f, V=V)


# ---------------------------------------------------------------- oracle (CPU)


def test_oracle_auto_canonical(water):
    E, *_ = _oracle_solve(water, water["C"])
    assert abs(E - E_CCSD_PSI4) < 1e-10


def test_oracle_auto_step_is_rccsd_step(water):
    w = water
    o = w["nocc"]
    C = w["C"]
    v = C.shape[1] - o
    tei = lambda a, b, c, d: O.tei_transform(w["g"], a, b, c, d)  # noqa: E731
    rng = np.random.default_rng(1)
    t1 = rng.standard_normal((o, v)) * 0.05
    t2 = rng.standard_normal((o, o, v, v)) * 0.05
    t2 = (t2 + t2.transpose(1, 0, 3, 2)) / 2
    f = O.get_fock(w["g"], w["H"], C, C[:, :o])
    fd = np.diag(f)
    fo = f - np.diag(fd)
    d = fd[:o, None] - fd[None, o:]
    D = fd[:o, None, None, None] + fd[None, :o, None, None] - fd[None, None, o:, None] - fd[None, None, None, o:]
    a1, a2, _, _ = O.auto_rccsd_update(t1, t2, (fo[:o, :o], fo[:o, o:], fo[o:, o:]),
                                       O.auto_rccsd_integrals(tei, C[:, :o], C[:, o:]), d, D)
    b1, b2 = O.rccsd15_update(t1, t2, O.make_rccsd15_integrals(tei, C[:, :o], C[:, o:]))
    assert np.abs(a1 - b1 / O.form_Dia(w["eps"], o, v)).max() < 1e-13
    assert np.abs(a2 - b2 / O.form_Dijab(w["eps"], o, v)).max() < 1e-13


def test_oracle_auto_rotation_invariance(water):
    C = _rotated(water, "oovv", 0.05, 1)
    E, _, _, f, _ = _oracle_solve(water, C)
    o = water["nocc"]
    assert np.abs(f[:o, :o] - np.diag(np.diag(f[:o, :o]))).max() > 0.1
    assert abs(E - E_CCSD_PSI4) < 1e-10


def test_oracle_auto_non_hf_vs_spin_orbital(water, non_hf):
    o = water["nocc"]
    assert np.abs(non_hf["f"][:o, o:]).max() > 0.1
    g_mo = O.tei_transform(water["g"], *(non_hf["C"],) * 4)
    assert abs(non_hf["E"] - O.ccsd_spin_orbital(non_hf["f"], g_mo, o)) < 1e-10


# ---------------------------------------------------------------- GPU


def _wfn(w, C):
    from jues.Wavefunction import Wfn
    return Wfn.from_orbitals(C, np.zeros(C.shape[1]), w["nocc"], np.asfortranarray(w["g"]), hao=w["H"])


@pytest.mark.gpu
def test_gpu_auto_canonical_and_pT(water):
    from jues import CoupledCluster as CC
    E, Et = CC.AutoRCCSD.do_rccsd(_wfn(water, water["C"]), do_pT=True, cc_e_conv=1e-13, cc_max_rms=1e-12,
                                  maxit=300)
    assert abs(E - E_CCSD_PSI4) < 1e-10
    w = water
    o = w["nocc"]
    Co, Cv = w["C"][:, :o], w["C"][:, o:]
    I = O.make_rccsd15_integrals(lambda a, b, c, d: O.tei_transform(w["g"], a, b, c, d), Co, Cv)
    _, T1, T2, _ = O.solve_rccsd(I, w["eps"], o, update=O.rccsd15_update, e_conv=1e-14, r_conv=1e-13, maxit=200)
    V = O.pt_integrals(lambda a, b, c, d: O.tei_transform(w["g"], a, b, c, d), Co, Cv)
    assert abs(Et - O.compute_pT(T1, T2, *V, w["eps"][:o], w["eps"][o:])) < 1e-12


@pytest.mark.gpu
def test_gpu_auto_rotated(water):
    from jues import CoupledCluster as CC
    C = _rotated(water, "oovv", 0.05, 1)
    E = CC.AutoRCCSD.do_rccsd(_wfn(water, C), cc_e_conv=1e-13, cc_max_rms=1e-12, maxit=400)
    assert abs(E - E_CCSD_PSI4) < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("prec,diis", [("fp64", False), ("fp64", True), ("fp32_mixed", True), ("fp16x3_mixed", True)])
def test_gpu_auto_non_hf(water, non_hf, prec, diis):
    from jues import CoupledCluster as CC
    E, T1, T2 = CC.AutoRCCSD.do_rccsd(_wfn(water, non_hf["C"]), prec=prec, diis=diis, return_T=True,
                                      cc_e_conv=1e-13, cc_max_rms=1e-12, maxit=400)
    assert abs(E - non_hf["E"]) < 1e-10
    assert np.abs(T1 - non_hf["T1"]).max() < 1e-9
    assert np.abs(T2 - non_hf["T2"]).max() < 1e-9


@pytest.mark.gpu
def test_gpu_auto_one_step(water, non_hf):
    """One update from random amplitudes on the non-HF reference, against AutoRCCSD.jl's terms."""
    from jues import CoupledCluster as CC
    o = water["nocc"]
    C = non_hf["C"]
    v = C.shape[1] - o
    rng = np.random.default_rng(3)
    t1 = rng.standard_normal((o, v)) * 0.05
    t2 = rng.standard_normal((o, o, v, v)) * 0.05
    t2 = (t2 + t2.transpose(1, 0, 3, 2)) / 2
    f = non_hf["f"]
    fd = np.diag(f)
    fo = f - np.diag(fd)
    d = fd[:o, None] - fd[None, o:]
    D = fd[:o, None, None, None] + fd[None, :o, None, None] - fd[None, None, o:, None] - fd[None, None, None, o:]
    a1, a2, _, _ = O.auto_rccsd_update(t1, t2, (fo[:o, :o], fo[:o, o:], fo[o:, o:]), non_hf["V"], d, D)
    E, g1, g2 = CC.AutoRCCSD.do_rccsd(_wfn(water, C), T1=t1, T2=t2, maxit=1, check_convergence=False,
                                      return_T=True)
    assert np.abs(g1 - a1).max() < 1e-12
    assert np.abs(g2 - a2).max() < 1e-12
    assert abs(E - O.auto_rccsd_energy(a1, a2, fo[:o, o:], non_hf["V"]["oovv"])) < 1e-12


@pytest.mark.gpu
def test_gpu_auto_needs_hao(water):
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn
    w = Wfn.from_orbitals(water["C"], water["eps"], water["nocc"], np.asfortranarray(water["g"]))
    with pytest.raises(ValueError):
        CC.AutoRCCSD.do_rccsd(w)
