"""Generate the golden fixtures in tests/golden/ from the CPU oracle.

    python tests/golden/make_golden.py

cfg1 (n=24, o=5, v=19, naux=72): RMP2, RCCD (mRCCD.jl equations), RCCSD with
both reference equation sets, the first three plain-Jacobi energies from t=0,
converged amplitudes.  cfg3 (n=256, o=40, v=216, naux=768): RMP2 energy and dE/dC at 64 seeded (mu, p)
entries, analytic (factor route), by central differences with h = 1e-5 and by the
five-point stencil at h = 1e-1 (exact for the quartic E along one entry of C).
cfg4 (n=400, o=60, v=340, naux=1200) and cfg5 (n=500, o=80, v=420, naux=1500): E_MP2
through the factor route.
cfg2 (n=120, o=20, v=100, naux=360): the full-strength
iteration diverges for these integrals, so RCCD and RCCSD are converged for the
integrals scaled by a coupling of 0.02 (factors scaled by sqrt(0.02)); norms and
sampled entries of the amplitudes are stored.  Oracles converge to
|dE| < 1e-13 and ||dt|| < 1e-12.
"""
import pathlib
import sys
import time

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path[:0] = [str(ROOT / "oracle"), str(ROOT / "jues.jl_amd")]
import jues_oracle as O  # noqa: E402
from jues import synthetic  # noqa: E402

SAMPLE_IDX = 64


def cfg1():
    s = synthetic.workload("cfg1_rccd_n24")
    g = s.ao_eri()
    oovv = O.get_eri(g, s.Co, s.Cv, "OOVV")
    Emp2 = O.rmp2(oovv, s.eps, s.o)
    I = O.make_rccd_integrals(g, s.Co, s.Cv)
    hist = []
    O.solve_rccd(I, s.eps, s.o, guess="zero", diis=False, maxit=3, history=hist)
    E_ccd, T_ccd, _ = O.solve_rccd(I, s.eps, s.o, guess="zero", max_vecs=8, maxit=400)
    Ic = O.make_rccsd_integrals(g, s.Co, s.Cv)
    E_m, T1_m, T2_m, _ = O.solve_rccsd(Ic, s.eps, s.o, max_vecs=8, maxit=400)
    b = O.make_rccsd15_integrals(lambda a, b_, c, d: O.tei_transform(g, a, b_, c, d), s.Co, s.Cv)
    E_r, T1_r, T2_r, _ = O.solve_rccsd(b, s.eps, s.o, max_vecs=8, maxit=400, update=O.rccsd15_update)
    K = O.tei_transform(g, s.Co, s.Cv, s.Co, s.Cv)
    np.savez_compressed(HERE / "cfg1.npz", E_mp2=Emp2, jacobi_E=np.array([h[0] for h in hist]),
                        E_rccd=E_ccd, T2_rccd=T_ccd, E_mrccsd=E_m, T1_mrccsd=T1_m, T2_mrccsd=T2_m,
                        E_rccsd=E_r, T1_rccsd=T1_r, T2_rccsd=T2_r,
                        iajb_sum=K.sum(), iajb_sumsq=(K * K).sum())
    print("cfg1", Emp2, E_ccd, E_m, E_r)


def cfg2(lam=0.02):
    s = synthetic.workload("cfg2_rccsd_n120")
    B = s.B * np.sqrt(lam)
    t0 = time.time()
    I = O.rccd_integrals_from_factors(B, s.Co, s.Cv)
    Emp2 = O.rmp2(I["oovv"], s.eps, s.o)
    E_ccd, T_ccd, it = O.solve_rccd(I, s.eps, s.o, guess="zero", max_vecs=8, maxit=200)
    print("cfg2 rccd", E_ccd, it, time.time() - t0, flush=True)
    Ic = O.rccsd_integrals_from_factors(B, s.Co, s.Cv)
    E_m, T1_m, T2_m, it2 = O.solve_rccsd(Ic, s.eps, s.o, max_vecs=8, maxit=200)
    print("cfg2 rccsd", E_m, it2, time.time() - t0, flush=True)
    rng = np.random.default_rng(7)
    idx = rng.integers(0, T_ccd.size, SAMPLE_IDX)
    np.savez_compressed(HERE / "cfg2_coupling0.02.npz", coupling=lam, E_mp2=Emp2,
                        E_rccd=E_ccd, T2_rccd_norm=np.linalg.norm(T_ccd), T2_rccd_sample=T_ccd.ravel(order="F")[idx],
                        E_mrccsd=E_m, T1_mrccsd=T1_m, T2_mrccsd_norm=np.linalg.norm(T2_m),
                        T2_mrccsd_sample=T2_m.ravel(order="F")[idx], sample_idx=idx)


def cfg3(h=1e-5, h5=1e-1, count=64):
    """E_MP2 and dE/dC at 64 seeded entries: analytic (factor route), central differences at
    the configured h = 1e-5, and the five-point stencil at h5 = 1e-1.  Along one entry of C,
    E is a polynomial of degree 4 in the step (K is quadratic in C, E quadratic in K), which
    the five-point stencil differentiates exactly at any step; its error is rounding alone,
    about 1e-9 / h5 ~ 1e-8 for energy sums carrying ~1e-9 Eh of rounding, where the central
    difference at 1e-5 carries ~1e-4."""
    s = synthetic.workload("cfg3_mp2grad_n256")
    t0 = time.time()
    E, dC = O.rmp2_grad_factors(s.B, s.C, s.eps, s.o)
    rng = np.random.default_rng(3)
    idx = rng.integers(0, s.n, (count, 2))
    an = np.array([dC[mu, p] for mu, p in idx])
    fd = np.empty(count)
    fd5 = np.empty(count)

    def E_at(mu, p, step):
        Cs = s.C.copy()
        Cs[mu, p] += step
        return O.rmp2_factors(s.B, Cs, s.eps, s.o)

    for k, (mu, p) in enumerate(idx):
        fd[k] = (E_at(mu, p, h) - E_at(mu, p, -h)) / (2 * h)
        fd5[k] = (-E_at(mu, p, 2 * h5) + 8 * E_at(mu, p, h5) - 8 * E_at(mu, p, -h5) + E_at(mu, p, -2 * h5)) / (12 * h5)
        print(f"cfg3 fd {k} ({mu},{p}) an {an[k]:.10f} fd {fd[k]:.10f} fd5 {fd5[k]:.10f} "
              f"({time.time() - t0:.0f} s)", flush=True)
    np.savez_compressed(HERE / "cfg3.npz", E_mp2=E, idx=idx, grad=an, fd=fd, fd_h=h, fd5=fd5, fd5_h=h5)
    print("cfg3", E, np.max(np.abs(an - fd)), np.max(np.abs(an - fd5)))


def cfg4():
    """E_MP2 of the o=60, v=340 factors through the factor route, summed in slabs of one
    occupied index: the first Jacobi step from t = 0 at coupling lambda has energy
    lambda^2 E_MP2."""
    s = synthetic.workload("cfg4_rccd_n400")
    t0 = time.time()
    E = O.rmp2_factors_slabs(s.B, s.C, s.eps, s.o)
    np.savez_compressed(HERE / "cfg4.npz", E_mp2=E)
    print("cfg4", E, time.time() - t0)


def cfg5():
    """E_MP2 of the o=80, v=420 factors (K would be 9 GB: summed in slabs of one j), as for cfg4."""
    s = synthetic.workload("cfg5_rccsd_n500")
    t0 = time.time()
    E = O.rmp2_factors_slabs(s.B, s.C, s.eps, s.o)
    np.savez_compressed(HERE / "cfg5.npz", E_mp2=E)
    print("cfg5", E, time.time() - t0)


if __name__ == "__main__":
    which = sys.argv[1:] or ["cfg1", "cfg2", "cfg3", "cfg4", "cfg5"]
    for name in ("cfg1", "cfg2", "cfg3", "cfg4", "cfg5"):
        if name in which:
            globals()[name]()
