"""STO-3G integrals and closed-shell RHF for small molecules -- TEST INFRASTRUCTURE ONLY.

Used only by tests/ to regenerate the inputs of the reference's own known-answer
tests (test/TestMollerPlesset.jl:35, test/TestCoupledCluster.jl:41-45,
test/testmp2.jl:3), which take water in STO-3G from psi4 or Lints.  Neither
integral engine is available here, so this module computes the integrals from
the published method:

* McMurchie-Davidson Hermite expansion for overlap, kinetic, nuclear attraction
  and electron repulsion over contracted Cartesian s/p Gaussians (Helgaker,
  Jorgensen, Olsen, "Molecular Electronic-Structure Theory", ch. 9);
* Boys function F_n(T) = 1F1(n+1/2; n+3/2; -T) / (2n+1) (scipy.special.hyp1f1);
* STO-3G exponents and contraction coefficients of Hehre, Stewart and Pople
  (J. Chem. Phys. 51, 2657 (1969)) as distributed in the EMSL/BSE library;
* a plain Roothaan-Hall SCF with DIIS converged to 1e-13 in the density for
  the psi4 cases (psi4 runs with d_convergence 10-14), and, for the Lints case,
  the reference's own loosely converged SCF loop restated step for step
  (``rhf_jues``), since the reference's MP2 known answer carries that loop's
  ~1e-9 incomplete convergence.

Third-party inputs restated here (absent from the reference tree): psi4's
geometry convention (CODATA 2014 bohr, psi4 >= 1.1) and Lints over libint_jll
2.7.0 (Manifest.toml:379-383), whose ``read_dotxyz`` converts Angstrom with
CODATA 2018 and whose sto-3g.g file carries 8-digit coefficients.
"""
from __future__ import annotations

import itertools
import math

import numpy as np
from scipy.special import hyp1f1

BOHR_PER_ANGSTROM = 1.0 / 0.52917721067  # CODATA 2014 (psi4 >= 1.1)
BOHR_PER_ANGSTROM_LIBINT = 1.0 / 0.529177210903  # CODATA 2018 (libint 2.7 read_dotxyz)

# element: list of (shell l, exponents, coefficients); the 2sp shell of O shares exponents
STO3G = {
    "H": [(0, [3.42525091, 0.62391373, 0.16885540], [0.15432897, 0.53532814, 0.44463454])],
    "O": [(0, [130.7093200, 23.8088610, 6.4436083], [0.15432897, 0.53532814, 0.44463454]),
          (0, [5.0331513, 1.1695961, 0.3803890], [-0.09996723, 0.39951283, 0.70011547]),
          (1, [5.0331513, 1.1695961, 0.3803890], [0.15591627, 0.60768372, 0.39195739])],
}
CHARGE = {"H": 1.0, "O": 8.0}


def _dfact(n):
    return 1.0 if n <= 0 else float(np.prod(np.arange(n, 0, -2)))


class CGTO:
    """Contracted Cartesian Gaussian x^i y^j z^k sum_p c_p exp(-a_p r^2), normalised."""

    def __init__(self, origin, lmn, exps, coefs):
        self.o = np.asarray(origin, float)
        self.lmn = lmn
        self.a = np.asarray(exps, float)
        l, m, n = lmn
        L = l + m + n
        # primitive normalisation
        self.norm = np.array([(2 * a / np.pi) ** 0.75 * (4 * a) ** (L / 2)
                              / math.sqrt(_dfact(2 * l - 1) * _dfact(2 * m - 1) * _dfact(2 * n - 1))
                              for a in self.a])
        self.c = np.asarray(coefs, float)
        # contraction normalisation
        s = 0.0
        for i, ai in enumerate(self.a):
            for j, aj in enumerate(self.a):
                s += (self.norm[i] * self.norm[j] * self.c[i] * self.c[j]
                      * _dfact(2 * l - 1) * _dfact(2 * m - 1) * _dfact(2 * n - 1)
                      / (2 * (ai + aj)) ** L * (np.pi / (ai + aj)) ** 1.5)
        self.c = self.c / math.sqrt(s)


def _E(i, j, t, Qx, a, b):
    """Hermite expansion coefficients (recursive)."""
    p = a + b
    q = a * b / p
    if t < 0 or t > i + j:
        return 0.0
    if i == j == t == 0:
        return math.exp(-q * Qx * Qx)
    if j == 0:
        return (1 / (2 * p) * _E(i - 1, j, t - 1, Qx, a, b) - (q * Qx / a) * _E(i - 1, j, t, Qx, a, b)
                + (t + 1) * _E(i - 1, j, t + 1, Qx, a, b))
    return (1 / (2 * p) * _E(i, j - 1, t - 1, Qx, a, b) + (q * Qx / b) * _E(i, j - 1, t, Qx, a, b)
            + (t + 1) * _E(i, j - 1, t + 1, Qx, a, b))


def boys(n, T):
    return hyp1f1(n + 0.5, n + 1.5, -T) / (2.0 * n + 1.0)


def _R(t, u, v, n, p, PCx, PCy, PCz, RPC):
    T = p * RPC * RPC
    if t == u == v == 0:
        return (-2 * p) ** n * boys(n, T)
    if t == u == 0:
        val = 0.0
        if v > 1:
            val += (v - 1) * _R(t, u, v - 2, n + 1, p, PCx, PCy, PCz, RPC)
        return val + PCz * _R(t, u, v - 1, n + 1, p, PCx, PCy, PCz, RPC)
    if t == 0:
        val = 0.0
        if u > 1:
            val += (u - 1) * _R(t, u - 2, v, n + 1, p, PCx, PCy, PCz, RPC)
        return val + PCy * _R(t, u - 1, v, n + 1, p, PCx, PCy, PCz, RPC)
    val = 0.0
    if t > 1:
        val += (t - 1) * _R(t - 2, u, v, n + 1, p, PCx, PCy, PCz, RPC)
    return val + PCx * _R(t - 1, u, v, n + 1, p, PCx, PCy, PCz, RPC)


def _prim_overlap(a, lmn1, A, b, lmn2, B):
    S1 = _E(lmn1[0], lmn2[0], 0, A[0] - B[0], a, b)
    S2 = _E(lmn1[1], lmn2[1], 0, A[1] - B[1], a, b)
    S3 = _E(lmn1[2], lmn2[2], 0, A[2] - B[2], a, b)
    return S1 * S2 * S3 * (np.pi / (a + b)) ** 1.5


def _prim_kinetic(a, lmn1, A, b, lmn2, B):
    l2, m2, n2 = lmn2
    t0 = b * (2 * (l2 + m2 + n2) + 3) * _prim_overlap(a, lmn1, A, b, lmn2, B)
    t1 = -2 * b ** 2 * (_prim_overlap(a, lmn1, A, b, (l2 + 2, m2, n2), B)
                        + _prim_overlap(a, lmn1, A, b, (l2, m2 + 2, n2), B)
                        + _prim_overlap(a, lmn1, A, b, (l2, m2, n2 + 2), B))
    t2 = -0.5 * (l2 * (l2 - 1) * _prim_overlap(a, lmn1, A, b, (l2 - 2, m2, n2), B)
                 + m2 * (m2 - 1) * _prim_overlap(a, lmn1, A, b, (l2, m2 - 2, n2), B)
                 + n2 * (n2 - 1) * _prim_overlap(a, lmn1, A, b, (l2, m2, n2 - 2), B))
    return t0 + t1 + t2


def _prim_nuclear(a, lmn1, A, b, lmn2, B, C):
    p = a + b
    P = (a * A + b * B) / p
    RPC = np.linalg.norm(P - C)
    val = 0.0
    for t in range(lmn1[0] + lmn2[0] + 1):
        for u in range(lmn1[1] + lmn2[1] + 1):
            for v in range(lmn1[2] + lmn2[2] + 1):
                val += (_E(lmn1[0], lmn2[0], t, A[0] - B[0], a, b) * _E(lmn1[1], lmn2[1], u, A[1] - B[1], a, b)
                        * _E(lmn1[2], lmn2[2], v, A[2] - B[2], a, b)
                        * _R(t, u, v, 0, p, P[0] - C[0], P[1] - C[1], P[2] - C[2], RPC))
    return val * 2 * np.pi / p


def _prim_eri(a, l1, A, b, l2, B, c, l3, Cc, d, l4, D):
    p, q = a + b, c + d
    alpha = p * q / (p + q)
    P = (a * A + b * B) / p
    Q = (c * Cc + d * D) / q
    RPQ = np.linalg.norm(P - Q)
    val = 0.0
    for t in range(l1[0] + l2[0] + 1):
        Et = _E(l1[0], l2[0], t, A[0] - B[0], a, b)
        for u in range(l1[1] + l2[1] + 1):
            Eu = _E(l1[1], l2[1], u, A[1] - B[1], a, b)
            for v in range(l1[2] + l2[2] + 1):
                Ev = _E(l1[2], l2[2], v, A[2] - B[2], a, b)
                for tau in range(l3[0] + l4[0] + 1):
                    Etau = _E(l3[0], l4[0], tau, Cc[0] - D[0], c, d)
                    for nu in range(l3[1] + l4[1] + 1):
                        Enu = _E(l3[1], l4[1], nu, Cc[1] - D[1], c, d)
                        for phi in range(l3[2] + l4[2] + 1):
                            Ephi = _E(l3[2], l4[2], phi, Cc[2] - D[2], c, d)
                            val += (Et * Eu * Ev * Etau * Enu * Ephi * (-1) ** (tau + nu + phi)
                                    * _R(t + tau, u + nu, v + phi, 0, alpha, P[0] - Q[0], P[1] - Q[1],
                                         P[2] - Q[2], RPQ))
    return val * 2 * np.pi ** 2.5 / (p * q * math.sqrt(p + q))


def basis(atoms):
    """atoms: list of (symbol, xyz in bohr) -> list of CGTO (s, then px py pz per p shell)."""
    bf = []
    for sym, xyz in atoms:
        for l, exps, coefs in STO3G[sym]:
            if l == 0:
                bf.append(CGTO(xyz, (0, 0, 0), exps, coefs))
            else:
                for lmn in [(1, 0, 0), (0, 1, 0), (0, 0, 1)]:
                    bf.append(CGTO(xyz, lmn, exps, coefs))
    return bf


def _contract(f, b1, b2):
    s = 0.0
    for i, a in enumerate(b1.a):
        for j, b in enumerate(b2.a):
            s += b1.norm[i] * b2.norm[j] * b1.c[i] * b2.c[j] * f(a, b1.lmn, b1.o, b, b2.lmn, b2.o)
    return s


def one_electron(bf, atoms):
    n = len(bf)
    S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1):
            S[i, j] = S[j, i] = _contract(_prim_overlap, bf[i], bf[j])
            T[i, j] = T[j, i] = _contract(_prim_kinetic, bf[i], bf[j])
            v = 0.0
            for sym, xyz in atoms:
                v -= CHARGE[sym] * _contract(lambda a, la, A, b, lb, B: _prim_nuclear(a, la, A, b, lb, B, xyz),
                                            bf[i], bf[j])
            V[i, j] = V[j, i] = v
    return S, T, V


def eri(bf):
    n = len(bf)
    g = np.zeros((n, n, n, n))
    for i, j, k, l in itertools.product(range(n), repeat=4):
        if j > i or l > k or (i * (i + 1) // 2 + j) < (k * (k + 1) // 2 + l):
            continue
        b1, b2, b3, b4 = bf[i], bf[j], bf[k], bf[l]
        s = 0.0
        for p1, a in enumerate(b1.a):
            for p2, b in enumerate(b2.a):
                for p3, c in enumerate(b3.a):
                    for p4, d in enumerate(b4.a):
                        s += (b1.norm[p1] * b2.norm[p2] * b3.norm[p3] * b4.norm[p4]
                              * b1.c[p1] * b2.c[p2] * b3.c[p3] * b4.c[p4]
                              * _prim_eri(a, b1.lmn, b1.o, b, b2.lmn, b2.o, c, b3.lmn, b3.o, d, b4.lmn, b4.o))
        for (p, q, r, t) in [(i, j, k, l), (j, i, k, l), (i, j, l, k), (j, i, l, k),
                             (k, l, i, j), (l, k, i, j), (k, l, j, i), (l, k, j, i)]:
            g[p, q, r, t] = s
    return g


def rhf(S, H, g, nocc, conv=1e-13, maxit=200):
    """Roothaan-Hall RHF with DIIS; returns (E_elec, C, eps)."""
    X = np.linalg.inv(np.linalg.cholesky(S)).T  # S^-1/2-like orthogonaliser (X^T S X = 1)
    F = H.copy()
    Dm = np.zeros_like(S)
    fs, es = [], []
    E_old = 0.0
    for it in range(maxit):
        Fp = X.T @ F @ X
        eps, Cp = np.linalg.eigh(Fp)
        C = X @ Cp
        Co = C[:, :nocc]
        Dn = Co @ Co.T
        J = np.einsum("mnls,ls->mn", g, Dn)
        K = np.einsum("mlns,ls->mn", g, Dn)
        F = H + 2 * J - K
        E = np.sum(Dn * (H + F))
        err = F @ Dn @ S - S @ Dn @ F
        fs.append(F)
        es.append(err)
        if len(fs) > 8:
            fs.pop(0)
            es.pop(0)
        m = len(fs)
        B = -np.ones((m + 1, m + 1))
        B[m, m] = 0
        for a in range(m):
            for b in range(m):
                B[a, b] = np.sum(es[a] * es[b])
        rhs = np.zeros(m + 1)
        rhs[m] = -1
        c = np.linalg.solve(B, rhs)
        F = sum(c[k] * fs[k] for k in range(m))
        dD = np.max(np.abs(Dn - Dm))
        Dm = Dn
        if it > 2 and dD < conv and abs(E - E_old) < 1e-14:
            break
        E_old = E
    # final canonical orbitals from the converged (unextrapolated) Fock matrix
    F = H + 2 * np.einsum("mnls,ls->mn", g, Dm) - np.einsum("mlns,ls->mn", g, Dm)
    eps, Cp = np.linalg.eigh(X.T @ F @ X)
    return E, X @ Cp, eps


def zmat_water(r_ang=1.1, theta_deg=104.0):
    """O at the origin, H1 on z at r, H2 at r with angle theta (psi4's "O; H 1 r; H 1 r 2 theta")."""
    r = r_ang * BOHR_PER_ANGSTROM
    th = math.radians(theta_deg)
    return [("O", np.zeros(3)), ("H", np.array([0.0, 0.0, r])), ("H", np.array([r * math.sin(th), 0.0, r * math.cos(th)]))]


def xyz_water_lints():
    """test/mp2_sto3g.dat: O (0,0,0), H (1,0,0), H (0,1,0) in Angstrom, read through libint."""
    c = BOHR_PER_ANGSTROM_LIBINT
    return [("O", np.zeros(3)), ("H", np.array([c, 0.0, 0.0])), ("H", np.array([0.0, c, 0.0]))]


def system(atoms, nelec=10, scf="tight"):
    """AO integrals and canonical RHF orbitals; scf="tight" (psi4-like, DIIS to
    1e-13) or scf="jues" (the reference's RHFCompute loop)."""
    bf = basis(atoms)
    S, T, V = one_electron(bf, atoms)
    g = eri(bf)
    solver = rhf if scf == "tight" else rhf_jues
    E, C, eps = solver(S, T + V, g, nelec // 2)
    return dict(S=S, H=T + V, g=g, C=C, eps=eps, nocc=nelec // 2, E_elec=E)


def rhf_jues(S, H, g, nocc, maxit=50, Etol=1e-7, Dtol=1e-7):
    """HartreeFock/RHF.jl:77-131 restated step for step (RHFCompute): symmetric
    orthogonaliser A = S^-1/2, core-Hamiltonian guess, no DIIS, and the
    reference's stopping test ``dE < Etol && sqrt(sum(dD)^2) < Dtol`` with E
    starting at 0.  The reference's E includes the nuclear repulsion; the first
    dE is negative either way and later ones are differences, so the loop stops
    on the same iteration with the electronic energy used here.  Returns
    (E_elec, C, eps) of the last iteration, as the reference stores them
    (RHF.jl:129-131).

    Pinned by test/testmp2.jl:3 (MP2 of its orbitals equals the reference's
    -0.04191336790304866 to 1e-16)."""
    w, U = np.linalg.eigh(S)
    A = U @ np.diag(w ** -0.5) @ U.T
    G = 2 * g - g.transpose(0, 2, 1, 3)
    e, Ct = np.linalg.eigh(A.T @ H @ A)
    C = A @ Ct
    Co = C[:, :nocc]
    D = Co @ Co.T
    E = 0.0
    for _ in range(maxit):
        F = H + np.einsum("rs,mnrs->mn", D, G)
        Eel = np.sum(D * (H + F))
        Ft = A.T @ F @ A
        Ft = 0.5 * (Ft + Ft.T)
        e, Ct = np.linalg.eigh(Ft)
        C = A @ Ct
        Co = C[:, :nocc]
        Dn = Co @ Co.T
        dD = Dn - D
        Drms = math.sqrt(np.sum(dD) ** 2)
        dE = Eel - E
        D, E = Dn, Eel
        if dE < Etol and Drms < Dtol:
            break
    return E, C, e


def uhf(S, H, g, na, nb, conv=1e-13, maxit=300):
    """Pople-Nesbet UHF with DIIS on the stacked alpha/beta commutators, core guess;
    returns (E_elec, Ca, Cb, epsa, epsb)."""
    X = np.linalg.inv(np.linalg.cholesky(S)).T
    Fa, Fb = H.copy(), H.copy()
    Da_old = np.zeros_like(S)
    hist = []
    E_old = 0.0

    def diag(F):
        e, Cp = np.linalg.eigh(X.T @ F @ X)
        return e, X @ Cp

    for it in range(maxit):
        _, Ca = diag(Fa)
        _, Cb = diag(Fb)
        Da = Ca[:, :na] @ Ca[:, :na].T
        Db = Cb[:, :nb] @ Cb[:, :nb].T
        J = np.einsum("mnls,ls->mn", g, Da + Db)
        Fa = H + J - np.einsum("mlns,ls->mn", g, Da)
        Fb = H + J - np.einsum("mlns,ls->mn", g, Db)
        E = 0.5 * np.sum((Da + Db) * H + Da * Fa + Db * Fb)
        err = np.concatenate([(Fa @ Da @ S - S @ Da @ Fa).ravel(), (Fb @ Db @ S - S @ Db @ Fb).ravel()])
        hist.append((Fa, Fb, err))
        if len(hist) > 10:
            hist.pop(0)
        m = len(hist)
        B = -np.ones((m + 1, m + 1))
        B[m, m] = 0
        for a in range(m):
            for b in range(m):
                B[a, b] = hist[a][2] @ hist[b][2]
        rhs = np.zeros(m + 1)
        rhs[m] = -1
        c = np.linalg.solve(B, rhs)
        Fa = sum(c[k] * hist[k][0] for k in range(m))
        Fb = sum(c[k] * hist[k][1] for k in range(m))
        dD = np.max(np.abs(Da - Da_old))
        Da_old = Da
        if it > 2 and dD < conv and abs(E - E_old) < 1e-14:
            break
        E_old = E
    J = np.einsum("mnls,ls->mn", g, Da + Db)
    ea, Ca = diag(H + J - np.einsum("mlns,ls->mn", g, Da))
    eb, Cb = diag(H + J - np.einsum("mlns,ls->mn", g, Db))
    return E, Ca, Cb, ea, eb


def system_uhf(atoms, na, nb):
    """AO integrals and UHF orbitals (psi4 "reference uhf", PK integrals)."""
    bf = basis(atoms)
    S, T, V = one_electron(bf, atoms)
    g = eri(bf)
    E, Ca, Cb, ea, eb = uhf(S, T + V, g, na, nb)
    return dict(S=S, H=T + V, g=g, Ca=Ca, Cb=Cb, epsa=ea, epsb=eb, na=na, nb=nb, E_elec=E)
