"""Seeded synthetic closed-shell systems for the correlation hot path.

Every workload is built from the same recipe so that any machine regenerates
bit-identical inputs from the seeds alone:

* factors ``B[mu, nu, P]`` (n x n x naux): ``X ~ N(0, 0.1^2)`` i.i.d., then
  ``B = (X + X^T_{mu nu}) / 2``.  The AO integrals are
  ``(mu nu|lam sig) = sum_P B[mu,nu,P] B[lam,sig,P]``, which has exact 8-fold
  permutational symmetry and is positive semidefinite as an n^2 x n^2 matrix.
* ``C``: the Q factor of a seeded n x n N(0,1) matrix with signs fixed so that
  ``diag(R) > 0``; columns ``0..o-1`` are occupied, ``o..n-1`` virtual.
* ``eps``: o sorted U(-1.5, -0.3) draws, then v sorted U(0.2, 3.0) draws.

Seeds are ``{B: 1000+k, C: 2000+k, eps: 3000+k}`` for workload ``k`` with
NumPy's PCG64 ``default_rng``.  Arrays are returned in Fortran (column-major)
order, which is the memory order the C ABI and the Julia callers use.
"""
from __future__ import annotations

import dataclasses

import numpy as np

# name -> (k, n, o, v, naux)
WORKLOADS = {
    "cfg1_rccd_n24": (1, 24, 5, 19, 72),
    "cfg2_rccsd_n120": (2, 120, 20, 100, 360),
    "cfg3_mp2grad_n256": (3, 256, 40, 216, 768),
    "cfg4_rccd_n400": (4, 400, 60, 340, 1200),
    "cfg5_rccsd_n500": (5, 500, 80, 420, 1500),
}


@dataclasses.dataclass
class System:
    name: str
    n: int
    o: int
    v: int
    naux: int
    B: np.ndarray      # (n, n, naux), Fortran order, symmetric in the first two indices
    C: np.ndarray      # (n, n), Fortran order, orthogonal
    eps: np.ndarray    # (n,)

    @property
    def Co(self) -> np.ndarray:
        return np.asfortranarray(self.C[:, : self.o])

    @property
    def Cv(self) -> np.ndarray:
        return np.asfortranarray(self.C[:, self.o:])

    def ao_eri(self, dtype=np.float64) -> np.ndarray:
        """Dense AO integrals g[mu,nu,lam,sig] (n^4, Fortran order) from the factors."""
        n = self.n
        Bm = self.B.reshape(n * n, self.naux, order="F").astype(dtype)
        g = Bm @ Bm.T
        return np.asfortranarray(g.reshape(n, n, n, n, order="F"))


def make_factors(n: int, naux: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    X = rng.normal(0.0, 0.1, size=(naux, n, n))
    Bp = 0.5 * (X + X.transpose(0, 2, 1))          # B[P, mu, nu]
    return np.asfortranarray(Bp.transpose(1, 2, 0))  # B[mu, nu, P]


def make_orbitals(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    G = rng.normal(size=(n, n))
    Q, R = np.linalg.qr(G)
    s = np.sign(np.diag(R))
    s[s == 0] = 1.0
    return np.asfortranarray(Q * s[None, :])


def make_eps(o: int, v: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    eo = np.sort(rng.uniform(-1.5, -0.3, size=o))
    ev = np.sort(rng.uniform(0.2, 3.0, size=v))
    return np.concatenate([eo, ev])


def make_system(n: int, o: int, v: int, naux: int, k: int, name: str = "custom") -> System:
    if n != o + v:
        raise ValueError(f"n={n} must equal o+v={o + v}")
    return System(name=name, n=n, o=o, v=v, naux=naux,
                  B=make_factors(n, naux, 1000 + k),
                  C=make_orbitals(n, 2000 + k),
                  eps=make_eps(o, v, 3000 + k))


def workload(name: str) -> System:
    k, n, o, v, naux = WORKLOADS[name]
    return make_system(n, o, v, naux, k, name=name)
