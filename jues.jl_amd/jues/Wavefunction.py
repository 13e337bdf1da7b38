"""The ``Wfn`` carrier (Backend/Wavefunction.jl:67-141), restricted closed shell.

Field names follow the Julia struct.  ``ao_eri`` is either a dense n^4 array
(chemist order) or a ``FactoredERI`` holding B[mu,nu,P] with
(mu nu|lam sig) = sum_P B[mu,nu,P] B[lam,sig,P]; the second form plays the
role the reference gives ``DiskFourTensor`` (Wavefunction.jl:87): an ERI carrier
for systems whose dense tensor does not fit.  A ``DFIntegrals`` holds the pair
``setup_df`` returns (Backend/DF.jl:30-51), pqP[mu,nu,P] and Jpqh = (P|Q)^(-1/2);
uploading it forms b = pqP.Jpqh on the device and from then on it is the
factored form.
"""
from __future__ import annotations

import dataclasses
from typing import Any, Optional, Union

import numpy as np

from . import _lib


@dataclasses.dataclass
class FactoredERI:
    B: np.ndarray  # (n, n, naux)

    @property
    def n(self) -> int:
        return self.B.shape[0]


@dataclasses.dataclass
class DFIntegrals:
    pqP: np.ndarray   # (n, n, naux) three-index integrals (mu nu|P)
    Jpqh: np.ndarray  # (naux, naux) inverse square root of the metric (P|Q)

    @property
    def n(self) -> int:
        return self.pqP.shape[0]


@dataclasses.dataclass
class Wfn:
    energy: float
    vnuc: float
    nalpha: int
    nbeta: int
    nvira: int
    nvirb: int
    nmo: int
    unrestricted: bool
    basis: Any
    molecule: Any
    Ca: np.ndarray
    Cb: np.ndarray
    Cao: np.ndarray
    Cav: np.ndarray
    Cbo: np.ndarray
    Cbv: np.ndarray
    hao: Optional[np.ndarray]
    epsa: np.ndarray
    epsb: np.ndarray
    ao_eri: Union[np.ndarray, FactoredERI, DFIntegrals]

    @classmethod
    def from_orbitals(cls, C: np.ndarray, eps: np.ndarray, nocc: int, ao_eri, energy: float = 0.0,
                      vnuc: float = 0.0, hao=None) -> "Wfn":
        """Restricted Wfn from MO coefficients, orbital energies and AO integrals (Wavefunction.jl:94-141)."""
        C = np.asfortranarray(C, dtype=np.float64)
        n = C.shape[0]
        nvir = n - nocc
        Co = np.asfortranarray(C[:, :nocc])
        Cv = np.asfortranarray(C[:, nocc:])
        eps = np.asarray(eps, dtype=np.float64)
        return cls(energy, vnuc, nocc, nocc, nvir, nvir, n, False, None, None, C, C, Co, Cv, Co, Cv, hao,
                   eps, eps, ao_eri)

    @classmethod
    def from_uhf(cls, Ca: np.ndarray, Cb: np.ndarray, epsa: np.ndarray, epsb: np.ndarray, nalpha: int, nbeta: int,
                 ao_eri, energy: float = 0.0, vnuc: float = 0.0, hao=None) -> "Wfn":
        """Unrestricted Wfn (Wavefunction.jl:94-141 with unrestricted = true)."""
        Ca = np.asfortranarray(Ca, dtype=np.float64)
        Cb = np.asfortranarray(Cb, dtype=np.float64)
        n = Ca.shape[0]
        return cls(energy, vnuc, nalpha, nbeta, n - nalpha, n - nbeta, n, True, None, None, Ca, Cb,
                   np.asfortranarray(Ca[:, :nalpha]), np.asfortranarray(Ca[:, nalpha:]),
                   np.asfortranarray(Cb[:, :nbeta]), np.asfortranarray(Cb[:, nbeta:]), hao,
                   np.asarray(epsa, dtype=np.float64), np.asarray(epsb, dtype=np.float64), ao_eri)

    @classmethod
    def from_synthetic(cls, system, dense: bool = True) -> "Wfn":
        eri = system.ao_eri() if dense else FactoredERI(system.B)
        return cls.from_orbitals(system.C, system.eps, system.o, eri)


def _ao_key(eri) -> tuple:
    if isinstance(eri, FactoredERI):
        return ("factors", id(eri.B), eri.B.shape)
    if isinstance(eri, DFIntegrals):
        return ("df", id(eri.pqP), id(eri.Jpqh), eri.pqP.shape)
    return ("dense", id(eri), eri.shape)


def ensure_ao(ctx: "_lib.Context", eri, force: bool = False) -> int:
    """Make ``eri`` the context's current AO integrals. Returns n.

    The device copy is reused while the same array object is passed again (the
    context holds a reference, so the identity cannot be recycled).  After
    modifying an ``ao_eri`` array in place, pass ``force=True`` (or call
    ``invalidate``) to upload it again.
    """
    key = _ao_key(eri)
    if force or ctx.ao_key != key or getattr(ctx, "ao_ref", None) is not _ao_obj(eri):
        if isinstance(eri, FactoredERI):
            ctx.set_ao_factors(eri.B, key)
        elif isinstance(eri, DFIntegrals):
            ctx.set_df_integrals(eri.pqP, eri.Jpqh, key)
        else:
            ctx.set_ao_eri(eri, key)
        ctx.ao_ref = _ao_obj(eri)
    return eri.n if isinstance(eri, (FactoredERI, DFIntegrals)) else eri.shape[0]


def is_factored(eri) -> bool:
    return isinstance(eri, (FactoredERI, DFIntegrals))


def _ao_obj(eri):
    if isinstance(eri, FactoredERI):
        return eri.B
    if isinstance(eri, DFIntegrals):
        return eri.pqP
    return eri


def invalidate(device: int = 0) -> None:
    """Forget the device copy of the AO integrals (the next call uploads again)."""
    ctx = _lib.context(device)
    ctx.ao_key = None
    ctx.ao_ref = None


def prec_code(prec: str) -> int:
    p = prec.lower()
    if p in ("fp64", "float64", "double"):
        return _lib.FP64
    if p in ("fp32_mixed", "fp32", "mixed"):
        return _lib.FP32_MIXED
    if p in ("fp16x3_mixed", "fp16x3"):
        return _lib.FP16X3_MIXED
    raise ValueError(f"unknown precision {prec!r}")
