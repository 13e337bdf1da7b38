"""Density fitting (Backend/DF.jl:1-60) on the GPU.

The reference's ``setup_df`` (DF.jl:30-51) asks psi4 for the three-index
integrals (mu nu|P) and the metric (P|Q) in an auxiliary basis and returns
``pqP, Jpqh`` with Jpqh = (P|Q)^(-1/2).  That integral front end is not part of
this library; its output pair is the input here.  ``make_bmunu`` (DF.jl:52-58,
``bμν[p,q,Q] := pqP[p,q,P]*Jpqh[P,Q]``) runs as one FP64 MFMA GEMM on the device.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .Wavefunction import DFIntegrals, FactoredERI, ensure_ao


def make_bmunu(pqP: np.ndarray, Jpqh: np.ndarray, *, device: int = 0) -> np.ndarray:
    """b[mu,nu,Q] = sum_P pqP[mu,nu,P] Jpqh[P,Q], formed on the GPU (DF.jl:52-58).

    The context keeps b as its current AO factors, so a following correlation
    call on a ``Wfn`` carrying the same ``DFIntegrals`` does not upload again.
    """
    ctx = _lib.context(device)
    eri = DFIntegrals(pqP, Jpqh)
    n = ensure_ao(ctx, eri, force=True)
    return ctx.get_ao_factors(n, pqP.shape[2])


make_bμν = make_bmunu


def factors(pqP: np.ndarray, Jpqh: np.ndarray, *, device: int = 0) -> FactoredERI:
    """b as a ``FactoredERI`` (the carrier the correlation methods take)."""
    return FactoredERI(np.asfortranarray(make_bmunu(pqP, Jpqh, device=device)))


__all__ = ["DFIntegrals", "make_bmunu", "make_bμν", "factors"]
