"""``tei_transform`` (Backend/Transformation.jl:15-93) on the GPU."""
from __future__ import annotations

import numpy as np

from . import _lib
from .Wavefunction import FactoredERI, ensure_ao, prec_code


def tei_transform(gao, C1, C2=None, C3=None, C4=None, name: str = "default", *, prec: str = "fp64",
                  device: int = 0) -> np.ndarray:
    """``out[p,q,r,s] = sum C1[mu,p] C2[nu,q] C3[lam,r] C4[sig,s] gao[mu,nu,lam,sig]`` (chemist).

    ``tei_transform(gao, C)`` is the one-C form (Transformation.jl:15-20).  ``gao``
    is a dense n^4 array or a ``FactoredERI``.  ``name`` is accepted for
    signature parity (the reference uses it only to name disk scratch files).
    """
    if isinstance(C2, str):
        name, C2 = C2, None
    if C2 is None:
        C2 = C3 = C4 = C1
    ctx = _lib.context(device)
    n = ensure_ao(ctx, gao)
    Cs = [_lib.f64(c) for c in (C1, C2, C3, C4)]
    for c in Cs:
        if c.shape[0] != n:
            raise ValueError("C has the wrong number of rows")
    d = [c.shape[1] for c in Cs]
    out = np.empty(d, dtype=np.float64, order="F")
    ctx.check(ctx.lib.jues_tei_transform(ctx.h, prec_code(prec), _lib.ptr(Cs[0]), d[0], _lib.ptr(Cs[1]), d[1],
                                         _lib.ptr(Cs[2]), d[2], _lib.ptr(Cs[3]), d[3], n, _lib.ptr(out)))
    return out


__all__ = ["tei_transform", "FactoredERI"]
