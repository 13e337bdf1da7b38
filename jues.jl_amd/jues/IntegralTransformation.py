"""``get_eri`` (Backend/IntegralTransformation.jl:38-101) and ``get_fock`` (:119-141) on the GPU."""
from __future__ import annotations

import numpy as np

from . import _lib
from .Wavefunction import Wfn, ensure_ao, prec_code


def get_eri(wfn: Wfn, eri_string: str, notation: str = "phys", fcn: int = 0, *, prec: str = "fp64",
            device: int = 0) -> np.ndarray:
    """MO integral block named by a 4-letter string over {o,O,v,V}.

    ``notation="phys"`` returns <pq|rs> (the default, as the reference);
    ``"chem"`` returns (pq|rs).  Upper-case letters are the alpha orbitals
    (``O`` -> ``Cao``, nalpha; ``V`` -> ``Cav``, nvira), lower-case the beta ones
    (``o`` -> ``Cbo``, nbeta; ``v`` -> ``Cbv``, nvirb), as IntegralTransformation.jl:52-68;
    they coincide for a restricted ``Wfn``.  ``fcn`` drops the first fcn occupied
    orbitals of either spin.  Raises ``JuesError`` (BAD_ARG) for an invalid string, like
    the reference's ``error``.
    """
    ctx = _lib.context(device)
    n = ensure_ao(ctx, wfn.ao_eri)
    s = eri_string if isinstance(eri_string, str) else str(eri_string)
    if len(s.encode()) != 4 or any(c not in "oOvV" for c in s):
        # same message the library returns; checked here to size the output
        raise _lib.JuesError(_lib.JUES_BAD_ARG, f"Invalid string given to get_eri: {s}")
    size = {"O": wfn.nalpha - fcn, "o": wfn.nbeta - fcn, "V": wfn.nvira, "v": wfn.nvirb}
    out = np.empty([size[c] for c in s], dtype=np.float64, order="F")
    Cao, Cav, Cbo, Cbv = (_lib.f64(x) for x in (wfn.Cao, wfn.Cav, wfn.Cbo, wfn.Cbv))
    ctx.check(ctx.lib.jues_get_eri(ctx.h, prec_code(prec), s.encode(), notation.encode(),
                                   _lib.ptr(Cao), wfn.nalpha, _lib.ptr(Cav), wfn.nvira,
                                   _lib.ptr(Cbo), wfn.nbeta, _lib.ptr(Cbv), wfn.nvirb, n, fcn, _lib.ptr(out)))
    return out


def get_fock(wfn: Wfn, spin: str = "alpha", *, device: int = 0) -> np.ndarray:
    """MO Fock matrix f = C^T (h + 2J - K) C of the spin's occupied density
    (IntegralTransformation.jl:119-141); needs ``wfn.hao``."""
    sp = spin.lower()
    if sp in ("alpha", "up", "a"):
        C, Co = wfn.Ca, wfn.Cao
    elif sp in ("beta", "down", "b"):
        C, Co = wfn.Cb, wfn.Cbo
    else:
        raise ValueError(f"Invalid Spin option given to JuES.IntegralTransformation.get_fock: {spin}")
    if wfn.hao is None:
        raise ValueError("get_fock needs the core Hamiltonian wfn.hao")
    ctx = _lib.context(device)
    n = ensure_ao(ctx, wfn.ao_eri)
    C, Co, h = _lib.f64(C), _lib.f64(Co), _lib.f64(wfn.hao)
    m = C.shape[1]
    f = np.zeros((m, m), order="F")
    ctx.check(ctx.lib.jues_get_fock(ctx.h, _lib.ptr(h), _lib.ptr(C), m, _lib.ptr(Co), Co.shape[1], n, _lib.ptr(f)))
    return f


__all__ = ["get_eri", "get_fock"]
