"""``do_rmp2`` (MollerPlesset/RMP2.jl:11-45), ``do_df_rmp2``
(MollerPlesset/DF-RMP2.jl:1-46), ``do_ump2`` (MollerPlesset/UMP2.jl:12-79) and the
MP2 adjoint on the GPU."""
from __future__ import annotations

import numpy as np

from . import _lib
from .Wavefunction import Wfn, DFIntegrals, ensure_ao, is_factored, prec_code


def do_rmp2(refWfn: Wfn, *, prec: str = "fp64", device: int = 0, **kwargs) -> float:
    """Closed-shell MP2 correlation energy (RMP2.jl:33-37)."""
    ctx = _lib.context(device)
    n = ensure_ao(ctx, refWfn.ao_eri)
    Co, Cv = _lib.f64(refWfn.Cao), _lib.f64(refWfn.Cav)
    eps = _lib.f64(refWfn.epsa)
    E = np.zeros(1)
    ctx.check(ctx.lib.jues_rmp2(ctx.h, prec_code(prec), _lib.ptr(Co), refWfn.nalpha, _lib.ptr(Cv), refWfn.nvira,
                                n, _lib.ptr(eps), _lib.ptr(E)))
    return float(E[0])


def do_rmp2_noalloc(refWfn: Wfn, *, device: int = 0) -> float:
    """``do_rmp2`` in FP64 through ``jues_rmp2_noalloc``: no device allocation during the
    call (every intermediate in the library's static per-GPU scratch array).  The AO
    integrals must be factored (``FactoredERI`` / ``DFIntegrals``); uploading them, on
    first use, is the only allocating step and happens before the call."""
    if not is_factored(refWfn.ao_eri):
        raise TypeError("do_rmp2_noalloc needs FactoredERI or DFIntegrals integrals")
    ctx = _lib.context(device)
    n = ensure_ao(ctx, refWfn.ao_eri)
    Co, Cv = _lib.f64(refWfn.Cao), _lib.f64(refWfn.Cav)
    eps = _lib.f64(refWfn.epsa)
    E = np.zeros(1)
    ctx.check(ctx.lib.jues_rmp2_noalloc(ctx.h, _lib.FP64, _lib.ptr(Co), refWfn.nalpha, _lib.ptr(Cv),
                                        refWfn.nvira, n, _lib.ptr(eps), _lib.ptr(E)))
    return float(E[0])


def do_df_rmp2(refWfn: Wfn, *, prec: str = "fp64", device: int = 0) -> float:
    """Density-fitted closed-shell MP2 (DF-RMP2.jl:1-46).

    ``refWfn.ao_eri`` must be a ``DFIntegrals`` (pqP, Jpqh), the pair the
    reference's ``setup_df`` returns (DF-RMP2.jl:7), or a ``FactoredERI``
    holding b[mu,nu,Q] already contracted with Jpqh (DF-RMP2.jl:13-15).  The
    energy sum over (i,j) pairs of (ia|jb)(2(ia|jb) - (ib|ja))/D with
    (ia|jb) = sum_Q b_ia^Q b_jb^Q (DF-RMP2.jl:22-43) is the factor route of
    ``do_rmp2``: b_ia^Q is formed on the device and never leaves it.
    """
    if not is_factored(refWfn.ao_eri):
        raise TypeError("do_df_rmp2 needs DFIntegrals (pqP, Jpqh) or FactoredERI integrals")
    return do_rmp2(refWfn, prec=prec, device=device)


def do_ump2(refWfn: Wfn, *, prec: str = "fp64", device: int = 0) -> float:
    """Unrestricted MP2 (UMP2.jl:12-79): alpha-alpha and beta-beta <ij||ab>^2/D over
    i > j, a > b, plus alpha-beta (ia|JB)^2/D.  On a restricted Wfn it equals do_rmp2."""
    ctx = _lib.context(device)
    n = ensure_ao(ctx, refWfn.ao_eri)
    Cao, Cav = _lib.f64(refWfn.Cao), _lib.f64(refWfn.Cav)
    Cbo, Cbv = _lib.f64(refWfn.Cbo), _lib.f64(refWfn.Cbv)
    ea, eb = _lib.f64(refWfn.epsa), _lib.f64(refWfn.epsb)
    E = np.zeros(1)
    ctx.check(ctx.lib.jues_ump2(ctx.h, prec_code(prec), _lib.ptr(Cao), Cao.shape[1], _lib.ptr(Cav), Cav.shape[1],
                                _lib.ptr(Cbo), Cbo.shape[1], _lib.ptr(Cbv), Cbv.shape[1], n, _lib.ptr(ea),
                                _lib.ptr(eb), _lib.ptr(E)))
    return float(E[0])


def rmp2_grad(refWfn: Wfn, *, want_dg: bool = True, prec: str = "fp64", device: int = 0):
    """(E, dE/dC (n x n), dE/dg (n^4 or None), dE/deps (n)) with every entry of C, g and
    eps independent (eps occupied first, as ``refWfn.epsa``).

    ``dE/dg`` needs dense AO integrals; with a ``FactoredERI`` it is None and the
    gradient takes the factor route (no n^2 o v intermediate).
    """
    ctx = _lib.context(device)
    n = ensure_ao(ctx, refWfn.ao_eri)
    C = _lib.f64(refWfn.Ca)
    eps = _lib.f64(refWfn.epsa)
    E = np.zeros(1)
    dC = np.zeros((n, n), order="F")
    deps = np.zeros(n)
    dg = None
    if want_dg and not is_factored(refWfn.ao_eri):
        dg = np.zeros((n, n, n, n), order="F")
    ctx.check(ctx.lib.jues_rmp2_grad(ctx.h, prec_code(prec), _lib.ptr(C), refWfn.nalpha, n, _lib.ptr(eps),
                                     _lib.ptr(E), _lib.ptr(dC), _lib.ptr(dg), _lib.ptr(deps)))
    return float(E[0]), dC, dg, deps


__all__ = ["do_rmp2", "do_rmp2_noalloc", "do_df_rmp2", "do_ump2", "rmp2_grad"]
