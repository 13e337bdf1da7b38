"""Closed-shell coupled cluster on the GPU: the ``RCCD``, ``mRCCD``, ``RCCSD``,
``mRCCSD``, ``DFRCCD`` and ``PerturbativeTriples`` modules of CoupledCluster/
(RCCD.jl:33, mRCCD.jl:38, RCCSD.jl:33, mRCCSD.jl:35, DF-RCCD.jl:11,
PerturbativeTriples.jl:35).

Each ``do_rccd`` / ``do_rccsd`` takes a ``Wfn`` and returns the correlation
energy (and the amplitudes when ``return_T2`` / ``return_T`` is set), like the
reference.  Unlike the reference's fixed 40 sweeps, the iteration runs to
convergence (|dE| < cc_e_conv and ||dt||_2 < cc_max_rms) with DIIS by default;
``diis=False`` gives the plain Jacobi iteration of RCCD.jl:55-74.  Option names
follow CoupledCluster.jl:36-43.  ``prec="fp32_mixed"`` runs FP32-MFMA
contractions and then FP64 refinement iterations from the FP32 amplitudes.
"""
from __future__ import annotations

import ctypes as C
import types

import numpy as np

from . import _lib
from .Wavefunction import Wfn, ensure_ao, is_factored, prec_code

defaults = dict(cc_max_iter=100, cc_max_rms=1e-9, cc_e_conv=1e-10, diis=True, fcn=0)


def _opts(kw: dict, method_default_guess: str) -> _lib.CCOpts:
    o = _lib.default_opts()
    o.maxit = int(kw.pop("maxit", kw.pop("cc_max_iter", defaults["cc_max_iter"])))
    o.e_conv = float(kw.pop("cc_e_conv", kw.pop("e_conv", defaults["cc_e_conv"])))
    o.r_conv = float(kw.pop("cc_max_rms", kw.pop("r_conv", defaults["cc_max_rms"])))
    o.diis = int(bool(kw.pop("diis", defaults["diis"])))
    o.diis_max = int(kw.pop("diis_max", 5))
    g = kw.pop("guess", method_default_guess)
    o.guess = _lib.GUESS[g] if isinstance(g, str) else int(g)
    o.ccsd_eqs = _lib.CCSD_EQS[kw.pop("ccsd_eqs", "mRCCSD")]
    o.fcn = int(kw.pop("fcn", defaults["fcn"]))
    o.fp32_r_conv = float(kw.pop("fp32_r_conv", 1e-5))
    o.refine_maxit = int(kw.pop("refine_maxit", 60))
    o.verbose = int(bool(kw.pop("doprint", False)))
    o.coupling = float(kw.pop("coupling", 1.0))
    o.continuation = int(kw.pop("continuation", 1))
    lad = kw.pop("ladder", "auto")
    o.ladder = _lib.LADDER[lad] if isinstance(lad, str) else int(lad)
    ref = kw.pop("refine", "restart")
    o.refine = _lib.REFINE[ref] if isinstance(ref, str) else int(ref)
    return o


def _run(method: int, refWfn: Wfn, kw: dict, default_guess: str, fock=None):
    kw = dict(kw)
    prec = kw.pop("prec", "fp64")
    device = kw.pop("device", 0)
    return_T = kw.pop("return_T2", False) or kw.pop("return_T", False)
    callback = kw.pop("callback", None)
    check = kw.pop("check_convergence", True)
    T1_in = kw.pop("T1", None)
    T2_in = kw.pop("T2", None)
    opts = _opts(kw, default_guess)
    if kw:
        raise TypeError(f"unexpected keyword arguments {sorted(kw)}")
    ctx = _lib.context(device)
    n = ensure_ao(ctx, refWfn.ao_eri)
    o, v = refWfn.nalpha, refWfn.nvira
    oa = o - opts.fcn
    Co, Cv, eps = _lib.f64(refWfn.Cao), _lib.f64(refWfn.Cav), _lib.f64(refWfn.epsa)
    T2 = np.zeros((oa, oa, v, v), order="F") if T2_in is None else _lib.f64(T2_in).copy(order="F")
    T1 = np.zeros((oa, v), order="F") if T1_in is None else _lib.f64(T1_in).copy(order="F")
    if T2_in is not None:
        opts.guess = _lib.GUESS["user"]
    E = np.zeros(1)
    history = []

    def _cb(user, info):
        rec = info.contents
        history.append(dict(iter=rec.iter, stage=rec.stage, energy=rec.energy, delta_e=rec.delta_e,
                            rnorm=rec.rnorm, diis_size=rec.diis_size, ms=rec.ms, coupling=rec.coupling))
        if callback is not None:
            callback(history[-1])

    cb = _lib.ITER_CB(_cb)
    allow = (_lib.JUES_OK,) if check else (_lib.JUES_OK, _lib.JUES_NOT_CONVERGED)
    if method == 0:
        st = ctx.lib.jues_rccd(ctx.h, prec_code(prec), C.byref(opts), _lib.ptr(Co), o, _lib.ptr(Cv), v, n,
                               _lib.ptr(eps), _lib.ptr(E), _lib.ptr(T2), cb, None)
    elif fock is None:
        st = ctx.lib.jues_rccsd(ctx.h, prec_code(prec), C.byref(opts), _lib.ptr(Co), o, _lib.ptr(Cv), v, n,
                                _lib.ptr(eps), _lib.ptr(E), _lib.ptr(T1), _lib.ptr(T2), cb, None)
    else:
        f = _lib.f64(fock)
        st = ctx.lib.jues_rccsd_fock(ctx.h, prec_code(prec), C.byref(opts), _lib.ptr(Co), o, _lib.ptr(Cv), v, n,
                                     _lib.ptr(f), _lib.ptr(E), _lib.ptr(T1), _lib.ptr(T2), cb, None)
    ctx.check(st, allow)
    _run.last_history = history
    if return_T:
        return (float(E[0]), T2) if method == 0 else (float(E[0]), T1, T2)
    return float(E[0])


def _rccd(refWfn: Wfn, **kwargs):
    """RCCD.jl:33-83: closed-shell CCD; default guess (ij|ab)/D as RCCD.jl:145-160."""
    return _run(0, refWfn, kwargs, "ovov")


def _mrccd(refWfn: Wfn, **kwargs):
    """mRCCD.jl:38-118: closed-shell CCD with DIIS from a zero guess (mRCCD.jl:245)."""
    return _run(0, refWfn, kwargs, "zero")


def _rccsd(refWfn: Wfn, **kwargs):
    """RCCSD.jl:33-101: closed-shell CCSD (Fae T1 term of RCCSD.jl:189), MP2 guess."""
    kwargs.setdefault("ccsd_eqs", "RCCSD")
    return _run(1, refWfn, kwargs, "mp2")


def _mrccsd(refWfn: Wfn, **kwargs):
    """mRCCSD.jl:35-101: closed-shell CCSD (Fae T1 term of mRCCSD.jl:140), MP2 guess."""
    kwargs.setdefault("ccsd_eqs", "mRCCSD")
    return _run(1, refWfn, kwargs, "mp2")


def _df_rccd(refWfn: Wfn, maxit: int = 40, doprint: bool = False, return_T2: bool = False,
              dfbname: str = "default", *, prec: str = "fp64", device: int = 0, callback=None, **kwargs):
    """DF-RCCD.jl:11-53: CCD from the 3-index factors b (``DFIntegrals`` or ``FactoredERI``).

    As the reference: the MP2 guess t = (ia|jb)/D (``T2_init!``, :113-135),
    exactly ``maxit`` plain Jacobi sweeps with no DIIS and no convergence test
    (:27-46), and the energy of the last amplitudes (:47-52).  The pp-ladder is
    recomputed from bvv.bvv every sweep (:207); ``ladder="stored"`` builds V+-
    once instead (same result).  The residual is RCCD.jl's (:242-256): the last
    exchange term of WmBeJ is -1/2 sum (ne|mf) t[n,j,f,b] (RCCD.jl:400-401),
    where DF-RCCD.jl:250 reads (me|nf).  ``prec`` other than "fp64" runs the
    same ``maxit`` sweeps with the contractions in that form (amplitudes,
    energies and denominators stay FP64; there is no refinement stage).
    ``dfbname`` names the auxiliary basis in the reference's psi4 front end;
    here the basis is fixed by the integrals passed in, so only "default" is
    accepted.
    """
    if dfbname not in ("default", None):
        raise ValueError("dfbname: the auxiliary basis is the one the DFIntegrals were built in")
    if not is_factored(refWfn.ao_eri):
        raise TypeError("do_df_rccd needs DFIntegrals (pqP, Jpqh) or FactoredERI integrals")
    kw = dict(kwargs)
    kw.setdefault("ladder", "recompute")
    kw.update(maxit=maxit, doprint=doprint, diis=False, guess="mp2", cc_e_conv=0.0, cc_max_rms=0.0)
    solver = Solver(refWfn, method="rccd", prec=prec, device=device, **kw)
    history = []
    for _ in range(maxit):
        info, _ = solver.iterate(1, prec=prec)
        history.append(info)
        if callback is not None:
            callback(info)
    _run.last_history = history
    E, T2 = solver.result()
    return (E, T2) if return_T2 else E


def _compute_pT(*, T1, T2, Vvvvo, Vvooo, Vvovo, fo, fv, prec: str = "fp64", device: int = 0,
                doprint: bool = False) -> float:
    """PerturbativeTriples.compute_pT (PerturbativeTriples.jl:35-138) on the GPU.

    Same keyword arguments as the reference: the CCSD amplitudes and the integral
    arrays AutoRCCSD.jl:295-297 passes (Vvvvo v x v x v x o, Vvooo v x o x o x o,
    Vvovo v x o x v x o), fo/fv the diagonal Fock elements.  ``prec="fp32_mixed"``
    forms W in FP32 MFMA; the energy sums stay FP64.
    """
    ctx = _lib.context(device)
    T1, T2 = _lib.f64(T1), _lib.f64(T2)
    o, v = T1.shape
    arrs = [_lib.f64(x) for x in (Vvvvo, Vvooo, Vvovo, fo, fv)]
    shapes = [(v, v, v, o), (v, o, o, o), (v, o, v, o), (o,), (v,)]
    for a, sh, nm in zip(arrs, shapes, ("Vvvvo", "Vvooo", "Vvovo", "fo", "fv")):
        if a.shape != sh:
            raise ValueError(f"{nm} has shape {a.shape}, expected {sh}")
    if T2.shape != (o, o, v, v):
        raise ValueError(f"T2 has shape {T2.shape}, expected {(o, o, v, v)}")
    Et = np.zeros(1)
    ctx.check(ctx.lib.jues_compute_pT(ctx.h, prec_code(prec), _lib.ptr(T1), _lib.ptr(T2),
                                      *[_lib.ptr(a) for a in arrs], o, v, int(doprint), _lib.ptr(Et)))
    return float(Et[0])


def _rccsd_pT(refWfn: Wfn, T1, T2, *, prec: str = "fp64", device: int = 0, doprint: bool = False) -> float:
    """(T) after a closed-shell CCSD on canonical orbitals (AutoRCCSD.jl:294-300): the
    Vvvvo/Vvooo/Vvovo blocks are built on the device from ``refWfn.ao_eri`` (dense or
    factored) and fo/fv = epsa."""
    ctx = _lib.context(device)
    n = ensure_ao(ctx, refWfn.ao_eri)
    T1, T2 = _lib.f64(T1), _lib.f64(T2)
    o, v = refWfn.nalpha, refWfn.nvira
    if T1.shape != (o, v) or T2.shape != (o, o, v, v):
        raise ValueError("T1/T2 do not match the Wfn's occupied/virtual sizes (no frozen core here)")
    Co, Cv, eps = _lib.f64(refWfn.Cao), _lib.f64(refWfn.Cav), _lib.f64(refWfn.epsa)
    Et = np.zeros(1)
    ctx.check(ctx.lib.jues_rccsd_pT(ctx.h, prec_code(prec), _lib.ptr(Co), o, _lib.ptr(Cv), v, n, _lib.ptr(eps),
                                    _lib.ptr(T1), _lib.ptr(T2), int(doprint), _lib.ptr(Et)))
    return float(Et[0])


PerturbativeTriples = types.SimpleNamespace(compute_pT=_compute_pT, rccsd_pT=_rccsd_pT)


def _auto_rccsd(wfn: Wfn, *, do_pT: bool = False, return_T: bool = False, diis: bool = False, **kwargs):
    """AutoRCCSD.do_rccsd (AutoRCCSD.jl:193-300): CCSD driven by the MO Fock matrix of the
    Wfn (``get_fock``, needs ``wfn.hao``), so non-canonical and non-Hartree-Fock orbitals
    are handled: the resolvents take the Fock diagonal, the off-diagonal blocks enter the
    equations and 2 f_ia t_ia the energy; MP2 guess with T1 = f_ia / d_ia.  Plain
    updates by default, as the reference (``diis=True`` accelerates, same fixed point).
    The stopping test is |dE| < cc_e_conv and ||dt||_2 < cc_max_rms.  Returns the
    correlation energy, with the (T) correction as a second value when ``do_pT``
    (fo, fv = the Fock diagonal, :294-300), and the amplitudes when ``return_T``."""
    from .IntegralTransformation import get_fock
    import dataclasses
    device = kwargs.get("device", 0)
    fock = get_fock(wfn, "alpha", device=device)
    kw = dict(kwargs)
    kw.setdefault("ccsd_eqs", "RCCSD")
    kw.update(diis=diis, return_T=True)
    E, T1, T2 = _run(1, wfn, kw, "mp2", fock=fock)
    out = [E]
    if do_pT:
        fcn = int(kwargs.get("fcn", 0))
        if fcn:
            raise ValueError("do_pT with frozen core is not supported here")
        w2 = dataclasses.replace(wfn, epsa=np.diag(fock).copy())
        out.append(_rccsd_pT(w2, T1, T2, prec=kwargs.get("prec", "fp64"), device=device))
    if return_T:
        out += [T1, T2]
    return out[0] if len(out) == 1 else tuple(out)


AutoRCCSD = types.SimpleNamespace(do_rccsd=_auto_rccsd)

RCCD = types.SimpleNamespace(do_rccd=_rccd)
mRCCD = types.SimpleNamespace(do_rccd=_mrccd)
DFRCCD = types.SimpleNamespace(do_df_rccd=_df_rccd)
RCCSD = types.SimpleNamespace(do_rccsd=_rccsd)
mRCCSD = types.SimpleNamespace(do_rccsd=_mrccsd)


def last_history():
    """Per-iteration records of the most recent solve (iter, stage, energy, delta_e, rnorm, diis_size, ms)."""
    return list(getattr(_run, "last_history", []))


class Solver:
    """Stepwise access to the device iteration (jues_cc_setup / jues_cc_iterate):
    integrals and amplitudes stay resident between calls."""

    def __init__(self, refWfn: Wfn, method: str = "rccd", prec: str = "fp64", device: int = 0, **kw):
        self.ctx = _lib.context(device)
        n = ensure_ao(self.ctx, refWfn.ao_eri)
        self.method = 0 if method.lower() == "rccd" else 1
        kw = dict(kw)
        self.opts = _opts(kw, "mp2")
        if kw:
            raise TypeError(f"unexpected keyword arguments {sorted(kw)}")
        o, v = refWfn.nalpha, refWfn.nvira
        self.o, self.v = o - self.opts.fcn, v
        Co, Cv, eps = _lib.f64(refWfn.Cao), _lib.f64(refWfn.Cav), _lib.f64(refWfn.epsa)
        self.ctx.check(self.ctx.lib.jues_cc_setup(self.ctx.h, self.method, prec_code(prec), C.byref(self.opts),
                                                  _lib.ptr(Co), o, _lib.ptr(Cv), v, n, _lib.ptr(eps), None, None))

    def iterate(self, nsteps: int, prec: str = "fp64", stop_on_conv: bool = False):
        info = _lib.IterInfo()
        conv = C.c_int(0)
        self.ctx.check(self.ctx.lib.jues_cc_iterate(self.ctx.h, prec_code(prec), nsteps, int(stop_on_conv),
                                                    C.byref(info), C.byref(conv)))
        return dict(iter=info.iter, stage=info.stage, energy=info.energy, delta_e=info.delta_e, rnorm=info.rnorm,
                    diis_size=info.diis_size, ms=info.ms, coupling=info.coupling), bool(conv.value)

    def reset_diis(self):
        self.ctx.check(self.ctx.lib.jues_cc_reset_diis(self.ctx.h))

    def refine(self, inner_prec: str = "fp16x3", max_anchors: int = 60, inner_maxit: int = 100):
        """Defect-correction refinement (RCCD): FP64 iterations anchoring inner solves of the
        correction equation in inner_prec; returns the last FP64 iteration's record (with
        n_fp64, the number of FP64 iterations) and whether it converged."""
        info = _lib.IterInfo()
        conv = C.c_int(0)
        n64 = C.c_int(0)
        self.ctx.check(self.ctx.lib.jues_cc_refine(self.ctx.h, prec_code(inner_prec), max_anchors, inner_maxit,
                                                   C.byref(info), C.byref(conv), C.byref(n64)))
        return dict(iter=info.iter, stage=info.stage, energy=info.energy, delta_e=info.delta_e, rnorm=info.rnorm,
                    diis_size=info.diis_size, ms=info.ms, coupling=info.coupling, n_fp64=n64.value), bool(conv.value)

    def energy(self) -> float:
        """Correlation energy of the resident amplitudes (no amplitude copy to the host)."""
        E = np.zeros(1)
        self.ctx.check(self.ctx.lib.jues_cc_get(self.ctx.h, _lib.ptr(E), None, None))
        return float(E[0])

    def result(self):
        E = np.zeros(1)
        T2 = np.zeros((self.o, self.o, self.v, self.v), order="F")
        T1 = np.zeros((self.o, self.v), order="F")
        self.ctx.check(self.ctx.lib.jues_cc_get(self.ctx.h, _lib.ptr(E), _lib.ptr(T1), _lib.ptr(T2)))
        return (float(E[0]), T2) if self.method == 0 else (float(E[0]), T1, T2)


__all__ = ["RCCD", "mRCCD", "RCCSD", "mRCCSD", "DFRCCD", "AutoRCCSD", "PerturbativeTriples", "Solver",
           "defaults", "last_history"]
