"""ctypes binding of libjues_hip.so (the C ABI declared in include/jues.h).

The library is required: if it is missing or cannot be loaded every entry point
raises ``JuesLibraryError``; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib
import threading

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = pathlib.Path(os.environ.get("JUES_HIP_LIB", _HERE.parent / "lib" / "libjues_hip.so"))

JUES_OK, JUES_NOT_CONVERGED, JUES_NONFINITE, JUES_OOM, JUES_BAD_ARG, JUES_HIP_ERROR = range(6)
NOALLOC_SCRATCH_BYTES = 512 << 20  # JUES_NOALLOC_SCRATCH_BYTES
STATUS_NAMES = {0: "OK", 1: "NOT_CONVERGED", 2: "NONFINITE", 3: "OOM", 4: "BAD_ARG", 5: "HIP_ERROR"}
FP64, FP32_MIXED, FP16X3_MIXED = 0, 1, 2
LADDER = {"auto": 0, "stored": 1, "recompute": 2}
REFINE = {"restart": 0, "defect": 1}
GUESS = {"zero": 0, "mp2": 1, "ovov": 2, "user": 3}
CCSD_EQS = {"mRCCSD": 0, "RCCSD": 1}

# names every build must export (checked by the CPU test-suite against include/jues.h)
EXPORTS = (
    "jues_create", "jues_destroy", "jues_last_error", "jues_cc_default_opts", "jues_device_bytes",
    "jues_release", "jues_set_ao_eri", "jues_set_ao_factors", "jues_set_df_integrals", "jues_get_ao_factors",
    "jues_tei_transform", "jues_get_eri",
    "jues_rmp2", "jues_rmp2_noalloc", "jues_rmp2_grad", "jues_rccd", "jues_rccsd", "jues_cc_setup", "jues_cc_iterate",
    "jues_cc_refine", "jues_cc_reset_diis", "jues_cc_get", "jues_kernel_stats", "jues_kernel_stats_reset", "jues_gemm_test", "jues_gemm_pair_test",
    "jues_compute_pT", "jues_rccsd_pT", "jues_ump2", "jues_get_fock", "jues_rccsd_fock",
)


class JuesLibraryError(RuntimeError):
    pass


class JuesError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"[{STATUS_NAMES.get(status, status)}] {msg}")
        self.status = status


class CCOpts(C.Structure):
    _fields_ = [
        ("maxit", C.c_int32), ("e_conv", C.c_double), ("r_conv", C.c_double), ("diis", C.c_int32),
        ("diis_max", C.c_int32), ("guess", C.c_int32), ("ccsd_eqs", C.c_int32), ("fcn", C.c_int32),
        ("fp32_r_conv", C.c_double), ("refine_maxit", C.c_int32), ("verbose", C.c_int32),
        ("coupling", C.c_double), ("continuation", C.c_int32), ("ladder", C.c_int32), ("refine", C.c_int32),
    ]


class IterInfo(C.Structure):
    _fields_ = [
        ("iter", C.c_int32), ("stage", C.c_int32), ("energy", C.c_double), ("delta_e", C.c_double),
        ("rnorm", C.c_double), ("diis_size", C.c_int32), ("ms", C.c_double), ("coupling", C.c_double),
    ]


ITER_CB = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(IterInfo))

_dp = C.POINTER(C.c_double)
_ctxp = C.c_void_p
_lib = None
_lock = threading.Lock()


def load(path: os.PathLike | str | None = None) -> C.CDLL:
    """Load (once) and return the library with prototypes attached."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        p = pathlib.Path(path) if path else LIB_PATH
        if not p.exists():
            raise JuesLibraryError(f"libjues_hip.so not found at {p}; build it with "
                                   f"`python jues.jl_amd/build_lib.py`")
        try:
            lib = C.CDLL(str(p))
        except OSError as e:  # pragma: no cover
            raise JuesLibraryError(f"cannot load {p}: {e}") from e
        i, d = C.c_int, C.c_double
        proto = {
            "jues_create": (i, [C.POINTER(_ctxp), i]),
            "jues_destroy": (None, [_ctxp]),
            "jues_last_error": (C.c_char_p, [_ctxp]),
            "jues_cc_default_opts": (None, [C.POINTER(CCOpts)]),
            "jues_device_bytes": (C.c_size_t, [_ctxp]),
            "jues_release": (i, [_ctxp]),
            "jues_set_ao_eri": (i, [_ctxp, _dp, i]),
            "jues_set_ao_factors": (i, [_ctxp, _dp, i, i]),
            "jues_set_df_integrals": (i, [_ctxp, _dp, _dp, i, i]),
            "jues_get_ao_factors": (i, [_ctxp, _dp, i, i]),
            "jues_compute_pT": (i, [_ctxp, i, _dp, _dp, _dp, _dp, _dp, _dp, _dp, i, i, i, _dp]),
            "jues_rccsd_pT": (i, [_ctxp, i, _dp, i, _dp, i, i, _dp, _dp, _dp, i, _dp]),
            "jues_ump2": (i, [_ctxp, i, _dp, i, _dp, i, _dp, i, _dp, i, i, _dp, _dp, _dp]),
            "jues_get_fock": (i, [_ctxp, _dp, _dp, i, _dp, i, i, _dp]),
            "jues_tei_transform": (i, [_ctxp, i, _dp, i, _dp, i, _dp, i, _dp, i, i, _dp]),
            "jues_get_eri": (i, [_ctxp, i, C.c_char_p, C.c_char_p, _dp, i, _dp, i, _dp, i, _dp, i, i, i, _dp]),
            "jues_rmp2": (i, [_ctxp, i, _dp, i, _dp, i, i, _dp, _dp]),
            "jues_rmp2_noalloc": (i, [_ctxp, i, _dp, i, _dp, i, i, _dp, _dp]),
            "jues_rmp2_grad": (i, [_ctxp, i, _dp, i, i, _dp, _dp, _dp, _dp, _dp]),
            "jues_rccd": (i, [_ctxp, i, C.POINTER(CCOpts), _dp, i, _dp, i, i, _dp, _dp, _dp, ITER_CB, C.c_void_p]),
            "jues_rccsd": (i, [_ctxp, i, C.POINTER(CCOpts), _dp, i, _dp, i, i, _dp, _dp, _dp, _dp, ITER_CB,
                               C.c_void_p]),
            "jues_rccsd_fock": (i, [_ctxp, i, C.POINTER(CCOpts), _dp, i, _dp, i, i, _dp, _dp, _dp, _dp, ITER_CB,
                                    C.c_void_p]),
            "jues_cc_setup": (i, [_ctxp, i, i, C.POINTER(CCOpts), _dp, i, _dp, i, i, _dp, _dp, _dp]),
            "jues_cc_iterate": (i, [_ctxp, i, i, i, C.POINTER(IterInfo), C.POINTER(C.c_int)]),
            "jues_cc_reset_diis": (i, [_ctxp]),
            "jues_cc_refine": (i, [_ctxp, i, i, i, C.POINTER(IterInfo), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "jues_cc_get": (i, [_ctxp, _dp, _dp, _dp]),
            "jues_kernel_stats": (i, [_ctxp, C.POINTER(C.c_char_p), _dp, _dp, _dp, _dp, i]),
            "jues_kernel_stats_reset": (None, [_ctxp, i]),
            "jues_gemm_test": (i, [_ctxp, i, C.c_char, C.c_char, i, i, i, d, _dp, i, _dp, i, d, _dp, i]),
            "jues_gemm_pair_test": (i, [_ctxp, i, C.c_char, i, i, i, _dp, i, _dp, i, _dp, i, i, i, i, _dp, i, _dp, i,
                                        _dp, i, d, d]),
        }
        for name, (res, args) in proto.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
        return lib


def f64(a: np.ndarray) -> np.ndarray:
    """Fortran-ordered float64 copy/view (the C ABI's memory order)."""
    return np.asfortranarray(a, dtype=np.float64)


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim <= 1)
    return a.ctypes.data_as(_dp)


def default_opts(**kw) -> CCOpts:
    o = CCOpts()
    load().jues_cc_default_opts(C.byref(o))
    for k, v in kw.items():
        if k == "guess" and isinstance(v, str):
            v = GUESS[v]
        if k == "ccsd_eqs" and isinstance(v, str):
            v = CCSD_EQS[v]
        setattr(o, k, v)
    return o


class Context:
    """One device context (one HIP stream and its device buffers)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        st = self.lib.jues_create(C.byref(h), device)
        if st != JUES_OK:
            raise JuesError(st, f"jues_create(device={device}) failed")
        self.h = h
        self.device = device
        self.ao_key = None  # identity of the AO integrals currently on the device

    def close(self):
        if getattr(self, "h", None):
            self.lib.jues_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def release(self):
        """Free every device buffer the context holds (integrals, CC state, scratch)."""
        self.check(self.lib.jues_release(self.h))
        self.ao_key = None
        self.ao_ref = None

    def check(self, st: int, allow=(JUES_OK,)) -> int:
        if st not in allow:
            msg = self.lib.jues_last_error(self.h)
            raise JuesError(st, msg.decode() if msg else "")
        return st

    # -- integrals -----------------------------------------------------------
    def set_ao_eri(self, gao: np.ndarray, key=None):
        g = f64(gao)
        n = g.shape[0]
        assert g.shape == (n, n, n, n)
        self.check(self.lib.jues_set_ao_eri(self.h, ptr(g), n))
        self.ao_key = key

    def set_ao_factors(self, B: np.ndarray, key=None):
        b = f64(B)
        n, n2, naux = b.shape
        assert n == n2
        self.check(self.lib.jues_set_ao_factors(self.h, ptr(b), n, naux))
        self.ao_key = key

    def set_df_integrals(self, pqP: np.ndarray, Jpqh: np.ndarray, key=None):
        t, J = f64(pqP), f64(Jpqh)
        n, n2, naux = t.shape
        if n != n2 or J.shape != (naux, naux):
            raise ValueError(f"pqP {t.shape} and Jpqh {J.shape} do not match")
        self.check(self.lib.jues_set_df_integrals(self.h, ptr(t), ptr(J), n, naux))
        self.ao_key = key

    def get_ao_factors(self, n: int, naux: int) -> np.ndarray:
        B = np.zeros((n, n, naux), order="F")
        self.check(self.lib.jues_get_ao_factors(self.h, ptr(B), n, naux))
        return B

    # -- kernel timing ---------------------------------------------------------
    def stats_reset(self, enable: bool = True):
        self.lib.jues_kernel_stats_reset(self.h, int(enable))

    def stats(self) -> dict:
        cap = 256
        names = (C.c_char_p * cap)()
        ms, calls, fl, by = (np.zeros(cap) for _ in range(4))
        k = self.lib.jues_kernel_stats(self.h, names, ptr(ms), ptr(calls), ptr(fl), ptr(by), cap)
        return {names[i].decode(): dict(ms=ms[i], calls=calls[i], flops=fl[i], bytes=by[i]) for i in range(k)}

    def gemm_test(self, prec, ta, tb, alpha, A, B, beta, Cm, K=None):
        """C = alpha op(A) op(B) + beta C with leading dimensions A.shape[0], B.shape[0]; an
        explicit K uses only the first K rows of a transposed A or untransposed B."""
        A, B, Cm = f64(A), f64(B), f64(Cm).copy(order="F")
        M, N = Cm.shape
        if K is None:
            K = A.shape[0] if ta == "T" else A.shape[1]
        self.check(self.lib.jues_gemm_test(self.h, prec, ta.encode(), tb.encode(), M, N, K, alpha, ptr(A),
                                           A.shape[0], ptr(B), B.shape[0], beta, ptr(Cm), Cm.shape[0]))
        return Cm


    def gemm_pair_test(self, prec, tb, alpha, A0, B0, C0, A1, B1, C1, beta):
        """Both products of one launch: returns (C0', C1')."""
        A0, B0, A1, B1 = f64(A0), f64(B0), f64(A1), f64(B1)
        C0, C1 = f64(C0).copy(order="F"), f64(C1).copy(order="F")
        (M0, N0), (M1, N1) = C0.shape, C1.shape
        self.check(self.lib.jues_gemm_pair_test(self.h, prec, tb.encode(), M0, N0, A0.shape[1], ptr(A0), A0.shape[0],
                                                ptr(B0), B0.shape[0], ptr(C0), C0.shape[0], M1, N1, A1.shape[1],
                                                ptr(A1), A1.shape[0], ptr(B1), B1.shape[0], ptr(C1), C1.shape[0],
                                                alpha, beta))
        return C0, C1


_ctx: dict[int, Context] = {}


def context(device: int = 0) -> Context:
    """The process-wide context for a device (created on first use)."""
    c = _ctx.get(device)
    if c is None:
        c = _ctx[device] = Context(device)
    return c
