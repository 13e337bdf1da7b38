"""CPU baseline timing for the RCCD iteration -- TEST/BENCH INFRASTRUCTURE ONLY.

Only ``bench.py``'s ``cpu_baseline`` leg uses this module.  It times, with
NumPy/OpenBLAS on the host cores, the GEMM sequence one mRCCD.jl iteration
performs (mRCCD.jl:159-166 with form_T2 :301-394 and the intermediates
:265-488): Fae, Fmi, Wmnij, Wabef (2 o^2 v^4 build), WmBeJ (2 GEMMs), WmBEj,
and in form_T2 two Fae, two Fmi, the Wmnij and Wabef (2 o^2 v^4) applications
and five ring GEMMs -- 4 o^2 v^4 + 16 o^3 v^3 + O(o^4 v^2) flops in all.

At o=60, v=340 one iteration is ~3.3e14 flops (minutes on a host) and its
Wabef alone is 107 GB, so the sample keeps every GEMM's shape except that the
output columns of the O(o^2 v^4) and O(o^3 v^3) GEMMs are cut to 1/k; the
sampled time is scaled by k.  The permutation copies mRCCD.jl performs between
GEMMs are not timed, so the figure is a lower bound on the reference's CPU time.
Operand values are seeded random numbers (timing does not depend on them).
"""
from __future__ import annotations

import os
import time

import numpy as np


def _blas_threads() -> int:
    try:
        import threadpoolctl
        info = threadpoolctl.threadpool_info()
        n = [d.get("num_threads", 0) for d in info if d.get("user_api") == "blas"]
        if n:
            return int(max(n))
    except Exception:
        pass
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def host_info() -> dict:
    """CPU model, sockets and RAM of the host the baseline runs on (from /proc)."""
    info = {}
    try:
        txt = open("/proc/cpuinfo").read()
        models = [ln.split(":", 1)[1].strip() for ln in txt.splitlines() if ln.startswith("model name")]
        if models:
            info["cpu_model"] = models[0]
        info["sockets"] = len({ln for ln in txt.splitlines() if ln.startswith("physical id")}) or 1
    except OSError:
        pass
    try:
        for ln in open("/proc/meminfo"):
            if ln.startswith("MemTotal"):
                info["ram_gib"] = round(int(ln.split()[1]) / 2**20, 1)
    except OSError:
        pass
    return info


def rccd_iteration_sample(o: int, v: int, k: int = 32, seed: int = 0) -> dict:
    rng = np.random.default_rng(seed)
    ov, oo, vv = o * v, o * o, v * v
    R = lambda *s: rng.random(s)  # noqa: E731
    # (label, M, N, K, cut): cut GEMMs use N // k output columns
    gemms = [
        ("Fae", v, v, oo * v, False), ("Fmi", o, o, ov * v, False),
        ("Wmnij", oo, oo, vv, False), ("Wabef", vv, vv, oo, True),
        ("WmBeJ_1", ov, ov, ov, True), ("WmBeJ_2", ov, ov, ov, True), ("WmBEj", ov, ov, ov, True),
        ("T2_Fae_1", oo * v, v, v, False), ("T2_Fae_2", oo * v, v, v, False),
        ("T2_Fmi_1", ov * v, o, o, False), ("T2_Fmi_2", ov * v, o, o, False),
        ("T2_Wmnij", oo, vv, oo, False), ("T2_Wabef", oo, vv, vv, True),
        ("ring_1", ov, ov, ov, True), ("ring_2", ov, ov, ov, True), ("ring_3", ov, ov, ov, True),
        ("ring_4", ov, ov, ov, True), ("ring_5", ov, ov, ov, True),
    ]
    total = 0.0
    flops = 0.0
    for label, M, N, K, cut in gemms:
        n = max(1, N // k) if cut else N
        A = R(M, K)
        B = R(K, n)
        t0 = time.perf_counter()
        C = A @ B
        dt = time.perf_counter() - t0
        del A, B, C
        scale = (N / n) if cut else 1.0
        total += dt * scale
        flops += 2.0 * M * N * K
    return dict(seconds_per_iter=total, flops_per_iter=flops, threads=_blas_threads(),
                sample=f"mRCCD.jl GEMM sequence at o={o}, v={v}; O(o^2v^4)/O(o^3v^3) GEMM outputs cut "
                       f"to 1/{k} of their columns and the time scaled by {k}; permutation copies not timed")
