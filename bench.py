"""RCCD iteration benchmark on one MI355X: FP32-mixed and split-FP16 vs plain FP64.

Workload: closed-shell RCCD at o=60, v=340 (n=400, naux=1200) on seeded
synthetic B-factor integrals; every MO integral block (including the packed
v^4 ladder block) and the amplitudes are resident in HBM before timing.  A
"step" is one full CC iteration (all contractions, the fused update, DIIS).
The FP32-mixed stage (FP32 MFMA contractions, FP64 update/energy/DIIS) is timed
first, then the split-FP16 stage (ring and pp-ladder products on FP16 MFMA with
hi+lo operand splitting, FP32 accumulation), then the plain FP64 stage, each from
the amplitudes the previous one left and with the DIIS history reset.  `value` is
the FP32-mixed seconds per iteration; the split-FP16 figures are reported beside
it under "fp16x3".

The synthetic integrals at o=60, v=340 are so strongly coupled that the CC
iteration has no convergent solution at full strength (E_MP2 ~ -1e4 Eh; DIIS and
coupling continuation both diverge).  Every residual term carries exactly one
integral, so the run uses the equations for integrals scaled by a coupling
lambda (default 0.005, reported in `config`), where the iteration converges; all
shapes, flops and bytes are those of the full workload and the cost of an
iteration does not depend on the values.  `--converge` runs both forms to
convergence and reports |dE| (FP32-mixed or split-FP16 + FP64 refinement vs
plain FP64).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload cfg4_rccd_n400]
                       [--coupling 0.005] [--converge] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "jues.jl_amd"))

METRIC = "RCCD s/iter + MFMA TFLOP/s, FP32-mixed vs FP64, |ΔE|≤1e-8 Eh, o=60 v=340"
PEAK = {"f64": 78.6, "f32": 157.3, "f16": 2500.0}  # dense MFMA TFLOP/s (MI355X datasheet)
TRAFFIC = ROOT / "profiles" / "rccd_o60_v340_hbm_traffic.json"


def traffic(stage: str, tag: str, workload: str):
    """HBM bytes per launch of a kernel, from the committed FETCH_SIZE / WRITE_SIZE
    profile of this workload (None when the profile does not cover it)."""
    if workload != "cfg4_rccd_n400" or not TRAFFIC.exists():
        return None
    rec = json.loads(TRAFFIC.read_text())["per_launch"].get(f"{stage}:{tag}")
    return None if rec is None else rec["total_bytes"]


def f_rccd(o: int, v: int) -> float:
    """Canonical RCCD flops per iteration (pp-ladder + ring + hh + Fock-like terms)."""
    return 2 * o**2 * v**4 + 12 * o**3 * v**3 + 4 * o**4 * v**2 + 4 * o**2 * v**3 + 4 * o**3 * v**2


def roofline(stats: dict, prec: str, workload: str) -> dict:
    """Roofline of the dominant kernel class (largest total device time)."""
    gemms = {k: s for k, s in stats.items() if s["flops"] > 0}
    name, s = max(gemms.items(), key=lambda kv: kv[1]["ms"])
    per_launch_flops = s["flops"] / s["calls"]
    avg_s = s["ms"] / s["calls"] / 1e3
    achieved = per_launch_flops / avg_s / 1e12
    return {"kernel": name, "bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK[prec],
            "unit": "TFLOP/s", "frac": round(achieved / PEAK[prec], 4), "traffic": traffic(prec, name, workload),
            "flops_per_launch": per_launch_flops, "avg_launch_ms": round(avg_s * 1e3, 3),
            "launches": int(s["calls"])}


def roofline_split(stats: dict, workload: str) -> dict:
    """Roofline of the dominant split-FP16 GEMM: each FP32-equivalent flop executes
    three FP16 MFMA products, priced against the dense FP16 peak."""
    g = {k: s for k, s in stats.items() if s["flops"] > 0 and k in ("cc_ring", "cc_pp_ladder")}
    name, s = max(g.items(), key=lambda kv: kv[1]["ms"])
    avg_s = s["ms"] / s["calls"] / 1e3
    eq = s["flops"] / s["calls"] / avg_s / 1e12
    return {"kernel": name, "bound": "mfma", "achieved": round(3 * eq, 2), "peak": PEAK["f16"], "unit": "TFLOP/s",
            "frac": round(3 * eq / PEAK["f16"], 4), "fp32_equivalent_tflops": round(eq, 2),
            "fp32_equivalent_frac_of_f32_peak": round(eq / PEAK["f32"], 4), "traffic": traffic("f16x3", name, workload),
            "flops_per_launch": s["flops"] / s["calls"], "avg_launch_ms": round(avg_s * 1e3, 3),
            "launches": int(s["calls"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cfg4_rccd_n400")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-k", type=int, default=32)
    ap.add_argument("--coupling", type=float, default=0.005)
    ap.add_argument("--converge", action="store_true")
    ap.add_argument("--fast-r-conv", type=float, default=1e-6,
                    help="--converge: the fast stage stops once ||dt|| is below this (or it stalls); at o=60 "
                         "v=340, 1e-6 beats both 1e-5 and 1e-7 in time to solution")
    ap.add_argument("--fast-r-conv-defect", type=float, default=1e-5,
                    help="--converge: the same for the defect-correction hand-over, whose refinement costs the "
                         "same wherever the fast stage stops (at o=60 v=340, 1e-5 beats 1e-6 and 1e-7)")
    ap.add_argument("--keep-diis", action="store_true",
                    help="--converge: keep the fast stage's DIIS history for the FP64 refinement")
    ap.add_argument("--forms", default="fp64,fp32_mixed,fp16x3", help="--converge: forms to run")
    ap.add_argument("--refine", default="restart,defect",
                    help="--converge: FP64 hand-overs to time after each fast stage (restart: FP64 iterations "
                         "with a fresh DIIS history; defect: defect correction)")
    ap.add_argument("--diis-max", type=int, default=8)
    ap.add_argument("--ladder", default="auto", choices=["auto", "stored", "recompute"],
                    help="pp-ladder integrals stored (packed V+/V-) or rebuilt from the factors every iteration")
    ap.add_argument("--mp2", action="store_true",
                    help="instead of RCCD: AO->MO transform, RMP2 and the MP2 adjoint at n=256 (dense AO)")
    ap.add_argument("--df", action="store_true",
                    help="instead of RCCD: density fitting at o=60 v=340 naux=1200 (make_bmunu, DF-RMP2, "
                         "DF-RCCD sweeps)")
    ap.add_argument("--triples", action="store_true",
                    help="instead of RCCD: (T) on the o=60 v=340 factors with seeded amplitudes")
    ap.add_argument("--pt-o", type=int, default=60, help="--triples: occupied orbitals used (first pt-o)")
    ap.add_argument("--pt-prec", default="fp64,fp32_mixed", help="--triples: forms to run")
    ap.add_argument("--n4", action="store_true",
                    help="instead of RCCD: UMP2, get_fock and Fock-driven CCSD iterations at o=60 v=340")
    ap.add_argument("--cfg2", action="store_true",
                    help="instead of RCCD: RCCSD at n=120 o=20 v=100 naux=360 (coupling 0.02) in the three forms, "
                         "per-iteration time and |dE| against the oracle's fixed point after each stage")
    ap.add_argument("--rccsd5", action="store_true",
                    help="instead of RCCD: RCCSD at o=80 v=420 (n=500, naux=1500), recomputed ladder, factored "
                         "(ia|bc); s/iter of the three forms")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one MI355X only; the iteration is not sharded across GPUs",
              file=sys.stderr)
        sys.exit(2)
    if args.mp2:
        print(json.dumps(mp2_bench()), flush=True)
        return
    if args.df:
        print(json.dumps(df_bench(args.coupling)), flush=True)
        return
    if args.triples:
        print(json.dumps(triples_bench(args.pt_o, args.pt_prec.split(","))), flush=True)
        return
    if args.n4:
        print(json.dumps(n4_bench()), flush=True)
        return
    if args.cfg2:
        print(json.dumps(cfg2_bench()), flush=True)
        return
    if args.rccsd5 and args.converge:
        # RCCSD at o=80 v=420 to convergence (the restart hand-over; RCCSD has no defect correction)
        from jues import synthetic
        from jues.Wavefunction import Wfn
        w5 = Wfn.from_synthetic(synthetic.workload("cfg5_rccsd_n500"), dense=False)
        res = converge(w5, args.coupling, args.fast_r_conv, args.keep_diis, args.forms.split(","), ("restart",),
                       method="rccsd", diis_max=min(args.diis_max, 3), ladder="recompute")
        print(json.dumps({"workload": "RCCSD o=80 v=420 n=500 naux=1500, recomputed ladder, (ia|bc) factored",
                          "convergence": res}), flush=True)
        return
    if args.rccsd5:
        print(json.dumps(rccsd5_bench(args.coupling, args.steps, args.warmup, args.diis_max)), flush=True)
        return

    from jues import synthetic, _lib
    from jues.CoupledCluster import Solver
    from jues.Wavefunction import Wfn

    s = synthetic.workload(args.workload)
    o, v = s.o, s.v
    w = Wfn.from_synthetic(s, dense=False)
    t0 = time.perf_counter()
    solver = Solver(w, method="rccd", prec="fp32_mixed", guess="zero", diis=True, diis_max=args.diis_max,
                    coupling=args.coupling, ladder=args.ladder)
    t_setup = time.perf_counter() - t0
    ctx = solver.ctx
    print(f"[bench] setup {t_setup:.1f} s, device bytes {ctx.lib.jues_device_bytes(ctx.h) / 2**30:.1f} GiB",
          file=sys.stderr, flush=True)

    res = {}
    # every timed step runs with the full DIIS history (diis_max - 1 old pairs), as mid-solve:
    # untimed iterations fill the subspace before the warmup steps
    fill = max(0, args.diis_max - 1 - args.warmup)
    for stage, prec in (("fp32", "fp32_mixed"), ("fp16x3", "fp16x3"), ("fp64", "fp64")):
        solver.reset_diis()
        if fill + args.warmup:
            solver.iterate(fill + args.warmup, prec=prec)
        ctx.stats_reset(True)
        t0 = time.perf_counter()
        info, _ = solver.iterate(args.steps, prec=prec)   # synchronous: returns after the stream drains
        dt = time.perf_counter() - t0
        st = ctx.stats()
        ctx.stats_reset(False)
        res[stage] = dict(s_per_iter=dt / args.steps, info=info, stats=st)
        print(f"[bench] {stage}: {dt / args.steps:.3f} s/iter, E={info['energy']:.10f} "
              f"|dt|={info['rnorm']:.3e}", file=sys.stderr, flush=True)

    F = f_rccd(o, v)
    m, d, h = res["fp32"], res["fp64"], res["fp16x3"]
    executed = lambda st: sum(x["flops"] for x in st.values())  # noqa: E731
    out = {
        "metric": METRIC,
        "value": round(m["s_per_iter"], 4),
        "unit": "s/iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(m["s_per_iter"] * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded B-factor ERIs, QR orbitals, sorted uniform orbital energies)",
        "config": {"workload": f"RCCD o={o} v={v} n={s.n} naux={s.naux}, B-factor integrals, "
                               "vvvv packed and resident in HBM",
                   "o": o, "v": v, "n": s.n, "naux": s.naux, "diis_max": args.diis_max, "coupling": args.coupling,
                   "ladder": args.ladder, "diis_fill": fill},
        "fp64_s_per_iter": round(d["s_per_iter"], 4),
        "speedup_mixed_vs_fp64": round(d["s_per_iter"] / m["s_per_iter"], 3),
        # canonical flops over time: the packed ladders execute about a quarter of their canonical
        # flops, so this effective rate can exceed the MFMA peak (tflops_executed is the MFMA work)
        "tflops_effective_canonical": {"fp32_mixed": round(F / m["s_per_iter"] / 1e12, 2),
                                       "fp64": round(F / d["s_per_iter"] / 1e12, 2), "flops_per_iter": F,
                                       "note": "canonical F_RCCD / time; not an MFMA rate"},
        "tflops_executed": {"fp32_mixed": round(executed(m["stats"]) / args.steps / m["s_per_iter"] / 1e12, 2),
                            "fp64": round(executed(d["stats"]) / args.steps / d["s_per_iter"] / 1e12, 2),
                            "flops_per_iter": executed(m["stats"]) / args.steps},
        "fp16x3": {"s_per_iter": round(h["s_per_iter"], 4),
                   "speedup_vs_fp64": round(d["s_per_iter"] / h["s_per_iter"], 3),
                   "tflops_effective_canonical": round(F / h["s_per_iter"] / 1e12, 2),
                   "roofline": roofline_split(h["stats"], args.workload)},
        "roofline": roofline(m["stats"], "f32", args.workload),
        "roofline_fp64": roofline(d["stats"], "f64", args.workload),
        "kernel_ms_per_iter": {"fp32": {k: round(x["ms"] / args.steps, 2) for k, x in m["stats"].items()},
                               "fp16x3": {k: round(x["ms"] / args.steps, 2) for k, x in h["stats"].items()},
                               "fp64": {k: round(x["ms"] / args.steps, 2) for k, x in d["stats"].items()}},
        "setup_s": round(t_setup, 2),
    }
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import cpu_baseline
        cb = cpu_baseline.rccd_iteration_sample(o, v, k=args.cpu_k)
        out["cpu_baseline"] = {"value": round(cb["seconds_per_iter"], 2), "unit": "s/iter",
                               "cores": cb["threads"], "kind": "port", "sample": cb["sample"],
                               "host_cpus": os.cpu_count(), **cpu_baseline.host_info()}
    if args.converge:
        out["convergence"] = converge(w, args.coupling, args.fast_r_conv, args.keep_diis, args.forms.split(","),
                                      args.refine.split(","), fast_r_conv_defect=args.fast_r_conv_defect)
    print(json.dumps(out), flush=True)


def mp2_flop_counts(n: int, o: int, v: int) -> tuple[int, int]:
    """Flops the dense (ia|jb) transform and the dense MP2 adjoint with dE/dg execute (GEMM products).

    Transform, in the order its dispatch takes for (Co, Cv, Co, Cv): for o < v mu, then lam (both
    occupied; lam batched over sig), then sig and nu; for o > v the mirror image (sig, nu, mu, lam);
    for o = v mu, sig, lam, nu.  Adjoint: its forward contracts sig second, because it keeps
    (i nu|lam b) for position 3 (F_T); then the half-transform (mu a|j sig) (2n^4 o + 2n^3 o v); the four dE/dC positions: three products of
    2n^2 o^2 v (positions 1 and 3, the occupied quarter of position 4) and four of 2n o^2 v^2; the
    back-transform to dE/dg, i contracted last (2n^4 o + 2n^3 o v + 2n^2 o^2 v).
    """
    if o < v:
        F_X = 2 * n**4 * o + 2 * n**3 * o**2 + 2 * n**2 * o**2 * v + 2 * n * o**2 * v**2
    elif o > v:
        F_X = 2 * n**4 * v + 2 * n**3 * v**2 + 2 * n**2 * o * v**2 + 2 * n * o**2 * v**2
    else:
        F_X = 2 * n**4 * o + 2 * n**3 * o * v + 2 * n**2 * o**2 * v + 2 * n * o**2 * v**2
    F_T = 2 * n**4 * o + 2 * n**3 * o * v + 2 * n**2 * o**2 * v + 2 * n * o**2 * v**2
    F_G = F_T + 4 * n**4 * o + 4 * n**3 * o * v + 8 * n**2 * o**2 * v + 8 * n * o**2 * v**2
    return F_X, F_G


def mp2_bench() -> dict:
    """Transform, RMP2 energy and MP2 adjoint at n=256, o=40, v=216 on dense AO integrals
    (34.4 GB FP64 resident in HBM), FP64 and FP32-mixed; device times from HIP events."""
    import numpy as np
    from jues import synthetic, MollerPlesset, Transformation, _lib
    from jues.Wavefunction import Wfn, ensure_ao
    s = synthetic.workload("cfg3_mp2grad_n256")
    n, o, v = s.n, s.o, s.v
    t0 = time.perf_counter()
    Bm = s.B.reshape(n * n, s.naux, order="F")
    # symmetric n^2 x n^2 product (its transpose is the F-ordered tensor), in row blocks: one BLAS
    # call with a 4.3e9-element output crashes the host BLAS
    G2 = np.empty((n * n, n * n))
    for r0 in range(0, n * n, 4096):
        G2[r0:r0 + 4096] = Bm[r0:r0 + 4096] @ Bm.T
    g = G2.T.reshape((n, n, n, n), order="F")
    print(f"[bench] dense AO integrals built on the host in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)
    ctx = _lib.context(0)
    w = Wfn.from_orbitals(s.C, s.eps, s.o, g)
    t0 = time.perf_counter()
    ensure_ao(ctx, g)
    t_up = time.perf_counter() - t0
    F_X, F_G = mp2_flop_counts(n, o, v)
    out = {"workload": f"RMP2 + adjoint, n={n} o={o} v={v}, dense AO integrals resident in HBM "
                       f"({n**4 * 8 / 1e9:.1f} GB FP64)", "upload_s": round(t_up, 2), "stages": {}}

    def timed(label, fn, flops, peak):
        fn()                                          # first call: one-time allocations / FP32 copies
        ctx.stats_reset(True)
        t0 = time.perf_counter()
        r = fn()
        wall = time.perf_counter() - t0
        st = ctx.stats()
        ctx.stats_reset(False)
        dev = sum(x["ms"] for x in st.values())
        out["stages"][label] = {"device_ms": round(dev, 2), "wall_s_incl_host_copies": round(wall, 3),
                                "flops": flops, "TFLOP/s": round(flops / dev / 1e9, 1),
                                "frac_of_peak": round(flops / dev / 1e9 / peak, 3), "peak": peak,
                                "kernels_ms": {k: round(x["ms"], 2) for k, x in st.items()}}
        print(f"[bench] {label}: device {dev:.1f} ms, {flops / dev / 1e9:.1f} TFLOP/s", file=sys.stderr, flush=True)
        return r

    Co, Cv = s.Co, s.Cv
    timed("transform_iajb_fp64", lambda: Transformation.tei_transform(g, Co, Cv, Co, Cv), F_X, 78.6)
    timed("transform_iajb_fp32", lambda: Transformation.tei_transform(g, Co, Cv, Co, Cv, prec="fp32"), F_X, 157.3)
    E64 = timed("rmp2_fp64", lambda: MollerPlesset.do_rmp2(w), F_X, 78.6)
    E32 = timed("rmp2_fp32_mixed", lambda: MollerPlesset.do_rmp2(w, prec="fp32_mixed"), F_X, 157.3)
    res = timed("rmp2_grad_fp64_with_dEdg", lambda: MollerPlesset.rmp2_grad(w), F_G, 78.6)
    res32 = timed("rmp2_grad_fp32_mixed_with_dEdg", lambda: MollerPlesset.rmp2_grad(w, prec="fp32_mixed"), F_G, 157.3)
    # FP32-mixed adjoint against FP64: max-abs differences relative to the FP64 maxima (dE/dg in chunks
    # of its 4.3e9 entries, so no second n^4 temporary is formed on the host)
    a, b = res[2].ravel(order="K"), res32[2].ravel(order="K")
    step = 1 << 26
    dg_diff = max(float(np.max(np.abs(a[i:i + step] - b[i:i + step]))) for i in range(0, a.size, step))
    dg_max = max(float(np.max(np.abs(a[i:i + step]))) for i in range(0, a.size, step))
    out["grad_fp32_mixed_vs_fp64"] = {
        "E_rel": abs(res32[0] - res[0]) / abs(res[0]),
        "dC_rel": float(np.max(np.abs(res32[1] - res[1])) / np.max(np.abs(res[1]))),
        "dg_rel": dg_diff / dg_max,
        "deps_rel": float(np.max(np.abs(res32[3] - res[3])) / np.max(np.abs(res[3])))}
    del a, b, res32
    out["E_mp2_fp64"] = E64
    out["E_mp2_fp32_mixed_abs_diff"] = abs(E32 - E64)
    out["grad_E"] = res[0]
    del g, G2, w
    ctx.release()
    out["factor_route"] = {name: mp2_factor_adjoint(name, fd_entries=4 if name.startswith("cfg5") else 0)
                           for name in ("cfg3_mp2grad_n256", "cfg5_rccsd_n500")}
    return out


def mp2_factor_adjoint(name: str, fd_entries: int = 0) -> dict:
    """The MP2 adjoint through the factor route on a seeded workload (no n^4 tensor, no n^2 o v
    half-transform), FP64 and FP32-mixed; with fd_entries, dE/dC at that many seeded entries
    against the five-point stencil of the device energy at h = 0.1 (exact for E, a quartic
    along one entry of C, up to rounding)."""
    import numpy as np
    from jues import synthetic, MollerPlesset, _lib
    from jues.Wavefunction import Wfn, FactoredERI
    s = synthetic.workload(name)
    n, o, v, naux = s.n, s.o, s.v, s.naux
    ctx = _lib.context(0)
    wf = lambda C: Wfn.from_orbitals(C, s.eps, o, FactoredERI(s.B))  # noqa: E731
    ov = o * v
    # mo factors, K = L L^T, dL = 2 G L, B_P Cv and B_P^T Co, and the two dC products
    F = (2 * n * n * o * naux + 2 * n * o * v * naux + 4 * ov * ov * naux + 2 * n * n * v * naux
         + 2 * n * n * o * naux + 4 * n * o * v * naux)
    rec = {"workload": f"RMP2 adjoint, factor route, n={n} o={o} v={v} naux={naux}", "flops": F}
    for prec, peak in (("fp64", 78.6), ("fp32_mixed", 157.3)):
        MollerPlesset.rmp2_grad(wf(s.C), want_dg=False, prec=prec)
        ctx.stats_reset(True)
        t0 = time.perf_counter()
        E, dC, _, de = MollerPlesset.rmp2_grad(wf(s.C), want_dg=False, prec=prec)
        wall = time.perf_counter() - t0
        st = ctx.stats()
        ctx.stats_reset(False)
        dev = sum(x["ms"] for x in st.values())
        rec[prec] = {"E": E, "device_ms": round(dev, 2), "wall_s_incl_host_copies": round(wall, 3),
                     "TFLOP/s": round(F / dev / 1e9, 1), "frac_of_peak": round(F / dev / 1e9 / peak, 3),
                     "kernels_ms": {k: round(x["ms"], 2) for k, x in st.items()}}
        if prec == "fp64":
            dC64, de64 = dC, de
        else:
            rec[prec]["rel_err_dC_vs_fp64"] = float(np.abs(dC - dC64).max() / np.abs(dC64).max())
            rec[prec]["rel_err_E_vs_fp64"] = abs(E - rec["fp64"]["E"]) / abs(rec["fp64"]["E"])
        print(f"[bench] {name} adjoint {prec}: device {dev:.1f} ms, {F / dev / 1e9:.1f} TFLOP/s", file=sys.stderr,
              flush=True)
    if fd_entries:
        rng = np.random.default_rng(5)
        h = 0.1
        checks = []
        for mu, p in rng.integers(0, n, (fd_entries, 2)):
            def E_at(step):
                C = s.C.copy()
                C[mu, p] += step
                return MollerPlesset.do_rmp2(wf(C))
            fd5 = (-E_at(2 * h) + 8 * E_at(h) - 8 * E_at(-h) + E_at(-2 * h)) / (12 * h)
            checks.append({"entry": [int(mu), int(p)], "analytic": float(dC64[mu, p]), "fd5": fd5,
                           "abs_diff": abs(fd5 - float(dC64[mu, p]))})
        rec["fd5_check_h0.1"] = checks
        print(f"[bench] {name} fd5 max |diff| {max(c['abs_diff'] for c in checks):.2e}", file=sys.stderr, flush=True)
    ctx.release()
    return rec


def df_bench(coupling: float) -> dict:
    """Density fitting on the o=60, v=340 factors (n=400, naux=1200): b = pqP.Jpqh on the
    device, DF-RMP2 in FP64 and FP32-mixed, and DF-RCCD sweeps (recomputed bvv.bvv ladder).
    pqP = b0.J^(1/2) for a seeded metric with eigenvalues in [0.3, 3]."""
    import numpy as np
    from jues import synthetic, DF, MollerPlesset, _lib
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn, DFIntegrals
    s = synthetic.workload("cfg4_rccd_n400")
    n, o, v, naux = s.n, s.o, s.v, s.naux
    rng = np.random.default_rng(11)
    Q, _ = np.linalg.qr(rng.standard_normal((naux, naux)))
    wv = np.linspace(0.3, 3.0, naux)
    Jh = (Q * np.sqrt(wv)) @ Q.T
    Jpqh = np.asfortranarray((Q / np.sqrt(wv)) @ Q.T)
    pqP = np.asfortranarray((s.B.reshape(n * n, naux, order="F") @ Jh).reshape(n, n, naux, order="F"))
    ctx = _lib.context(0)
    out = {"workload": f"density fitting, n={n} o={o} v={v} naux={naux}", "stages": {}}

    def timed(label, fn, flops, peak, reps=2):
        for _ in range(reps - 1):
            fn()
        ctx.stats_reset(True)
        r = fn()
        st = ctx.stats()
        ctx.stats_reset(False)
        dev = sum(x["ms"] for x in st.values())
        out["stages"][label] = {"device_ms": round(dev, 2), "flops": flops, "TFLOP/s": round(flops / dev / 1e9, 1),
                                "frac_of_peak": round(flops / dev / 1e9 / peak, 3), "peak": peak,
                                "kernels_ms": {k: round(x["ms"], 2) for k, x in st.items()}}
        print(f"[bench] {label}: device {dev:.1f} ms, {flops / dev / 1e9:.1f} TFLOP/s", file=sys.stderr, flush=True)
        return r

    b = timed("make_bmunu_fp64", lambda: DF.make_bmunu(pqP, Jpqh), 2 * n * n * naux * naux, 78.6)
    out["bmunu_max_abs_err"] = float(np.abs(b - s.B).max())
    del b
    w = Wfn.from_orbitals(s.C, s.eps, o, DFIntegrals(pqP, Jpqh))
    F_mp2 = 2 * n * n * o * naux + 2 * n * o * v * naux + 2 * (o * v) ** 2 * naux
    E64 = timed("df_rmp2_fp64", lambda: MollerPlesset.do_df_rmp2(w), F_mp2, 78.6)
    E32 = timed("df_rmp2_fp32_mixed", lambda: MollerPlesset.do_df_rmp2(w, prec="fp32_mixed"), F_mp2, 157.3)
    out["E_df_mp2_fp64"] = E64
    out["E_df_mp2_fp32_mixed_rel_diff"] = abs(E32 - E64) / abs(E64)
    for prec in ("fp64", "fp32_mixed", "fp16x3_mixed"):
        hist = []
        t0 = time.perf_counter()
        CC.DFRCCD.do_df_rccd(w, maxit=3, prec=prec, coupling=coupling, callback=hist.append)
        wall = time.perf_counter() - t0
        out["stages"][f"df_rccd_sweep_{prec}"] = {"ms_per_sweep": [round(h["ms"], 1) for h in hist],
                                                  "wall_s_incl_setup": round(wall, 1)}
        print(f"[bench] df_rccd {prec}: {[round(h['ms']) for h in hist]} ms/sweep", file=sys.stderr, flush=True)
    return out


def triples_bench(o: int, forms) -> dict:
    """(T) on the factors of the n=400, v=340 system with the first ``o`` occupied
    orbitals and seeded pair-symmetric amplitudes (|t1| ~ |t2| ~ 1e-2); the integral
    blocks are built on the device.  Device times from HIP events."""
    import numpy as np
    from jues import synthetic, _lib
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn, FactoredERI
    s = synthetic.workload("cfg4_rccd_n400")
    v = s.v
    import dataclasses
    w = Wfn.from_orbitals(s.C, s.eps, s.o, FactoredERI(s.B))
    w = dataclasses.replace(w, nalpha=o, nbeta=o, Cao=np.asfortranarray(s.C[:, :o]), Cbo=np.asfortranarray(s.C[:, :o]),
                            epsa=np.concatenate([s.eps[:o], s.eps[s.o:]]))
    rng = np.random.default_rng(17)
    T1 = rng.standard_normal((o, v)) * 1e-2
    T2 = rng.standard_normal((o, o, v, v)) * 1e-2
    T2 = (T2 + T2.transpose(1, 0, 3, 2)) / 2
    ctx = _lib.context(0)
    ntrip = o * (o + 1) * (o + 2) // 6
    flops_w = ntrip * (12.0 * v**4 + 12.0 * v**3 * o)
    out = {"workload": f"(T), o={o} v={v} (n=400 factors), {ntrip} occupied triples i>=j>=k",
           "flops_W": flops_w, "forms": {}}
    for prec in forms:
        ctx.stats_reset(True)
        t0 = time.perf_counter()
        Et = CC.PerturbativeTriples.rccsd_pT(w, T1, T2, prec=prec, doprint=True)
        wall = time.perf_counter() - t0
        st = ctx.stats()
        ctx.stats_reset(False)
        gw = sum(x["ms"] for k, x in st.items() if k.startswith("pt_w"))
        peak = 78.6 if prec == "fp64" else 157.3
        out["forms"][prec] = {"Et": Et, "wall_s": round(wall, 2), "W_gemm_ms": round(gw, 1),
                              "W_TFLOP/s": round(flops_w / gw / 1e9, 1),
                              "W_frac_of_peak": round(flops_w / gw / 1e9 / peak, 3),
                              "ms_per_triple": round(wall * 1e3 / ntrip, 3),
                              "kernels_ms": {k: round(x["ms"], 1) for k, x in st.items()}}
        print(f"[bench] (T) {prec}: Et={Et:.12e} wall {wall:.1f} s, W GEMMs {flops_w / gw / 1e9:.1f} TFLOP/s",
              file=sys.stderr, flush=True)
    if "fp64" in out["forms"] and "fp32_mixed" in out["forms"]:
        e64, e32 = out["forms"]["fp64"]["Et"], out["forms"]["fp32_mixed"]["Et"]
        out["fp32_rel_diff"] = abs(e32 - e64) / abs(e64)
    return out


def n4_bench() -> dict:
    """On the n=400 factors: UMP2 (alpha = the o=60 orbitals, beta = a rotated set with 59
    occupied), get_fock, and CCSD iterations (RCCSD equations, o=60, v=340, integrals
    scaled by 0.005) with a canonical and a non-canonical Fock matrix (diag(eps) plus a
    seeded symmetric 1e-3 perturbation); device times from HIP events."""
    import numpy as np
    from scipy.linalg import expm
    from jues import synthetic, MollerPlesset, _lib
    from jues import CoupledCluster as CC
    from jues.IntegralTransformation import get_fock
    from jues.Wavefunction import Wfn, FactoredERI
    s = synthetic.workload("cfg4_rccd_n400")
    n, o = s.n, s.o
    rng = np.random.default_rng(23)
    K = rng.standard_normal((n, n)) * 0.02
    Cb = s.C @ expm(K - K.T)
    h = rng.standard_normal((n, n))
    h = (h + h.T) / 2
    eri = FactoredERI(s.B)
    ctx = _lib.context(0)
    out = {"workload": f"n={n} factors (naux={s.naux}), o={o} v={s.v}", "stages": {}}

    def timed(label, fn, reps=2):
        for _ in range(reps - 1):
            fn()
        ctx.stats_reset(True)
        r = fn()
        st = ctx.stats()
        ctx.stats_reset(False)
        dev = sum(x["ms"] for x in st.values())
        out["stages"][label] = {"device_ms": round(dev, 2), "kernels_ms": {k: round(x["ms"], 2) for k, x in st.items()}}
        print(f"[bench] {label}: device {dev:.1f} ms", file=sys.stderr, flush=True)
        return r

    wu = Wfn.from_uhf(s.C, Cb, s.eps, s.eps, o, o - 1, eri, hao=h)
    out["E_ump2_fp64"] = timed("ump2_fp64", lambda: MollerPlesset.do_ump2(wu))
    out["E_ump2_fp32_rel_diff"] = abs(timed("ump2_fp32_mixed", lambda: MollerPlesset.do_ump2(wu, prec="fp32_mixed"))
                                      - out["E_ump2_fp64"]) / abs(out["E_ump2_fp64"])
    timed("get_fock", lambda: get_fock(wu))
    # integrals scaled by 0.005 (the factors by its square root), where the iteration converges
    wr = Wfn.from_orbitals(s.C, s.eps, o, FactoredERI(np.asfortranarray(s.B * np.sqrt(0.005))))
    P = rng.standard_normal((n, n)) * 1e-3
    fock = np.diag(s.eps) + (P + P.T) / 2
    for label, fk in (("rccsd_canonical", None), ("rccsd_fock_noncanonical", fock)):
        for prec in ("fp64", "fp32_mixed"):
            hist = []
            kw = dict(prec=prec, maxit=3, refine_maxit=0, diis=False, check_convergence=False, ccsd_eqs="RCCSD",
                      callback=hist.append)
            CC._run(1, wr, kw, "mp2", fock=fk)
            ms = [round(x["ms"], 1) for x in hist if x["stage"] == (0 if prec == "fp64" else 1)]
            out["stages"][f"{label}_{prec}"] = {"ms_per_iteration": ms}
            print(f"[bench] {label} {prec}: {ms} ms/iter", file=sys.stderr, flush=True)
    return out


def cfg2_bench() -> dict:
    """RCCSD (mRCCSD equations) at n=120, o=20, v=100, naux=360 with the factors scaled by
    sqrt(0.02), the integrals the committed fixture tests/golden/cfg2_coupling0.02.npz was
    converged for by the CPU oracle: each form from t = 0 to the FP64 fixed point, with the
    device time per iteration of each stage and |E - E_oracle| after each stage."""
    import numpy as np
    from jues import synthetic
    from jues import CoupledCluster as CC
    from jues.Wavefunction import Wfn, FactoredERI
    gold = np.load(ROOT / "tests" / "golden" / "cfg2_coupling0.02.npz")
    lam = float(gold["coupling"])
    s = synthetic.workload("cfg2_rccsd_n120")
    w = Wfn.from_orbitals(s.C, s.eps, s.o, FactoredERI(np.asfortranarray(s.B * np.sqrt(lam))))
    E_ref = float(gold["E_mrccsd"])
    out = {"workload": f"RCCSD n={s.n} o={s.o} v={s.v} naux={s.naux}, factors scaled by sqrt({lam})",
           "E_oracle": E_ref, "forms": {}}
    names = {0: "fp64", 1: "fp32", 2: "fp16x3"}
    for prec in ("fp64", "fp32_mixed", "fp16x3"):
        hist = []
        CC.mRCCSD.do_rccsd(w, prec=prec, guess="zero", diis_max=8, cc_e_conv=1e-12, cc_max_rms=1e-11,
                           fp32_r_conv=1e-6, maxit=300, callback=hist.append)
        stages = {}
        for h in hist:
            st = stages.setdefault(names[h["stage"]], {"iterations": 0, "ms": []})
            st["iterations"] += 1
            st["ms"].append(h["ms"])
            st["E_end"] = h["energy"]
        rec = {}
        for name, st in stages.items():
            rec[name] = {"iterations": st["iterations"], "median_ms_per_iter": round(float(np.median(st["ms"])), 2),
                         "abs_dE_vs_oracle_at_stage_end": abs(st["E_end"] - E_ref)}
        out["forms"][prec] = rec
        print(f"[bench] cfg2 {prec}: " + ", ".join(f"{k} {v['iterations']} it {v['median_ms_per_iter']} ms "
                                                     f"|dE| {v['abs_dE_vs_oracle_at_stage_end']:.1e}"
                                                     for k, v in rec.items()), file=sys.stderr, flush=True)
    return out


def rccsd5_bench(coupling: float, steps: int, warmup: int, diis_max: int) -> dict:
    """RCCSD at o=80, v=420 (n=500, naux=1500) on seeded factors: the pp-ladder rebuilt from the
    factors every iteration and (ia|bc) never stored.  One resident state, timed per form as the
    RCCD benchmark does (FP32-mixed, split-FP16, FP64); device memory held after each stage."""
    from jues import synthetic, _lib
    from jues.CoupledCluster import Solver
    from jues.Wavefunction import Wfn
    s = synthetic.workload("cfg5_rccsd_n500")
    o, v, naux = s.o, s.v, s.naux
    w = Wfn.from_synthetic(s, dense=False)
    t0 = time.perf_counter()
    solver = Solver(w, method="rccsd", prec="fp16x3", guess="zero", diis=True, diis_max=min(diis_max, 3),
                    coupling=coupling, ladder="recompute")
    ctx = solver.ctx
    out = {"workload": f"RCCSD o={o} v={v} n={s.n} naux={naux}, B-factor integrals, pp-ladder recomputed "
                       "per iteration, (ia|bc) factored", "coupling": coupling, "setup_s": round(time.perf_counter() - t0, 1),
           "device_gib_after_setup": round(ctx.lib.jues_device_bytes(ctx.h) / 2**30, 1), "forms": {}}
    print(f"[bench] rccsd5 setup {out['setup_s']} s, {out['device_gib_after_setup']} GiB", file=sys.stderr, flush=True)
    for stage, prec in (("fp32", "fp32_mixed"), ("fp16x3", "fp16x3"), ("fp64", "fp64")):
        solver.reset_diis()
        if warmup:
            solver.iterate(warmup, prec=prec)
        ctx.stats_reset(True)
        t0 = time.perf_counter()
        info, _ = solver.iterate(steps, prec=prec)
        dt = time.perf_counter() - t0
        st = ctx.stats()
        ctx.stats_reset(False)
        ex = sum(x["flops"] for x in st.values()) / steps
        out["forms"][stage] = {"s_per_iter": round(dt / steps, 3), "E": info["energy"], "rnorm": info["rnorm"],
                               "executed_tflops": round(ex / (dt / steps) / 1e12, 1),
                               "device_gib": round(ctx.lib.jues_device_bytes(ctx.h) / 2**30, 1),
                               "kernel_ms_per_iter": {k: round(x["ms"] / steps, 1) for k, x in st.items()}}
        print(f"[bench] rccsd5 {stage}: {dt / steps:.3f} s/iter, E={info['energy']:.10f}, "
              f"{out['forms'][stage]['device_gib']} GiB", file=sys.stderr, flush=True)
    f = out["forms"]
    out["speedup_vs_fp64"] = {k: round(f["fp64"]["s_per_iter"] / f[k]["s_per_iter"], 2) for k in ("fp32", "fp16x3")}
    return out


def converge(w, coupling: float, fast_r_conv: float = 1e-6, keep_diis: bool = False,
             forms=("fp64", "fp32_mixed", "fp16x3"), refines=("restart", "defect"), method: str = "rccd",
             diis_max: int = 8, ladder: str = "auto", fast_r_conv_defect: float | None = None) -> dict:
    """Each form to convergence: plain FP64, and each fast form (FP32-mixed, split-FP16) followed
    by each FP64 hand-over: FP64 iterations (DIIS history reset unless keep_diis), or defect
    correction (FP64 iterations anchoring inner solves in the fast form)."""
    from jues.CoupledCluster import Solver
    fr_dc = fast_r_conv if fast_r_conv_defect is None else fast_r_conv_defect
    res = {"method": method, "coupling": coupling, "fast_r_conv": fast_r_conv, "fast_r_conv_defect": fr_dc,
           "keep_diis": keep_diis, "diis_max": diis_max}
    kw = dict(guess="zero", diis=True, diis_max=diis_max, coupling=coupling, cc_e_conv=1e-10, cc_max_rms=1e-9,
              fp32_r_conv=fast_r_conv, doprint=True, ladder=ladder)
    if "fp64" in forms:
        sd = Solver(w, method=method, prec="fp64", **kw)
        t0 = time.perf_counter()
        d64, cd = sd.iterate(200, prec="fp64", stop_on_conv=True)
        td = time.perf_counter() - t0
        Ed = sd.energy()
        print(f"[bench] fp64: {d64['iter']} it {td:.1f} s E={Ed:.12f} conv={cd}", file=sys.stderr, flush=True)
        res["fp64"] = dict(iters=d64["iter"], time_to_solution_s=round(td, 2), E=Ed, converged=cd,
                           final_rnorm=d64["rnorm"], final_dE=d64["delta_e"])
        del sd
    for form in ("fp32_mixed", "fp16x3"):
        if form not in forms:
            continue
        for mode in refines:
            sm = Solver(w, method=method, prec=form, **dict(kw, fp32_r_conv=fr_dc if mode == "defect" else fast_r_conv))
            t0 = time.perf_counter()
            i1, _ = sm.iterate(200, prec=form, stop_on_conv=True)
            t1 = time.perf_counter() - t0
            E1 = sm.energy()
            t0 = time.perf_counter()
            if mode == "defect":
                # stage 0 records are the FP64 iterations; the rest are inner correction steps
                i2, c2 = sm.refine(form, max_anchors=20, inner_maxit=100)
            else:
                if not keep_diis:
                    sm.reset_diis()
                i2, c2 = sm.iterate(100, prec="fp64", stop_on_conv=True)
            t2 = time.perf_counter() - t0
            E2 = sm.energy()
            n64 = i2["n_fp64"] if mode == "defect" else i2["iter"] - i1["iter"]
            print(f"[bench] {form}+{mode}: fast {i1['iter']} it {t1:.1f} s E={E1:.12f}; refine "
                  f"{i2['iter'] - i1['iter']} it ({n64} FP64) {t2:.1f} s E={E2:.12f} conv={c2}",
                  file=sys.stderr, flush=True)
            rec = dict(fast_iters=i1["iter"], fast_s=round(t1, 2), E_after_fast=E1,
                       refine_iters=i2["iter"] - i1["iter"], refine_fp64_iters=n64, refine_s=round(t2, 2), E=E2,
                       converged=c2, time_to_solution_s=round(t1 + t2, 2), final_rnorm=i2["rnorm"],
                       final_dE=i2["delta_e"])
            if "fp64" in res:
                rec["abs_dE_vs_fp64"] = abs(E2 - res["fp64"]["E"])
                rec["abs_dE_fast_stage_vs_fp64"] = abs(E1 - res["fp64"]["E"])
                rec["speedup_time_to_solution_vs_fp64"] = round(res["fp64"]["time_to_solution_s"] / (t1 + t2), 2)
            res[f"{form}+{mode}"] = rec
            del sm
    return res


if __name__ == "__main__":
    main()
