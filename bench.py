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
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one MI355X; multi-GPU runs are not part of it "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from jues import synthetic, _lib
    from jues.CoupledCluster import Solver
    from jues.Wavefunction import Wfn

    s = synthetic.workload(args.workload)
    o, v = s.o, s.v
    w = Wfn.from_synthetic(s, dense=False)
    t0 = time.perf_counter()
    solver = Solver(w, method="rccd", prec="fp32_mixed", guess="zero", diis=True, diis_max=8,
                    coupling=args.coupling)
    t_setup = time.perf_counter() - t0
    ctx = solver.ctx
    print(f"[bench] setup {t_setup:.1f} s, device bytes {ctx.lib.jues_device_bytes(ctx.h) / 2**30:.1f} GiB",
          file=sys.stderr, flush=True)

    res = {}
    for stage, prec in (("fp32", "fp32_mixed"), ("fp16x3", "fp16x3"), ("fp64", "fp64")):
        solver.reset_diis()
        if args.warmup:
            solver.iterate(args.warmup, prec=prec)
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
                   "o": o, "v": v, "n": s.n, "naux": s.naux, "diis_max": 8, "coupling": args.coupling},
        "fp64_s_per_iter": round(d["s_per_iter"], 4),
        "speedup_mixed_vs_fp64": round(d["s_per_iter"] / m["s_per_iter"], 3),
        "tflops_canonical": {"fp32_mixed": round(F / m["s_per_iter"] / 1e12, 2),
                             "fp64": round(F / d["s_per_iter"] / 1e12, 2), "flops_per_iter": F},
        "tflops_executed": {"fp32_mixed": round(executed(m["stats"]) / args.steps / m["s_per_iter"] / 1e12, 2),
                            "fp64": round(executed(d["stats"]) / args.steps / d["s_per_iter"] / 1e12, 2),
                            "flops_per_iter": executed(m["stats"]) / args.steps},
        "fp16x3": {"s_per_iter": round(h["s_per_iter"], 4),
                   "speedup_vs_fp64": round(d["s_per_iter"] / h["s_per_iter"], 3),
                   "tflops_canonical": round(F / h["s_per_iter"] / 1e12, 2),
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
                               "host_cpus": os.cpu_count()}
    if args.converge:
        out["convergence"] = converge(w, args.coupling)
    print(json.dumps(out), flush=True)


def converge(w, coupling: float) -> dict:
    """Both forms to convergence: FP32-mixed (FP32 stage, DIIS reset, FP64 refinement) and plain FP64."""
    from jues.CoupledCluster import Solver
    res = {}
    kw = dict(guess="zero", diis=True, diis_max=8, coupling=coupling, cc_e_conv=1e-10, cc_max_rms=1e-9,
              fp32_r_conv=1e-5, doprint=True)
    t0 = time.perf_counter()
    sm = Solver(w, method="rccd", prec="fp32_mixed", **kw)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    i32, _ = sm.iterate(200, prec="fp32_mixed", stop_on_conv=True)
    t32 = time.perf_counter() - t0
    E32 = sm.result()[0]
    sm.reset_diis()
    t0 = time.perf_counter()
    i64, c64 = sm.iterate(100, prec="fp64", stop_on_conv=True)
    tref = time.perf_counter() - t0
    Em = sm.result()[0]
    print(f"[bench] mixed: fp32 {i32['iter']} it {t32:.1f} s E={E32:.12f}; refine {i64['iter'] - i32['iter']} it "
          f"{tref:.1f} s E={Em:.12f} conv={c64}", file=sys.stderr, flush=True)
    res["mixed"] = dict(fp32_iters=i32["iter"], fp32_s=round(t32, 2), E_after_fp32=E32,
                        refine_iters=i64["iter"] - i32["iter"], refine_s=round(tref, 2), E=Em, converged=c64,
                        time_to_solution_s=round(t32 + tref, 2), final_rnorm=i64["rnorm"], final_dE=i64["delta_e"])
    del sm
    t0 = time.perf_counter()
    sh = Solver(w, method="rccd", prec="fp16x3", **kw)
    t0 = time.perf_counter()
    i16, _ = sh.iterate(200, prec="fp16x3", stop_on_conv=True)
    t16 = time.perf_counter() - t0
    E16 = sh.result()[0]
    sh.reset_diis()
    t0 = time.perf_counter()
    j64, cj = sh.iterate(100, prec="fp64", stop_on_conv=True)
    trh = time.perf_counter() - t0
    Eh = sh.result()[0]
    print(f"[bench] fp16x3: split {i16['iter']} it {t16:.1f} s E={E16:.12f}; refine {j64['iter'] - i16['iter']} it "
          f"{trh:.1f} s E={Eh:.12f} conv={cj}", file=sys.stderr, flush=True)
    res["fp16x3"] = dict(split_iters=i16["iter"], split_s=round(t16, 2), E_after_split=E16,
                         refine_iters=j64["iter"] - i16["iter"], refine_s=round(trh, 2), E=Eh, converged=cj,
                         time_to_solution_s=round(t16 + trh, 2), final_rnorm=j64["rnorm"], final_dE=j64["delta_e"])
    del sh
    t0 = time.perf_counter()
    sd = Solver(w, method="rccd", prec="fp64", **kw)
    t0 = time.perf_counter()
    d64, cd = sd.iterate(200, prec="fp64", stop_on_conv=True)
    td = time.perf_counter() - t0
    Ed = sd.result()[0]
    print(f"[bench] fp64: {d64['iter']} it {td:.1f} s E={Ed:.12f} conv={cd}", file=sys.stderr, flush=True)
    res["fp64"] = dict(iters=d64["iter"], time_to_solution_s=round(td, 2), E=Ed, converged=cd,
                       final_rnorm=d64["rnorm"], final_dE=d64["delta_e"])
    res["abs_dE_mixed_vs_fp64"] = abs(Em - Ed)
    res["abs_dE_fp32_stage_vs_fp64"] = abs(E32 - Ed)
    res["abs_dE_fp16x3_vs_fp64"] = abs(Eh - Ed)
    res["abs_dE_split_stage_vs_fp64"] = abs(E16 - Ed)
    res["coupling"] = coupling
    res["setup_s"] = round(t_setup, 2)
    return res


if __name__ == "__main__":
    main()
