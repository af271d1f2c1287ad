"""Benchmark: RK4 steps per second of one periodic water surface of N = 2048 nodes in FP64.

Workload: nodes x_j = 2 pi j / N, y = sum_{k=1..8} a_k cos(k x + theta_k) with a_k ~ U(0, 0.02/k),
theta_k ~ U(0, 2 pi) (numpy default_rng(20261202)), phi = 0, gravity 1, infinite depth,
dt = 0.002.  A step is one classical RK4 step (four right-hand sides: spectral derivatives,
dipole-matrix assembly, dense solve, velocity + Bernoulli + stage update), state resident in HBM.

Prints one JSON line.  `roofline` is measured live on the dominant kernel (the FP64 MFMA
trailing update of the LU) with hipEvents around each of its launches on the solver's stream;
`cpu_baseline` times the CPU FP64 oracle (OpenMP, all host threads) on a bounded sample of the
same workload and also serves as the accuracy reference for the same number of steps.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "superfluid-dynamics_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "RK4 steps/s (N=2048 FP64; batched mixed-prec) at ≤1e-9 err vs FP64 oracle"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
N = 2048
DT = 0.002
SEED = 20261202


def surface(n, seed, amp=0.02, modes=8):
    rng = np.random.default_rng(seed)
    x = 2 * np.pi * np.arange(n) / n
    a = rng.uniform(0, amp / np.arange(1, modes + 1))
    th = rng.uniform(0, 2 * np.pi, modes)
    return x + 1j * sum(a[k] * np.cos((k + 1) * x + th[k]) for k in range(modes))


def cpu_baseline(Z0, P0, budget_s=12.0, max_steps=12):
    from oracle import oracle as O
    O.build()
    Z, P = Z0.copy(), P0.copy()
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps:
        Z, P = O.rk4(O.Props(), Z, P, DT, 1)
        steps += 1
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    return steps, el, Z, P, O.num_threads()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--solver", choices=("lu", "gmres"), default="lu")
    ap.add_argument("--max-iter", type=int, default=16, help="GMRES iteration cap per solve")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the stepper runs one "
              f"surface ensemble per GPU and this build has no multi-GPU path)", file=sys.stderr)
        sys.exit(2)

    import cusuperhelium as c

    Z0 = surface(N, SEED)
    P0 = np.zeros(N)
    solver = c.GMRES if args.solver == "gmres" else c.LU
    e = c.Ensemble([N], c.Props(), precision=c.FP64, solver=solver, tol=1e-12, max_iter=args.max_iter)
    e.set_state(Z0, P0)
    e.step(DT, args.warmup)
    e.sync()
    t0 = time.perf_counter()
    e.step(DT, args.steps)
    e.sync()
    el = time.perf_counter() - t0
    value = args.steps / el
    ms_per_step = 1e3 * el / args.steps

    # live kernel timing: a few steps with hipEvents around every launch
    e.kernel_timing(True)
    kt_steps = 3
    e.step(DT, kt_steps)
    e.sync()
    reps = {t: e.kernel_report(t) for t in ("deriv", "assemble", "lu_panel", "lu_update", "lu_backsolve",
                                            "gm_init", "gm_arnoldi", "gm_extend", "gm_finish",
                                            "mu_deriv", "velocity")}
    st = e.stats()
    e.kernel_timing(False)
    if args.solver == "gmres":
        # dominant solver kernel: w = M v over the stored FP64 matrix (8 N^2 bytes per launch)
        dom, bound, peak, unit, key, scale = "gm_arnoldi", "hbm", PEAK_HBM_GBS, "GB/s", "bytes", 1e9
        kname = "gm_arnoldi_kernel (FP64 GMRES matrix-vector product + Arnoldi dots)"
    else:
        dom, bound, peak, unit, key, scale = "lu_update", "mfma", PEAK_FP64_TFLOPS, "TFLOP/s", "flops", 1e12
        kname = "lu_update_kernel (FP64 MFMA trailing update)"
    r = reps[dom]
    avg_ms = r["ms"] / max(r["launches"], 1)
    work_per_launch = r[key] / max(r["launches"], 1)
    achieved = work_per_launch / (avg_ms * 1e-3) / scale if avg_ms > 0 else 0.0
    stages = {t: round(v["ms"] / kt_steps, 4) for t, v in reps.items() if v["launches"]}
    e.close()

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "RK4 steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic: seeded 8-mode surface (numpy default_rng({SEED})), phi = 0",
        "config": {"workload": "single periodic water surface, N=2048 nodes, FP64, gravity 1, infinite depth, "
                               f"dt=0.002, {'GMRES (tol 1e-12)' if args.solver == 'gmres' else 'blocked LU'} solve",
                   "nodes": N, "dt": DT, "solver": args.solver, "precision": "fp64"},
        "roofline": {"bound": bound, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
                     "frac": round(achieved / peak, 4), "traffic": None, "kernel": kname,
                     "avg_launch_ms": round(avg_ms, 5), f"{key}_per_launch": work_per_launch,
                     "launches_timed": r["launches"]},
        "stage_ms_per_step": stages,
        "solver_stats": {"gmres_iterations_last": st["last_iterations"], "gmres_unconverged": st["unconverged"],
                         "gepp_fallbacks": st["fallback_solves"]},
    }
    if not args.no_cpu:
        steps, cel, Zc, Pc, threads = cpu_baseline(Z0, P0)
        out["cpu_baseline"] = {"value": round(steps / cel, 4), "unit": "RK4 steps/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{steps} RK4 steps of the same N=2048 workload on the CPU FP64 oracle "
                                         f"(OpenMP, {threads} threads), {cel:.1f} s"}
        g = c.Ensemble([N], c.Props(), precision=c.FP64, solver=solver, tol=1e-12, max_iter=args.max_iter)
        g.set_state(Z0, P0)
        g.step(DT, steps)
        Zg, Pg = g.get_state()
        g.close()
        out["accuracy"] = {"steps_compared": steps, "max_abs_dz_vs_oracle": float(np.abs(Zg - Zc).max()),
                           "max_abs_dphi_vs_oracle": float(np.abs(Pg - Pc).max())}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
