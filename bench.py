"""Benchmark: RK4 steps per second of one periodic water surface of N = 2048 nodes in FP64, plus
the batched mixed-precision ensemble leg and the other workloads.

Headline workload (`value`): nodes x_j = 2 pi j / N, y = sum_{k=1..8} a_k cos(k x + theta_k) with
a_k ~ U(0, 0.02/k), theta_k ~ U(0, 2 pi) (numpy default_rng(20261202)), phi = 0, gravity 1, infinite
depth, dt = 0.002.  A step is one classical RK4 step (four right-hand sides: spectral derivatives,
dipole-matrix assembly, solve, velocity + Bernoulli + stage update), state resident in HBM, steps
replayed from a captured hipGraph.  Default solver: FP64 GMRES (relative residual 1e-12, warm
started); --solver lu selects the blocked FP64 LU.  The warm-up runs the requested W steps and,
for GMRES, continues (untimed) until the solver is in its steady state: 5 steps fill the warm-start
history, 1 sets the trimmed launch sequence, and the chunks of steps checked for unconverged
solves grow 2, 4, 8, 16 steps to the steady check interval of 32 steps (36 steps in all), so the
timed K steps run the steady-state graph with one host check per 32 steps.

Batched leg (`batched`): 512 independent surfaces with N drawn uniformly from {128, 256, 512, 1024}
(default_rng(20261202)), surface i drawn with seed 20261202 + i as above, phi = 0, dt = 0.002; one
persistent launch per RK stage, mixed precision (FP32 matrix, FP64-refined mu) and FP64.  It is
compared with the plain FP64 stepper run over the same 512 surfaces one surface at a time:
ensemble steps/s = 1 / sum_i t_step(N_i).

Other legs: `c2_lu` (the headline surface with the drop-in LU solver), `c1` (N = 256,
y = 0.1 cos x, dt = 0.01, 100 steps, LU and GMRES), `film` (N = 4096 helium film with the image
term) and `adjoint` (gradient through the adjoint).

Accuracy: every workload is compared with golden states of the CPU FP64 oracle stored in
tests/golden/ (the same seeded inputs): max |z - z_oracle| and max |mu - mu_oracle| (mu of the
right-hand side at the final state), per instance for the ensemble.

Prints one JSON line.  `roofline` is measured live on the dominant kernel of the headline solver
with a pair of hipEvents per launch on the solver's stream, recorded by the launch itself
(hipExtLaunchKernelGGL: the kernel's own start and end, without the dispatch); `cpu_baseline` times the CPU
FP64 oracle (OpenMP, all host threads) on a bounded sample of the headline workload.
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
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")
GOLDEN = os.path.join(ROOT, "tests", "golden")
# GMRES: history fill (5) + calibration (1) + the checked chunks growing 2, 4, 8, 16 steps to the steady
# check interval of 32 steps
CALIB_STEPS = 36

from cusuperhelium import workloads as W  # noqa: E402  (pure numpy: no device, no library)

DT, SEED = W.DT, W.SEED
surface = W.surface


_T0 = time.perf_counter()


def note(msg):
    """Progress on stderr (stdout carries only the result line)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def golden(name):
    """Golden oracle states (tests/golden/*.npz): steps, Z, phi, mu[, sizes, index]."""
    with np.load(os.path.join(GOLDEN, name)) as f:
        return {k: f[k] for k in f.files}


def errors(Z, P, mu, Zr, Pr, mur):
    return {"max_abs_dz": float(np.abs(Z - Zr).max()), "max_abs_dphi": float(np.abs(P - Pr).max()),
            "max_abs_dmu": float(np.abs(mu - mur).max())}


def trajectory_errors(e, g, dt):
    """Runs the handle from its current state through the fixture's step marks; errors at each mark,
    mu of the right-hand side at the last one."""
    out, done, dz, dp = {}, 0, [], []
    for k, m in enumerate(g["steps"]):
        e.step(dt, int(m) - done)
        done = int(m)
        Z, P = e.get_state()
        dz.append(float(np.abs(Z - g["Z"][k]).max()))
        dp.append(float(np.abs(P - g["phi"][k]).max()))
    _, _, d = e.rhs(diagnostics=True)
    out["steps"] = [int(m) for m in g["steps"]]
    out["max_abs_dz"] = dz
    out["max_abs_dphi"] = dp
    out["max_abs_dmu_final"] = float(np.abs(d["mu"] - g["mu"][-1]).max())
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(Z0, P0, budget_s=25.0, max_steps=40):
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


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary, if present."""
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f).get(kernel)
    except (OSError, ValueError):
        return None


def timed_steps(e, steps, warmup, calib=0):
    """W warm-up steps (at least `calib`), then the wall time of exactly `steps` steps."""
    e.step(DT, max(warmup, calib))
    e.sync()
    t0 = time.perf_counter()
    e.step(DT, steps)
    e.sync()
    return time.perf_counter() - t0


def batched_leg(c, steps, with_accuracy):
    sizes, Zs = W.c3()
    Z0 = np.concatenate(Zs)
    P0 = np.zeros(len(Z0))
    out = {"workload": "512 water surfaces, N uniform in {128,256,512,1024} (seeded), 8-mode draws, phi=0, "
                       "g=1, deep water, dt=0.002; one persistent launch per RK stage",
           "surfaces": len(sizes), "nodes_total": int(sum(sizes)), "steps": steps}
    for prec, name in ((c.MIXED, "mixed"), (c.FP64, "fp64")):
        with c.Ensemble(sizes, c.Props(), precision=prec, solver=c.BATCHED, tol=1e-12, max_iter=16) as e:
            e.set_state(Z0, P0)
            el = timed_steps(e, steps, 2)
            st = e.stats()
            e.kernel_timing(True)
            e.step(DT, 1)
            r = e.kernel_report("stage")
            e.kernel_timing(False)
        out[name] = {"ensemble_steps_per_s": round(steps / el, 3), "ms_per_step": round(1e3 * el / steps, 3),
                     "stage_kernel_ms": round(r["ms"] / max(r["launches"], 1), 4),
                     "gmres_iterations_last": st["max_iterations"], "refinement_sweeps_last": st["refinements"],
                     "unconverged": st["unconverged"]}
    # the plain FP64 stepper (the drop-in: blocked LU; and FP64 GMRES) one surface at a time
    plain = {}
    for solver, name, k, mi, cal in ((c.LU, "lu", 6, 8, 0), (c.GMRES, "gmres", 20, 8, CALIB_STEPS)):
        t = {}
        for n in sorted(set(sizes)):
            with c.Ensemble([n], c.Props(), solver=solver, tol=1e-12, max_iter=mi) as e:
                e.set_state(surface(n, SEED), np.zeros(n))
                t[n] = timed_steps(e, k, 2, cal) / k
        ens = 1.0 / sum(t[n] for n in sizes)
        plain[name] = {"ensemble_steps_per_s": round(ens, 4), "t_step_ms": {str(n): round(1e3 * v, 4) for n, v in t.items()}}
    out["plain_fp64_one_at_a_time"] = plain
    out["speedup_mixed_vs_plain_lu"] = round(out["mixed"]["ensemble_steps_per_s"] / plain["lu"]["ensemble_steps_per_s"], 1)
    out["speedup_mixed_vs_plain_gmres"] = round(out["mixed"]["ensemble_steps_per_s"] / plain["gmres"]["ensemble_steps_per_s"], 1)
    if with_accuracy:
        # every one of the 512 instances after 20 steps against the oracle's golden states, and two
        # surfaces of each N over 500 steps
        g = golden("c3_ensemble_512_20steps.npz")
        off = np.concatenate([[0], np.cumsum(sizes)])
        for prec, name in ((c.MIXED, "accuracy_mixed"), (c.FP64, "accuracy_fp64")):
            with c.Ensemble(sizes, c.Props(), precision=prec, solver=c.BATCHED, tol=1e-12, max_iter=16) as e:
                e.set_state(Z0, P0)
                e.step(DT, int(g["steps"]))
                Z, _ = e.get_state()
                _, _, d = e.rhs(diagnostics=True)
            dz = np.array([np.abs(Z[a:b] - g["Z"][a:b]).max() for a, b in zip(off[:-1], off[1:])])
            dm = np.array([np.abs(d["mu"][a:b] - g["mu"][a:b]).max() for a, b in zip(off[:-1], off[1:])])
            out[name] = {"instances": len(sizes), "steps": int(g["steps"]),
                         "max_abs_dz": {"max": float(dz.max()), "median": float(np.median(dz)),
                                        "worst_instance": int(np.argmax(dz))},
                         "max_abs_dmu": {"max": float(dm.max()), "median": float(np.median(dm))},
                         "instances_within_1e-9": int((dz <= 1e-9).sum())}
        # every instance over the config's 500 steps: the mixed form against the FP64 form (both
        # pinned to the oracle above), per instance
        fin = {}
        for prec, name in ((c.MIXED, "mixed"), (c.FP64, "fp64")):
            with c.Ensemble(sizes, c.Props(), precision=prec, solver=c.BATCHED, tol=1e-12, max_iter=16) as e:
                e.set_state(Z0, P0)
                e.step(DT, 500)
                fin[name] = e.get_state()
                _, _, d = e.rhs(diagnostics=True)
                fin[name] += (d["mu"],)
        dz = np.array([np.abs(fin["mixed"][0][a:b] - fin["fp64"][0][a:b]).max() for a, b in zip(off[:-1], off[1:])])
        dm = np.array([np.abs(fin["mixed"][2][a:b] - fin["fp64"][2][a:b]).max() for a, b in zip(off[:-1], off[1:])])
        out["mixed_vs_fp64_500_steps"] = {"instances": len(sizes), "steps": 500,
                                          "max_abs_dz": {"max": float(dz.max()), "median": float(np.median(dz)),
                                                         "worst_instance": int(np.argmax(dz))},
                                          "max_abs_dmu": {"max": float(dm.max()), "median": float(np.median(dm))},
                                          "instances_within_1e-9": int((dz <= 1e-9).sum())}
        g = golden("c3_sample_500steps.npz")
        idx = [int(i) for i in g["index"]]
        sub = [sizes[i] for i in idx]
        with c.Ensemble(sub, c.Props(), precision=c.MIXED, solver=c.BATCHED, tol=1e-12, max_iter=16) as e:
            e.set_state(np.concatenate([Zs[i] for i in idx]), np.zeros(sum(sub)))
            out["accuracy_mixed_500_steps"] = dict(surfaces=[f"#{i} N={sizes[i]}" for i in idx],
                                                   **trajectory_errors(e, g, DT))
    return out


def c1_leg(c, with_accuracy):
    """The reference's CPU-runnable case: N = 256, y = 0.1 cos x, phi = 0, water with g = 1 in deep
    water, 100 RK4 steps of dt = 0.01 from a fresh state (for GMRES this includes the steps that
    fill the warm-start history and set the trimmed sequence, and the graph captures), with the
    drop-in LU solver and with GMRES; the error of the final surface against the oracle's."""
    Z0, P0, dt, steps = W.c1()
    n = len(Z0)
    out = {"workload": "N=256 water surface y=0.1 cos x, phi=0, g=1, deep water, dt=0.01, 100 RK4 steps, FP64",
           "nodes": n, "steps": steps}
    g = golden("c1_water_n256.npz") if with_accuracy else None
    for solver, name in ((c.LU, "lu"), (c.GMRES, "gmres")):
        with c.Ensemble([n], c.Props(), solver=solver, tol=1e-12) as e:
            e.set_state(Z0, P0)
            e.step(dt, 2)   # warm-up: graph capture, first launches
            e.set_state(Z0, P0)
            e.sync()
            t0 = time.perf_counter()
            e.step(dt, steps)
            e.sync()
            el = time.perf_counter() - t0
            out[name] = {"steps_per_s": round(steps / el, 2), "ms_per_step": round(1e3 * el / steps, 4)}
            if g is not None:
                Z, P = e.get_state()
                _, _, d = e.rhs(diagnostics=True)
                out[name].update(errors(Z, P, d["mu"], g["Z"][0], g["phi"][0], g["mu"][0]))
    return out


def c2_lu_leg(c, steps):
    """The headline surface with the drop-in solver (blocked FP64 LU on MFMA): steps/s, the MFMA
    roofline of its panel kernel (each launch: one panel and the previous panel's trailing update)
    and the time of each LU kernel per step, and the error after 100 steps against the oracle's
    golden state."""
    Z0, P0 = W.c2()
    out = {"workload": "N=2048 water surface (the headline input), FP64, blocked LU (getrf/getrs semantics)"}
    with c.Ensemble([N], c.Props(), solver=c.LU) as e:
        e.set_state(Z0, P0)
        el = timed_steps(e, steps, 2)
        e.kernel_timing(True)
        e.step(DT, 1)
        e.sync()
        reps = {t: e.kernel_report(t) for t in ("deriv", "assemble", "lu_step", "lu_backsolve",
                                                "mu_deriv", "velocity")}
        e.kernel_timing(False)
        st = e.stats()
    r = reps["lu_step"]
    upd_ms = r["ms"] / max(r["launches"], 1)
    lu_flops = 4 * (2 / 3 * N ** 3 + 2 * N ** 2)
    lu_ms = sum(reps[t]["ms"] for t in ("lu_step", "lu_backsolve"))
    out.update({"steps_per_s": round(steps / el, 3), "ms_per_step": round(1e3 * el / steps, 4), "steps": steps,
                "kernel_ms_per_step": {t: round(v["ms"], 4) for t, v in reps.items() if v["launches"]},
                "lu_step_roofline": {"bound": "mfma", "achieved": round(r["flops"] / (r["ms"] * 1e-3) / 1e12, 3),
                                       "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                                       "frac": round(r["flops"] / (r["ms"] * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4),
                                       "avg_launch_ms": round(upd_ms, 5), "launches": r["launches"]},
                "lu_solve_tflops": round(lu_flops / (lu_ms * 1e-3) / 1e12, 3),
                "lu_solve_frac_of_fp64_mfma_peak": round(lu_flops / (lu_ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4),
                "gepp_fallbacks": st["fallback_solves"]})
    return out


def film_leg(c, steps, with_accuracy):
    """Helium film over a flat substrate: N = 4096, depth h0 = 1 with the image term, restoring
    acceleration 3 A / (h0 + y)^4 with A = 0.05, i.e. g_eff = 3 A / h0^4 = 0.15 scaling the reference's
    van der Waals term; 8-mode seeded surface with a_k ~ U(0, 0.05 / k); phi = 0; dt = 0.002."""
    props = W.C5_PROPS
    Z0, P0 = W.c5()
    n = len(Z0)
    out = {"workload": "helium film N=4096, h0=1 (image term), g_eff=3A/h0^4=0.15 (A=0.05), 8-mode seeded "
                       "surface a_k~U(0,0.05/k), phi=0, dt=0.002; FP64 GMRES, mixed GMRES (FP32 matrix, "
                       "FP64 r0 and refinement residual), FP64 LU", "nodes": n}
    for solver, prec, name, mi, k, cal in ((c.GMRES, c.FP64, "gmres", 24, steps, CALIB_STEPS),
                                           (c.GMRES, c.MIXED, "gmres_mixed", 24, steps, CALIB_STEPS),
                                           (c.LU, c.FP64, "lu", 8, max(2, steps // 25), 0)):
        with c.Ensemble([n], c.Props(**props), precision=prec, solver=solver, tol=1e-12, max_iter=mi) as e:
            e.set_state(Z0, P0)
            el = timed_steps(e, k, 1, cal)
            st = e.stats()
        out[name] = {"steps_per_s": round(k / el, 3), "ms_per_step": round(1e3 * el / k, 3), "steps": k,
                     "gmres_iterations_last": st["max_iterations"], "unconverged": st["unconverged"],
                     "redone_steps": st["redone_steps"], "launches_per_solve": st["sequence_length"],
                     "gepp_fallbacks": st["fallback_solves"]}
    if with_accuracy:
        g = golden("c5_film_n4096.npz")
        for prec, name in ((c.FP64, "accuracy_gmres"), (c.MIXED, "accuracy_gmres_mixed")):
            with c.Ensemble([n], c.Props(**props), precision=prec, solver=c.GMRES, tol=1e-12, max_iter=24) as e:
                e.set_state(Z0, P0)
                out[name] = trajectory_errors(e, g, DT)
        with c.Ensemble([n], c.Props(**props), solver=c.LU) as e:
            e.set_state(Z0, P0)
            e.step(DT, int(g["steps"][0]))
            Z, P = e.get_state()
            _, _, d = e.rhs(diagnostics=True)
            out["accuracy_lu"] = dict(steps=int(g["steps"][0]),
                                      **errors(Z, P, d["mu"], g["Z"][0], g["phi"][0], g["mu"][0]))
        # the config's 500 steps: the mixed form and FP64 GMRES against the plain FP64 stepper (the LU,
        # pinned to the oracle's golden states above)
        fin = {}
        for solver, prec, name in ((c.LU, c.FP64, "lu"), (c.GMRES, c.FP64, "gmres"), (c.GMRES, c.MIXED, "gmres_mixed")):
            with c.Ensemble([n], c.Props(**props), precision=prec, solver=solver, tol=1e-12, max_iter=24) as e:
                e.set_state(Z0, P0)
                e.step(DT, 500)
                Z, P = e.get_state()
                _, _, d = e.rhs(diagnostics=True)
                fin[name] = (Z, P, d["mu"])
        out["vs_plain_lu_500_steps"] = {k: errors(*fin[k], *fin["lu"]) for k in ("gmres", "gmres_mixed")}
        out["vs_plain_lu_500_steps"]["max_abs_y_lu"] = float(np.abs(fin["lu"][0].imag).max())
    return out


def adjoint_leg(c):
    """The differentiable-stepper workload: N = 1024 water surface y0 = sum_k c_k cos(k x) with 16
    seeded amplitudes c_k ~ U(0, 0.01), loss = 1/2 sum_j (y_j(T) - y*_j)^2 against a seeded target,
    60 RK4 steps of dt = 0.005 with a checkpoint every 20; the gradient d loss / d c through the
    hand-written adjoint (torch.autograd), timed, and checked against central differences (h = 1e-6)
    of the GPU stepper, all 32 perturbed surfaces advanced as one batched ensemble."""
    import torch
    from cusuperhelium.adjoint import Trajectory, evolve
    dt, steps, every, h = 0.005, 60, 20, 1e-6
    x, amp, tgt, basis = W.c4()
    n, modes = len(x), len(amp)
    target = torch.tensor(tgt @ basis)
    B = torch.tensor(basis)

    split = [0.0, 0.0]   # forward (trajectory with checkpoints) and backward (reverse sweep) seconds

    def gradient(tr):
        t0 = time.perf_counter()
        ct = torch.tensor(amp, dtype=torch.float64, requires_grad=True)
        Z0 = torch.tensor(x, dtype=torch.complex128) + 1j * (ct @ B).to(torch.complex128)
        ZT, _ = evolve(Z0, torch.zeros(n, dtype=torch.float64), tr, dt, steps)   # returns host tensors
        loss = 0.5 * torch.sum((ZT.imag - target) ** 2)
        t1 = time.perf_counter()
        loss.backward()
        t2 = time.perf_counter()
        split[0] += t1 - t0
        split[1] += t2 - t1
        return float(loss.detach()), ct.grad.numpy()

    def timed(solver, reps):
        # one handle for the warm-up (first launches, adjoint workspace) and the timed gradients
        with Trajectory([n], checkpoint_every=every, solver=solver) as tr:
            gradient(tr)
            split[0] = split[1] = 0.0
            t0 = time.perf_counter()
            for _ in range(reps):
                out = gradient(tr)
            return (time.perf_counter() - t0) / reps, out

    # the reverse pass over many workgroups (GMRES handle), and over one workgroup per surface (batched)
    el, (loss, grad) = timed(c.GMRES, 3)
    fwd_s, bwd_s = split[0] / 3, split[1] / 3
    el_b, (_, grad_b) = timed(c.BATCHED, 1)
    Z0s = []
    for k in range(modes):
        for sgn in (+1, -1):
            cc = amp.copy()
            cc[k] += sgn * h
            Z0s.append(x + 1j * (cc @ basis))
    with c.Ensemble([n] * (2 * modes), c.Props(), solver=c.BATCHED, tol=1e-14, max_iter=48) as e:
        e.set_state(np.concatenate(Z0s), np.zeros(2 * modes * n))
        e.step(dt, steps)
        ZT, _ = e.get_state()
    L = 0.5 * np.sum((ZT.imag.reshape(2 * modes, n) - target.numpy()) ** 2, axis=1)
    fd = (L[0::2] - L[1::2]) / (2 * h)
    return {"workload": "N=1024 water surface, 16 cosine amplitudes, loss 1/2 |y(T) - y*|^2, 60 RK4 steps of "
                        "dt=0.005, checkpoint every 20, gradient by the hand-written adjoint (reverse pass over "
                        "many workgroups; GMRES forward; device-resident checkpoints and cotangent; one handle, "
                        "mean of 3 gradients after a warm-up)",
            "gradient_s": round(el, 4), "ms_per_step_forward_and_backward": round(1e3 * el / steps, 3),
            "forward_s": round(fwd_s, 4), "backward_s": round(bwd_s, 4),
            "backward_over_forward": round(bwd_s / fwd_s, 2) if fwd_s > 0 else None,
            "loss": loss, "max_rel_err_vs_fd": float(np.abs(grad - fd).max() / np.abs(fd).max()),
            "one_workgroup_per_surface": {"gradient_s": round(el_b, 4),
                                          "ms_per_step_forward_and_backward": round(1e3 * el_b / steps, 3),
                                          "max_rel_err_vs_fd": float(np.abs(grad_b - fd).max() / np.abs(fd).max())}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the comparisons with the golden states")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched ensemble leg")
    ap.add_argument("--batched-steps", type=int, default=50)
    ap.add_argument("--no-c1", action="store_true", help="skip the N=256 100-step leg")
    ap.add_argument("--no-lu", action="store_true", help="skip the N=2048 LU leg")
    ap.add_argument("--lu-steps", type=int, default=20)
    ap.add_argument("--no-film", action="store_true", help="skip the N=4096 helium film leg")
    ap.add_argument("--film-steps", type=int, default=50)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-stepper leg")
    ap.add_argument("--solver", choices=("lu", "gmres"), default="gmres")
    ap.add_argument("--max-iter", type=int, default=32, help="GMRES iteration cap per solve")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        a = torch.empty(args.copy_probe << 20, dtype=torch.uint8, device="cuda").fill_(1)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": a.numel(), "copy_probe_bytes_written": b.numel()}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the stepper runs one "
              f"surface ensemble per GPU and this build has no multi-GPU path)", file=sys.stderr)
        sys.exit(2)

    import cusuperhelium as c

    acc = not args.no_accuracy
    Z0, P0 = W.c2()
    solver = c.GMRES if args.solver == "gmres" else c.LU
    calib = CALIB_STEPS if args.solver == "gmres" else 0
    e = c.Ensemble([N], c.Props(), precision=c.FP64, solver=solver, tol=1e-12, max_iter=args.max_iter)
    e.set_state(Z0, P0)
    el = timed_steps(e, args.steps, args.warmup, calib)
    value = args.steps / el
    ms_per_step = 1e3 * el / args.steps
    st_main = e.stats()
    note(f"headline: {value:.1f} steps/s")

    # live kernel timing: a few steps with a hipEvent pair on every launch (its start and end)
    e.kernel_timing(True)
    kt_steps = 3
    e.step(DT, kt_steps)
    e.sync()
    reps = {t: e.kernel_report(t) for t in ("deriv", "assemble", "lu_step", "lu_backsolve",
                                            "gm_iter", "gm_close", "gm_close_mu",
                                            "mu_deriv", "velocity")}
    st = e.stats()
    e.kernel_timing(False)
    if args.solver == "gmres":
        # dominant kernel: one Arnoldi step, w = M v over the stored FP64 matrix (8 N^2 B) plus the
        # basis vectors of the Gram-Schmidt step, (8 N^2 + 8 N (j+3)) B per launch
        dom, bound, peak, unit, key, scale = "gm_iter", "hbm", PEAK_HBM_GBS, "GB/s", "bytes", 1e9
        kname = "gm_iter_kernel"
    else:
        # one panel of the LU with the previous panel's trailing update, on FP64 MFMA
        dom, bound, peak, unit, key, scale = "lu_step", "mfma", PEAK_FP64_TFLOPS, "TFLOP/s", "flops", 1e12
        kname = "lu_step_kernel"
    r = reps[dom]
    avg_ms = r["ms"] / max(r["launches"], 1)
    work_per_launch = r[key] / max(r["launches"], 1)
    achieved = work_per_launch / (avg_ms * 1e-3) / scale if avg_ms > 0 else 0.0
    stages = {t: round(v["ms"] / kt_steps, 4) for t, v in reps.items() if v["launches"]}
    e.close()
    # FP64 rooflines of the other kernels of a stage, from their algorithmic operations per launch
    # (FMA = 2, multiply, add and reciprocal = 1) over the live launch durations:
    #   pair cotangent parts (sin/cos/exp differences, p, q, 1/(sa^2 + sh^2) with two Newton steps): 24
    #   assembly per ordered pair: 24 + the entry (4) + the M x0 accumulation (2) = 30, and 8 B stored
    #   velocity per ordered pair: 24 + the weight and two sums (5) + the mu' circulant pair (4) = 33
    #   derivatives per (row, column): 7 multiply-adds of the three circulant products = 14
    per_pair = {"deriv": 14.0, "assemble": 30.0, "velocity": 33.0}
    stage_roofs = {}
    if args.solver == "gmres":
        for t, fl in per_pair.items():
            rr = reps[t]
            if not rr["launches"]:
                continue
            ms = rr["ms"] / rr["launches"]
            tf = fl * N * N / (ms * 1e-3) / 1e12
            stage_roofs[t] = {"bound": "fp64_valu", "achieved": round(tf, 3), "peak": PEAK_FP64_TFLOPS,
                              "unit": "TFLOP/s", "frac": round(tf / PEAK_FP64_TFLOPS, 4),
                              "avg_launch_ms": round(ms, 5), "flops_per_launch": fl * N * N}
        if "assemble" in stage_roofs:
            ms = stage_roofs["assemble"]["avg_launch_ms"]
            stage_roofs["assemble"]["stores_gbs"] = round(8.0 * N * N / (ms * 1e-3) / 1e9, 1)
            stage_roofs["assemble"]["stores_frac_of_hbm"] = round(8.0 * N * N / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "RK4 steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "warmup_steps_run": max(args.warmup, calib),
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
                     "frac": round(achieved / peak, 4), "traffic": measured_traffic(kname), "kernel": kname,
                     "avg_launch_ms": round(avg_ms, 5), f"{key}_per_launch": work_per_launch,
                     "launches_timed": r["launches"]},
        "stage_ms_per_step": stages,
        "stage_rooflines": stage_roofs,
        "solver_stats": {"gmres_iterations_last": st["last_iterations"], "gmres_unconverged": st_main["unconverged"],
                         "redone_steps": st_main["redone_steps"], "launches_per_solve": st_main["sequence_length"],
                         "gepp_fallbacks": st["fallback_solves"]},
    }
    if acc:
        # the headline path over the workload's full 1000 steps against the oracle's golden states
        g = golden("c2_water_n2048.npz")
        with c.Ensemble([N], c.Props(), precision=c.FP64, solver=solver, tol=1e-12, max_iter=args.max_iter) as e:
            e.set_state(Z0, P0)
            out["accuracy"] = trajectory_errors(e, g, DT)
        note("headline accuracy done")
    if not args.no_lu and args.solver == "gmres":
        out["c2_lu"] = c2_lu_leg(c, args.lu_steps)
        if acc:
            g = golden("c2_water_n2048.npz")
            with c.Ensemble([N], c.Props(), solver=c.LU) as e:
                e.set_state(Z0, P0)
                e.step(DT, int(g["steps"][0]))
                Z, P = e.get_state()
            out["c2_lu"]["accuracy"] = {"steps": int(g["steps"][0]), "max_abs_dz": float(np.abs(Z - g["Z"][0]).max()),
                                        "max_abs_dphi": float(np.abs(P - g["phi"][0]).max())}
        note("C2 LU leg done")
    if not args.no_batched:
        out["batched"] = batched_leg(c, args.batched_steps, acc)
        note("batched leg done")
    if not args.no_c1:
        out["c1"] = c1_leg(c, acc)
        note("C1 leg done")
    if not args.no_film:
        out["film"] = film_leg(c, args.film_steps, acc)
        note("film leg done")
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg(c)
        note("adjoint leg done")
    if not args.no_cpu:
        steps, cel, Zc, Pc, threads = cpu_baseline(Z0, P0)
        out["cpu_baseline"] = {"value": round(steps / cel, 4), "unit": "RK4 steps/s", "cores": threads,
                               "kind": "port", "cpu_model": cpu_model(), "host_threads": os.cpu_count(),
                               "sample": f"{steps} RK4 steps of the same N=2048 workload on the CPU FP64 oracle "
                                         f"(OpenMP, {threads} threads, {cpu_model()}), {cel:.1f} s"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
