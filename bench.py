"""Barnes-Hut body-steps/s at N = 2^20 on one MI355X.

Workload (the Barnes-Hut configuration of the metric): N = 1,048,576 equal-mass
Plummer sphere (total mass 1, scale radius 1 AU, SplitMix64 seed 42, truncated
at the 0.999 mass fraction, centre-of-mass frame), softening 1e-4, 30-bit
Morton keys, Verlet dt = 1e-3.  A step is one full SoAVerletIntegrator.Step
on the device: kick-drift, tree rebuild (bbox, keys, sort, octree, mass) and
force traversal, kick-rotate.

The fast path (FP32 split precision) and the plain FP64 GPU path (same sort,
tree and traversal, FP64 arithmetic) are each timed over K steps after W
warm-up steps, bracketed by device synchronisation.  Force accuracy is the
median per-body relative error against an FP64 direct sum on a seeded
4096-body sample, evaluated on the final state.

Prints one JSON line (value = fast-path body-steps/s).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "celestial-mechanics-desktop_amd"))

import numpy as np  # noqa: E402

PEAK_FP32_VALU_TFLOPS = 157.3   # MI355X published FP32 vector peak
PEAK_FP64_VALU_TFLOPS = 78.6
FLOPS_PER_INTERACTION = 20


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_path(cmhip, soa, precision, theta, steps, warmup, eps, dt, sample):
    n = soa.count
    be = cmhip.HipPhysicsBackend(n, precision=precision, theta=theta, morton_levels=10,
                                 leaf_capacity=16, group_size=64)
    be.upload(soa)
    be.step(eps, dt, warmup)
    be.reset_stats()
    t0 = time.perf_counter()
    be.step(eps, dt, steps)          # synchronises at the end
    wall = time.perf_counter() - t0
    st = be.stats()
    # accuracy on the final state: the last step's accelerations vs FP64 direct sum
    out = cmhip.BodySoA(n)
    be.download(out, fields=("acc_x", "acc_y", "acc_z"))
    ex = be.direct_sample(eps, sample)
    ax, ay, az = out.acc_x[sample], out.acc_y[sample], out.acc_z[sample]
    num = np.sqrt((ax - ex[0]) ** 2 + (ay - ex[1]) ** 2 + (az - ex[2]) ** 2)
    den = np.sqrt(ex[0] ** 2 + ex[1] ** 2 + ex[2] ** 2)
    err = num / den
    be.close()
    return {
        "body_steps_per_s": n * steps / wall,
        "ms_per_step": 1e3 * wall / steps,
        "force_err_median": float(np.median(err)),
        "force_err_mean": float(np.mean(err)),
        "force_err_max": float(np.max(err)),
        "interactions_per_body": st["acc_interactions"] / max(1, st["acc_evaluations"]) / n,
        "ms_force_kernel_avg": st["acc_ms_force"] / max(1, st["acc_evaluations"]),
        "ms_force_eval_avg": st["acc_ms_total"] / max(1, st["acc_evaluations"]),
        "acc_interactions": st["acc_interactions"],
        "acc_ms_force": st["acc_ms_force"],
        "evaluations": st["acc_evaluations"],
        "n_nodes": st["n_nodes"],
    }


def cpu_baseline(soa, eps, dt, threads):
    """Reference CPU path restated in C (oracle): BarnesHutBackend with serial
    insertion + OpenMP traversal inside SoAVerletIntegrator.Step, one step."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    n = soa.count
    s = {"px": soa.pos_x[:n].copy(), "py": soa.pos_y[:n].copy(), "pz": soa.pos_z[:n].copy(),
         "vx": soa.vel_x[:n].copy(), "vy": soa.vel_y[:n].copy(), "vz": soa.vel_z[:n].copy(),
         "ax": np.zeros(n), "ay": np.zeros(n), "az": np.zeros(n),
         "oax": np.zeros(n), "oay": np.zeros(n), "oaz": np.zeros(n),
         "mass": soa.mass[:n].copy(), "radius": np.zeros(n),
         "active": np.ones(n, dtype=np.uint8)}
    t0 = time.perf_counter()
    O.verlet_step(s, dt, lambda st: O.bh_csharp(st, eps, 0.5, nthreads=threads)[0])
    t = time.perf_counter() - t0
    return {"value": n / t, "unit": "body-steps/s", "cores": threads, "kind": "port",
            "sample": f"1 Verlet step of the same N={n} Plummer state, C# BarnesHutBackend "
                      f"restated in C (serial insertion, OpenMP traversal, theta=0.5)",
            "seconds": t}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--theta-tuned", type=float, default=None,
                    help="also time the fast path at this theta")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import cmhip
    eps, dt = 1e-4, 1e-3
    soa = cmhip.generate_plummer(args.n, 1.0, 1.0, seed=42, mass_fraction_cut=0.999,
                                 com_frame=True)
    sample = np.sort(np.random.default_rng(4096).choice(args.n, min(4096, args.n),
                                                         replace=False)).astype(np.int32)
    log(f"fast path (FP32 split, theta={args.theta}) ...")
    fast = run_path(cmhip, soa, cmhip.PREC_FP32_SPLIT, args.theta, args.steps, args.warmup,
                    eps, dt, sample)
    log(json.dumps(fast))
    base = None
    if not args.no_fp64:
        log("plain FP64 GPU path (theta=0.5) ...")
        base = run_path(cmhip, soa, cmhip.PREC_FP64_FAST, 0.5, args.steps, args.warmup,
                        eps, dt, sample)
        log(json.dumps(base))
    tuned = None
    if args.theta_tuned:
        log(f"fast path at theta={args.theta_tuned} ...")
        tuned = run_path(cmhip, soa, cmhip.PREC_FP32_SPLIT, args.theta_tuned, args.steps,
                         args.warmup, eps, dt, sample)
        log(json.dumps(tuned))

    achieved = (FLOPS_PER_INTERACTION * fast["acc_interactions"]
                / (fast["acc_ms_force"] * 1e-3) / 1e12)
    line = {
        "metric": "BH body-steps/s at N=1M (theta=0.5, rel force err<=1e-3) vs plain FP64 GPU path",
        "value": fast["body_steps_per_s"],
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fast["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-split (f64 state)",
        "data": "synthetic",
        "config": {"workload": "plummer_bh_verlet", "n_bodies": args.n, "theta": args.theta,
                   "softening": eps, "dt": dt, "morton_bits": 30, "leaf_capacity": 16,
                   "plummer_mass": 1.0, "plummer_a_au": 1.0, "mass_fraction_cut": 0.999,
                   "seed": 42},
        "force_err_median": fast["force_err_median"],
        "force_err_max": fast["force_err_max"],
        "interactions_per_body": fast["interactions_per_body"],
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_VALU_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_VALU_TFLOPS,
                     "traffic": None, "kernel": "k_traverse<FP32_SPLIT>",
                     "ms_kernel_avg": fast["ms_force_kernel_avg"]},
    }
    if base:
        line["fp64_baseline"] = {"value": base["body_steps_per_s"],
                                 "ms_per_step": base["ms_per_step"],
                                 "force_err_median": base["force_err_median"],
                                 "interactions_per_body": base["interactions_per_body"],
                                 "ms_force_kernel_avg": base["ms_force_kernel_avg"]}
        line["speedup_vs_fp64_gpu"] = fast["body_steps_per_s"] / base["body_steps_per_s"]
    if tuned:
        line["theta_tuned"] = {"theta": args.theta_tuned, "value": tuned["body_steps_per_s"],
                               "force_err_median": tuned["force_err_median"],
                               "ms_per_step": tuned["ms_per_step"]}
        if base:
            line["theta_tuned"]["speedup_vs_fp64_gpu"] = (tuned["body_steps_per_s"]
                                                          / base["body_steps_per_s"])
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline on {threads} threads ...")
        cb = cpu_baseline(soa, eps, dt, threads)
        line["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
