"""Barnes-Hut body-steps/s at N = 2^20 on one MI355X.

Workload (the Barnes-Hut configuration of the metric): N = 1,048,576 equal-mass
Plummer sphere (total mass 1, scale radius 1 AU, SplitMix64 seed 42, truncated
at the 0.999 mass fraction, centre-of-mass frame), softening 1e-4, 30-bit
Morton keys, Verlet dt = 1e-3.  A step is one full SoAVerletIntegrator.Step
on the device: kick-drift, tree rebuild (bbox, keys, sort, octree, mass) and
force traversal, kick-rotate.  Inputs are resident in HBM before timing.

Timed lines, each over K steps after W warm-up steps, bracketed by device
synchronisation.  Before its warm-up steps each line first evaluates the
forces of its resident state, untimed, for 0.3 s: after an idle spell the
GPU's clocks take about 20 ms of load to come back up (the traversal takes
1.75 ms on the first step after 1 s idle, 1.63 ms ten steps later and 1.55 ms
at once after 0.3 s of load), which would otherwise fall into the first timed
steps.  Force evaluations leave the trajectory unchanged: a step's first kick
uses the previous accelerations, which only the steps themselves rotate in.
  * value: FP32 split precision at theta = 0.5;
  * fp64_baseline: the plain FP64 GPU path (same sort, tree and traversal,
    FP64 arithmetic, theta = 0.5, no learned term) -- the ratio's denominator;
  * theta_tuned: FP32 split at a wider theta (default 0.65);
  * learned: FP32 split with the learned far-field correction at
    --theta-learned, the network trained by the library's own kernels on
    FP64 direct-sum residuals of a seeded subsample of about 65,536 target
    bodies (whole target groups of the same state's tree).
Force accuracy: median per-body relative error against an FP64 direct sum on
a seeded 4096-body sample, evaluated on the final state.

--workload merger: the learned-correction configuration instead (two Plummer
spheres of 2^21 bodies each, mass 0.5, a = 1 AU, centres 10 AU apart,
approaching at 0.3 x the escape speed), theta in {0.5, 0.7, 0.9} with and
without the network.

Prints one JSON line (value = fast-path body-steps/s).
"""
import argparse
import ctypes
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "celestial-mechanics-desktop_amd"))

import numpy as np  # noqa: E402

PEAK_FP32_VALU_TFLOPS = 157.3   # MI355X published FP32 vector peak
FLOPS_PER_INTERACTION = 20
FLOPS_PER_FAR_CELL = 76         # Taylor coefficients of one (group, cell) far-field term
FAR_FIELD_RATIO = 0.25          # group far-field expansion ratio of the FP32 split path (swept: 0.15-0.3)
LEAF_CAPACITY = 8               # bodies per leaf bucket (measured best at N = 2^20, theta 0.5-0.65)
GROUP_SIZE = 64                 # bodies per target group (one wavefront's targets)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_traverse_plummer1M_traffic.md")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_backend(cmhip, n, precision, theta, params=None, far=0.0):
    be = cmhip.HipPhysicsBackend(n, precision=precision, theta=theta, morton_levels=10,
                                 leaf_capacity=LEAF_CAPACITY, group_size=GROUP_SIZE,
                                 far_field_ratio=far)
    if params is not None:
        be.mlp_set_params(params)
        be.configure(learned_correction=1)
    return be


def warm_clocks(be, eps, seconds=0.3):
    """Untimed force evaluations of the resident state for `seconds`: the
    GPU's clocks ramp up over the first ~20 ms of load after idling."""
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        be.compute_forces_resident(eps)


def run_path(cmhip, soa, precision, theta, steps, warmup, eps, dt, sample, params=None,
             far=0.0):
    n = soa.count
    be = make_backend(cmhip, n, precision, theta, params, far)
    be.upload(soa)
    warm_clocks(be, eps)
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
    keep = den >= 1e-30     # ForceErrorAnalyzer.cs:108 skips near-zero exact forces
    err = num[keep] / den[keep]
    be.close()
    return {
        "body_steps_per_s": n * steps / wall,
        "ms_per_step": 1e3 * wall / steps,
        "force_err_median": float(np.median(err)),
        "force_err_mean": float(np.mean(err)),
        "force_err_max": float(np.max(err)),
        "interactions_per_body": st["acc_interactions"] / max(1, st["acc_evaluations"]) / n,
        # the library times every 4th evaluation with HIP events on its stream
        "ms_force_kernel_avg": st["acc_ms_force"] / max(1, st["acc_timed_evaluations"]),
        "ms_force_eval_avg": st["acc_ms_total"] / max(1, st["acc_timed_evaluations"]),
        "acc_interactions": st["acc_interactions"],
        "acc_far_cells": st["acc_far_cells"],
        "far_cells_per_eval": st["acc_far_cells"] / max(1, st["acc_evaluations"]),
        "acc_ms_force": st["acc_ms_force"],
        "evaluations": st["acc_evaluations"],
        "timed_evaluations": st["acc_timed_evaluations"],
        "n_nodes": st["n_nodes"],
        "n_groups": st["n_groups"],
        "cell_fraction": st["cell_interactions"] / max(1, st["cell_interactions"]
                                                          + st["body_interactions"]),
    }


def host_roundtrip(cmhip, soa, theta, eps, reps=5, far=0.0):
    """IPhysicsComputeBackend.ComputeForces as the C# host calls it every step:
    upload the host BodySoA, evaluate forces (FP32 split), download AccX/Y/Z.
    PCIe-inclusive; reported beside the device-resident value, never as it."""
    n = soa.count
    be = make_backend(cmhip, n, cmhip.PREC_FP32_SPLIT, theta, far=far)
    work = soa.copy()
    be.compute_forces(work, eps)
    t0 = time.perf_counter()
    for _ in range(reps):
        be.compute_forces(work, eps)
    wall = (time.perf_counter() - t0) / reps
    be.close()
    return {"ms_per_force_eval": 1e3 * wall, "body_evals_per_s": n / wall}


def train_network(cmhip, soa, theta, eps, n_sub=65536, seed=2024, epochs=20):
    """The learned correction's network, trained by the library's kernels on
    direct-sum residuals for a seeded subsample of about n_sub target bodies:
    whole target groups of the full-resolution tree of `soa`, drawn by seed, so
    the cells seen in training are the cells the traversal accepts."""
    n = soa.count
    be = make_backend(cmhip, n, cmhip.PREC_FP32_SPLIT, theta)
    be.upload(soa)
    be.compute_forces_resident(eps)
    n_groups = be.stats()["n_groups"]
    max_groups = max(1, int(round(n_sub * n_groups / n)))
    be.mlp_init(seed)
    t0 = time.perf_counter()
    npairs = be.mlp_collect(eps, max_groups=max_groups, seed=seed, max_pairs=1 << 23)
    losses = be.mlp_train(epochs=epochs, batch=1024, lr=2e-3, seed=seed)
    secs = time.perf_counter() - t0
    params = be.mlp_params()
    be.close()
    return params, {"pairs": npairs, "epochs": epochs, "loss_first": float(losses[0]),
                    "loss_last": float(losses[-1]), "seconds": secs,
                    "target_groups": max_groups,
                    "target_bodies_approx": int(round(max_groups * n / n_groups))}


def host_cpu():
    """Host cores available to this process and the CPU model."""
    nproc = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = nproc
    threads = avail
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        threads = min(threads, int(env))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": nproc, "affinity": avail, "threads": threads, "model": model}


def cpu_baseline(soa, eps, dt, host, steps=3, warmup=1):
    """Reference CPU path restated in C (oracle): SoAVerletIntegrator.Step around
    BarnesHutBackend.ComputeForces (serial insertion, OpenMP traversal on every
    available host core standing in for Parallel.For), `warmup` untimed steps
    then `steps` timed steps on the same N = 2^20 state."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    n = soa.count
    s = {"px": soa.pos_x[:n].copy(), "py": soa.pos_y[:n].copy(), "pz": soa.pos_z[:n].copy(),
         "vx": soa.vel_x[:n].copy(), "vy": soa.vel_y[:n].copy(), "vz": soa.vel_z[:n].copy(),
         "ax": np.zeros(n), "ay": np.zeros(n), "az": np.zeros(n),
         "oax": np.zeros(n), "oay": np.zeros(n), "oaz": np.zeros(n),
         "mass": soa.mass[:n].copy(), "radius": np.zeros(n),
         "active": np.ones(n, dtype=np.uint8)}
    threads = host["threads"]

    def force(st):
        return O.bh_csharp(st, eps, 0.5, nthreads=threads)[0]

    for _ in range(warmup):
        O.verlet_step(s, dt, force)
    t0 = time.perf_counter()
    for _ in range(steps):
        O.verlet_step(s, dt, force)
    t = (time.perf_counter() - t0) / steps
    return {"value": n / t, "unit": "body-steps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} timed Verlet steps after {warmup} warm-up of the same N={n} "
                      f"Plummer state, C# BarnesHutBackend restated in C (serial insertion, "
                      f"OpenMP traversal on {threads} threads, theta=0.5)",
            "seconds_per_step": t, "nproc": host["nproc"], "affinity_cores": host["affinity"],
            "cpu_model": host["model"]}


def config1_direct(cmhip, steps=1000, cpu_steps=100):
    """The plain method (CPU-runnable configuration): N = 1000 Plummer
    (M = 1, a = 1, SplitMix64 seed 42, centre-of-mass frame), FP64 direct sum
    in CpuSingleThreadBackend's Newton-3 order (k_direct_pairwise), softening
    1e-4, Verlet dt = 1e-3, `steps` device-resident steps: ms/step, energy
    drift, and the force error of the first evaluation against the FP64 full
    sweep (CpuParallelBackend order) and the oracle's Newton-3 restatement
    (bit for bit).  Beside it the oracle's single-thread restatement of
    CpuSingleThreadBackend + SoAVerletIntegrator, timed over `cpu_steps`."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    n, eps, dt = 1000, 1e-4, 1e-3
    soa = cmhip.generate_plummer(n, seed=42)
    st = {"px": soa.pos_x[:n].copy(), "py": soa.pos_y[:n].copy(), "pz": soa.pos_z[:n].copy(),
          "mass": soa.mass[:n].copy(), "radius": np.zeros(n), "active": np.ones(n, np.uint8)}
    be = cmhip.HipPhysicsBackend(n, kind=cmhip.BACKEND_DIRECT_PAIRWISE, precision=cmhip.PREC_FP64)
    work = soa.copy()
    be.compute_forces(work, eps)
    got = (work.acc_x[:n], work.acc_y[:n], work.acc_z[:n])
    sweep = O.direct_gk(st, eps)
    err = O.force_error(sweep, got)
    bit_exact = all(np.array_equal(a, b) for a, b in zip(got, O.direct_n3(st, eps)))
    be.upload(soa)
    e = be.energy(eps, 1)
    e0 = e["kinetic"] + e["potential"]
    be.step(eps, dt, 10)                  # warm-up (kernels loaded)
    be.upload(soa)
    t0 = time.perf_counter()
    be.step(eps, dt, steps)               # synchronises at the end
    wall = time.perf_counter() - t0
    e = be.energy(eps, 1)
    drift = (e["kinetic"] + e["potential"] - e0) / abs(e0)
    be.close()
    s = {"px": st["px"].copy(), "py": st["py"].copy(), "pz": st["pz"].copy(),
         "vx": soa.vel_x[:n].copy(), "vy": soa.vel_y[:n].copy(), "vz": soa.vel_z[:n].copy(),
         "ax": np.zeros(n), "ay": np.zeros(n), "az": np.zeros(n),
         "oax": np.zeros(n), "oay": np.zeros(n), "oaz": np.zeros(n),
         "mass": st["mass"].copy(), "active": np.ones(n, np.uint8)}
    t0 = time.perf_counter()
    for _ in range(cpu_steps):
        O.verlet_step(s, dt, lambda x: O.direct_n3(x, eps))
    cpu = (time.perf_counter() - t0) / cpu_steps
    pairs = n * (n - 1) // 2
    return {"workload": "plummer1000_direct_fp64_verlet", "n_bodies": n, "steps": steps,
            "ms_per_step": 1e3 * wall / steps, "body_steps_per_s": n * steps / wall,
            "energy_drift": drift, "force_err_vs_fp64_sweep": {k: err[k] for k in ("median", "max")},
            "bit_exact_vs_cpu_single_thread": bit_exact,
            "cpu_baseline": {"value": n / cpu, "unit": "body-steps/s", "cores": 1, "kind": "port",
                             "ms_per_step": 1e3 * cpu,
                             "sample": f"{cpu_steps} steps of CpuSingleThreadBackend "
                                       f"(Newton-3, {pairs} pairs) + SoAVerletIntegrator "
                                       f"restated in C, one thread"}}


def traffic_from_profile():
    """HBM bytes per k_traverse<FP32_SPLIT> launch measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy of known size), or None."""
    try:
        with open(TRAFFIC_FILE) as fh:
            m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", fh.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def copy_calibration(nbytes=1 << 28, reps=5):
    """Device-to-device copies of a known size (the PMC scale reference)."""
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0   # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": nbytes, "copies": reps}), flush=True)


def disk(args, cmhip):
    """Collision configuration: N = 262,144 cold thin disk (1 Msun star, a
    log-uniform 0.5-5 AU, Rayleigh e 0.02 / i 0.01, masses log-uniform 1e-9 -
    1e-6 Msun, radii from 3 g/cm^3 inflated 100x), FP64 state; a cycle is
    one Verlet step (FP64 Barnes-Hut, theta 0.5, the bit-exact path), collision
    detection, resolution in --collision-mode (MergeOnly by default) and
    compaction (in place on the device; the survivors' old indices stay
    there, as the cycle does not use them)."""
    eps, dt = 1e-4, 1e-3
    n = args.n if args.n_set else 262144
    soa = cmhip.generate_disk(n, seed=42, a_min=0.5, a_max=5.0, sigma_e=0.02, sigma_i=0.01,
                              m_min=1e-9, m_max=1e-6,
                              density=cmhip.ROCK_DENSITY_INFLATED)
    be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP64, theta=0.5,
                                 leaf_capacity=args.disk_leaf_capacity)
    mode = {"merge": cmhip.COLLISION_MERGE_ONLY, "bounce": cmhip.COLLISION_BOUNCE_ONLY,
            "realistic": cmhip.COLLISION_REALISTIC}[args.collision_mode]
    be.configure_collisions(mode=mode)
    be.upload(soa)
    warm_clocks(be, eps)

    def cycle():
        be.step(eps, dt, 1)
        _, ne = be.detect_collisions(0)
        m = be.resolve_collisions()
        be.compact(with_old_index=False)   # the cycle needs no host copy of the index map
        return ne, m

    for _ in range(args.warmup):
        cycle()
    events = merged = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ne, m = cycle()
        events += ne
        merged += m
    wall = time.perf_counter() - t0
    line = {"metric": f"collision cycles (step + detect + {args.collision_mode} resolve + compact) per second",
            "value": args.steps / wall, "unit": "cycles/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"disk_collisions_{args.collision_mode}", "n_bodies": n,
                       "theta": 0.5, "leaf_capacity": args.disk_leaf_capacity,
                       "softening": eps, "dt": dt, "seed": 42},
            "body_steps_per_s": n * args.steps / wall, "events": events, "merged": merged,
            "bodies_left": be.count}
    print(json.dumps(line), flush=True)


def ensemble(args, cmhip):
    """Batched ensemble: 4096 planetary systems, N uniform in [8, 2048] (a 1 Msun
    star plus bodies of log-uniform mass 1e-7 - 1e-3 Msun, a 0.1 - 30 AU, Rayleigh
    e 0.05), direct softened gravity, FP32 split pair terms with FP64 kick-drift,
    graph-replayed persistent launches (autotuned); per-system energy drift over
    the timed steps, FP64 path beside it."""
    eps, dt = 1e-4, 1e-3
    soa, off = cmhip.generate_ensemble(4096, seed=42)
    nb = np.diff(off).astype(np.int64)
    pairs = int((nb * (nb - 1)).sum())
    res = {}
    for name, prec in (("fp32_split", cmhip.PREC_FP32_SPLIT), ("fp64", cmhip.PREC_FP64)):
        ens = cmhip.HipEnsemble(off, precision=prec, threads=0, steps_per_launch=100)
        ens.upload(soa)
        ens.step(eps, dt, args.warmup)
        e0 = ens.energy(eps)
        t0 = time.perf_counter()
        ens.step(eps, dt, args.steps)
        wall = time.perf_counter() - t0
        e1 = ens.energy(eps)
        drift = np.abs((e1 - e0) / e0)
        st = ens.stats()
        res[name] = {"ms_per_step": 1e3 * wall / args.steps,
                     "interactions_per_s": pairs * args.steps / wall,
                     "tflops_20_per_pair": 20 * pairs * args.steps / wall / 1e12,
                     "energy_drift_median": float(np.median(drift)),
                     "energy_drift_max": float(drift.max()),
                     "threads": st["threads"], "cost_order": st["tuned_cost_order"]}
        log(json.dumps({name: res[name]}))
    f = res["fp32_split"]
    line = {"metric": "ensemble body-steps/s (4096 systems, direct)", "unit": "body-steps/s",
            "value": float(nb.sum()) * 1e3 / f["ms_per_step"], "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": f["ms_per_step"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32-split (f64 state)", "data": "synthetic",
            "config": {"workload": "planetary_ensemble_direct", "systems": 4096,
                       "bodies": int(nb.sum()), "pair_terms_per_step": pairs,
                       "softening": eps, "dt": dt, "seed": 42},
            "fp32_split": f, "fp64": res["fp64"],
            "speedup_vs_fp64": res["fp64"]["ms_per_step"] / f["ms_per_step"]}
    print(json.dumps(line), flush=True)


PEAK_FP64_VALU_TFLOPS = 78.6   # MI355X published FP64 vector peak
FLOPS_PER_PN_PAIR = 81          # PostNewtonian1Correction pair term, sqrt and divide as 1 each


def _time_forces(be, reps):
    be.compute_forces_resident(1e-4)
    be.stats()                               # synchronises
    t0 = time.perf_counter()
    for _ in range(reps):
        be.compute_forces_resident(1e-4)
    be.stats()
    return (time.perf_counter() - t0) / reps


def extensions(args, cmhip):
    """The rows that widen the hot path, at their measurement sizes.
    Direct FP64 sum at N = 65,536 (Plummer, masses log-uniform in
    [1e-9, 1], 16 body types) with the engine's global, per-type and adaptive
    softening; the 1PN decorator (k_pn) on top of it, velocities ~1 % of c;
    the Yoshida-4 step on the headline Barnes-Hut workload (N = 2^20, FP32
    split, theta = 0.5) beside the Verlet step.  Kernel times are wall-clock
    differences over `reps` device-resident evaluations."""
    import oracle as O   # the CPU baselines only
    eps = 1e-4
    n = args.n or 65536
    reps = 10
    soa = cmhip.generate_plummer(n, 1.0, 1.0, seed=42, mass_fraction_cut=0.999)
    rng = np.random.default_rng(42)
    soa.mass[:n] = 10.0 ** rng.uniform(-9, 0, n)
    soa.body_type[:n] = rng.integers(0, 16, n)
    c = 2.998e8 * 5.0226e6 / 1.496e11   # PhysicalConstants.C_Sim
    for f in ("vel_x", "vel_y", "vel_z"):
        getattr(soa, f)[:n] = rng.normal(0, 0.01 * c, n)
    table = rng.uniform(1e-4, 1e-2, 16)
    be = cmhip.HipPhysicsBackend(n, kind=cmhip.BACKEND_DIRECT, precision=cmhip.PREC_FP64)
    be.upload(soa)
    pairs = float(n) * (n - 1)
    res = {}
    for name, mode in (("global", cmhip.SOFTENING_GLOBAL),
                       ("per_type", cmhip.SOFTENING_PER_TYPE),
                       ("adaptive", cmhip.SOFTENING_ADAPTIVE)):
        be.set_softening(mode, type_softening=table, adaptive_scale=0.01)
        t = _time_forces(be, reps)
        res[name] = {"ms": 1e3 * t, "pairs_per_s": pairs / t}
        log(json.dumps({name: res[name]}))
    be.set_softening(cmhip.SOFTENING_GLOBAL)
    be.set_post_newtonian(True)
    t_pn = _time_forces(be, reps) - res["global"]["ms"] * 1e-3
    _, nw = be.pn_warnings(1)
    be.close()
    pn_tflops = FLOPS_PER_PN_PAIR * pairs / t_pn / 1e12
    res["pn"] = {"ms_k_pn": 1e3 * t_pn, "pairs_per_s": pairs / t_pn, "tflops": pn_tflops,
                 "warnings": nw,
                 "roofline": {"bound": "valu", "achieved": pn_tflops,
                              "peak": PEAK_FP64_VALU_TFLOPS, "unit": "TFLOP/s",
                              "frac": pn_tflops / PEAK_FP64_VALU_TFLOPS}}
    log(json.dumps({"pn": res["pn"]}))
    # CPU: the oracle's restatements (the reference's loops in C), one thread,
    # on the first 4096 bodies of the same state
    k = 4096
    st = {"px": soa.pos_x[:k].copy(), "py": soa.pos_y[:k].copy(), "pz": soa.pos_z[:k].copy(),
          "vx": soa.vel_x[:k].copy(), "vy": soa.vel_y[:k].copy(), "vz": soa.vel_z[:k].copy(),
          "mass": soa.mass[:k].copy(), "active": np.ones(k, dtype=np.uint8),
          "body_type": soa.body_type[:k].copy()}
    z = np.zeros(k)
    t0 = time.perf_counter()
    O.pn_apply(st, eps, z, z, z)
    t_cpu_pn = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.direct_softened(st, 2, scale=0.01)
    t_cpu_ad = time.perf_counter() - t0
    cpu_pairs = float(k) * (k - 1)
    res["pn"]["cpu_baseline"] = {"value": cpu_pairs / t_cpu_pn, "unit": "pairs/s", "cores": 1,
                                 "kind": "port",
                                 "sample": "PostNewtonian1Correction.ApplyCorrections on 4096 "
                                           "bodies of the same state"}
    res["adaptive"]["cpu_baseline"] = {"value": cpu_pairs / t_cpu_ad, "unit": "pairs/s",
                                       "cores": 1, "kind": "port",
                                       "sample": "Engine::compute_cpu_brute_forces (adaptive) "
                                                 "on 4096 bodies of the same state"}
    # Yoshida-4 vs Verlet on the headline Barnes-Hut workload
    nb = 1 << 20
    big = cmhip.generate_plummer(nb, 1.0, 1.0, seed=42, mass_fraction_cut=0.999)
    steps = max(2, args.steps // 4)
    yo = {}
    for name in ("verlet", "yoshida4"):
        b = make_backend(cmhip, nb, cmhip.PREC_FP32_SPLIT, 0.5)
        b.upload(big)
        fn = b.step if name == "verlet" else b.yoshida_step
        fn(eps, 1e-3, 1)
        b.stats()
        t0 = time.perf_counter()
        fn(eps, 1e-3, steps)
        b.stats()
        t = (time.perf_counter() - t0) / steps
        yo[name] = {"ms_per_step": 1e3 * t, "body_steps_per_s": nb / t}
        b.close()
    yo["yoshida4"]["force_evaluations_per_step"] = 3
    res["yoshida4_bh"] = yo
    log(json.dumps({"yoshida4_bh": yo}))
    # potential energy of the same state: Barnes-Hut (the engine's
    # compute_potential_bh_kernel, theta = 0.5) against the exact pair sum
    b = make_backend(cmhip, nb, cmhip.PREC_FP32_SPLIT, 0.5)
    b.upload(big)
    pe = {}
    for name, mode, r in (("barnes_hut", 2, 5), ("exact", 1, 1)):
        b.energy(eps, mode)
        t0 = time.perf_counter()
        for _ in range(r):
            e = b.energy(eps, mode)
        pe[name] = {"ms": 1e3 * (time.perf_counter() - t0) / r, "potential": e["potential"]}
    pe["barnes_hut"]["rel_err"] = abs(pe["barnes_hut"]["potential"] - pe["exact"]["potential"]) \
        / abs(pe["exact"]["potential"])
    # the snapshot's kinetic energy, momenta and centre of mass in one pass
    # (57 bytes per body read), and the maximum acceleration (25 bytes per
    # body); each call ends with a host read of the result
    b.compute_forces_resident(eps)
    for name, fn, per_body in (("energy_com", lambda: b.energy(eps, 0), 57),
                               ("max_acceleration", b.max_acceleration, 25)):
        fn()
        t0 = time.perf_counter()
        for _ in range(20):
            fn()
        t = (time.perf_counter() - t0) / 20
        pe[name] = {"ms": 1e3 * t, "GB_per_s": per_body * nb / t / 1e9}
    b.close()
    res["potential_energy"] = pe
    log(json.dumps({"potential_energy": pe}))
    line = {"metric": "1PN pair terms/s, FP64 direct sum at N=65536 (k_pn)",
            "value": res["pn"]["pairs_per_s"], "unit": "pairs/s", "n_gpus": 1,
            "steps": reps, "warmup": 1, "ms_per_step": res["pn"]["ms_k_pn"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": "pn_softening_yoshida", "n_bodies": n, "softening": eps,
                       "velocity_sigma_c": 0.01, "seed": 42},
            "roofline": res["pn"]["roofline"], "cpu_baseline": res["pn"]["cpu_baseline"],
            "softening_modes": {k: res[k] for k in ("global", "per_type", "adaptive")},
            "yoshida4_bh": yo, "potential_energy_1m": pe}
    print(json.dumps(line), flush=True)


def merger(args, cmhip):
    eps, dt = 1e-4, 1e-3
    n = args.n if args.n_set else 1 << 22
    soa = cmhip.generate_merger(n, seed=42)
    sample = np.sort(np.random.default_rng(4096).choice(n, 4096, replace=False)).astype(np.int32)
    res = {}
    for theta in (0.5, 0.7, 0.9):
        log(f"merger theta={theta} plain ...")
        plain = run_path(cmhip, soa, cmhip.PREC_FP32_SPLIT, theta, args.steps, args.warmup,
                         eps, dt, sample)
        log(f"merger theta={theta} with the group far field (ratio {args.far_field_ratio}) ...")
        farf = run_path(cmhip, soa, cmhip.PREC_FP32_SPLIT, theta, args.steps, args.warmup,
                        eps, dt, sample, far=args.far_field_ratio)
        log(f"merger theta={theta}: training on a 65,536-body subsample ...")
        params, tr = train_network(cmhip, soa, theta, eps, epochs=args.train_epochs)
        learned = run_path(cmhip, soa, cmhip.PREC_FP32_SPLIT, theta, args.steps, args.warmup,
                           eps, dt, sample, params=params)
        res[str(theta)] = {
            "plain": {k: plain[k] for k in ("body_steps_per_s", "ms_per_step",
                                            "force_err_median", "interactions_per_body")},
            "far_field": {k: farf[k] for k in ("body_steps_per_s", "ms_per_step",
                                               "force_err_median", "interactions_per_body",
                                               "far_cells_per_eval")},
            "learned": {k: learned[k] for k in ("body_steps_per_s", "ms_per_step",
                                                "force_err_median", "interactions_per_body")},
            "training": tr}
        log(json.dumps(res[str(theta)]))
    log("plain FP64 GPU path (theta=0.5) ...")
    base = run_path(cmhip, soa, cmhip.PREC_FP64_FAST, 0.5, args.steps, args.warmup, eps, dt,
                    sample)
    line = {"metric": "BH body-steps/s, learned far-field correction (two-Plummer merger)",
            "value": res["0.5"]["plain"]["body_steps_per_s"], "unit": "body-steps/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": res["0.5"]["plain"]["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32-split (f64 state)",
            "data": "synthetic",
            "config": {"workload": "two_plummer_merger_bh_verlet", "n_bodies": n,
                       "softening": eps, "dt": dt, "separation_au": 10.0,
                       "speed_fraction_of_escape": 0.3, "far_field_ratio": 0.0, "seed": 42},
            "fp64_baseline": {"value": base["body_steps_per_s"],
                              "force_err_median": base["force_err_median"]},
            "by_theta": res}
    print(json.dumps(line), flush=True)


def only_line(args, cmhip):
    """One timed line on the headline workload, for profiling runs: the plain
    FP64 GPU path (theta 0.5) or the learned correction at --theta-learned
    (network trained first)."""
    n = args.n or (1 << 20)
    eps, dt = 1e-4, 1e-3
    soa = cmhip.generate_plummer(n, 1.0, 1.0, seed=42, mass_fraction_cut=0.999,
                                 com_frame=True)
    sample = np.sort(np.random.default_rng(4096).choice(n, min(4096, n),
                                                         replace=False)).astype(np.int32)
    if args.only == "fp64":
        rec = run_path(cmhip, soa, cmhip.PREC_FP64_FAST, 0.5, args.steps, args.warmup, eps, dt,
                       sample)
        kernel = "k_traverse<FP64_FAST>"
        theta = 0.5
    else:
        theta = args.theta_learned or 0.75
        params, tr = train_network(cmhip, soa, theta, eps, epochs=args.train_epochs)
        rec = run_path(cmhip, soa, cmhip.PREC_FP32_SPLIT, theta, args.steps, args.warmup, eps,
                       dt, sample, params=params)
        kernel = "k_traverse<FP32_SPLIT, learned>"
    line = {"metric": f"BH body-steps/s at N=1M, {args.only} line only",
            "value": rec["body_steps_per_s"], "unit": "body-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": rec["ms_per_step"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64" if args.only == "fp64" else "f32-split (f64 state)",
            "data": "synthetic",
            "config": {"workload": "plummer_bh_verlet", "n_bodies": n, "theta": theta,
                       "leaf_capacity": LEAF_CAPACITY, "group_size": GROUP_SIZE, "seed": 42},
            "kernel": kernel, "ms_kernel_avg": rec["ms_force_kernel_avg"],
            "interactions_per_body": rec["interactions_per_body"],
            "force_err_median": rec["force_err_median"]}
    print(json.dumps(line), flush=True)


def main():
    global LEAF_CAPACITY, GROUP_SIZE
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--far-field-ratio", type=float, default=FAR_FIELD_RATIO,
                    help="group far-field expansion ratio of the FP32 split path (0: off)")
    ap.add_argument("--theta-tuned", type=float, default=0.65,
                    help="also time the fast path at this theta (0 to skip)")
    ap.add_argument("--theta-learned", type=float, default=0.75,
                    help="also time the learned correction at this theta (0 to skip)")
    ap.add_argument("--workload", choices=("plummer", "merger", "disk", "ensemble",
                                           "extensions"),
                    default="plummer")
    ap.add_argument("--collision-mode", choices=("merge", "bounce", "realistic"), default="merge",
                    help="CollisionResolver mode of --workload disk")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--fast-only", action="store_true",
                    help="time only the FP32 theta = 0.5 line (profiling runs)")
    ap.add_argument("--leaf-capacity", type=int, default=LEAF_CAPACITY,
                    help="bodies per leaf bucket of the Barnes-Hut lines")
    ap.add_argument("--disk-leaf-capacity", type=int, default=8,
                    help="bodies per leaf bucket of the disk workload's Barnes-Hut step "
                         "(8 measured fastest of 8-32)")
    ap.add_argument("--group-size", type=int, default=GROUP_SIZE,
                    help="bodies per target group of the Barnes-Hut paths (1..64)")
    ap.add_argument("--train-epochs", type=int, default=20,
                    help="epochs of the learned correction's training")
    ap.add_argument("--only", choices=("fast", "fp64", "learned"), default=None,
                    help="time only one line: the fast path, the FP64 baseline or the "
                         "learned correction (profiling runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run device copies of known size (PMC scale reference)")
    args = ap.parse_args()
    LEAF_CAPACITY = args.leaf_capacity
    GROUP_SIZE = args.group_size
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    args.n_set = args.n is not None
    import cmhip
    if args.workload == "ensemble":
        ensemble(args, cmhip)
        return
    if args.workload == "disk":
        disk(args, cmhip)
        return
    if args.workload == "merger":
        merger(args, cmhip)
        return
    if args.workload == "extensions":
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        extensions(args, cmhip)
        return
    if args.only == "fast":
        args.fast_only = True
    if args.only in ("fp64", "learned"):
        return only_line(args, cmhip)
    if args.fast_only:
        args.no_fp64 = True
        args.no_cpu_baseline = True
        args.theta_tuned = 0.0
        args.theta_learned = 0.0
    n = args.n or (1 << 20)
    eps, dt = 1e-4, 1e-3
    soa = cmhip.generate_plummer(n, 1.0, 1.0, seed=42, mass_fraction_cut=0.999,
                                 com_frame=True)
    sample = np.sort(np.random.default_rng(4096).choice(n, min(4096, n),
                                                         replace=False)).astype(np.int32)
    far = args.far_field_ratio
    log(f"fast path (FP32 split, theta={args.theta}, far-field ratio {far}) ...")
    fast = run_path(cmhip, soa, cmhip.PREC_FP32_SPLIT, args.theta, args.steps, args.warmup,
                    eps, dt, sample, far=far)
    log(json.dumps(fast))
    roundtrip = None
    plain = None
    if not args.fast_only:
        log("host round trip (upload, forces, download) ...")
        roundtrip = host_roundtrip(cmhip, soa, args.theta, eps, far=far)
        log(json.dumps(roundtrip))
        if far > 0.0:
            log(f"FP32 split without the far-field expansion (theta={args.theta}) ...")
            plain = run_path(cmhip, soa, cmhip.PREC_FP32_SPLIT, args.theta, args.steps,
                             args.warmup, eps, dt, sample)
            log(json.dumps(plain))
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
                         args.warmup, eps, dt, sample, far=far)
        log(json.dumps(tuned))
    learned = None
    if args.theta_learned:
        log(f"learned correction at theta={args.theta_learned}: training ...")
        params, tr = train_network(cmhip, soa, args.theta_learned, eps, epochs=args.train_epochs)
        log(json.dumps(tr))
        learned = run_path(cmhip, soa, cmhip.PREC_FP32_SPLIT, args.theta_learned, args.steps,
                           args.warmup, eps, dt, sample, params=params)
        learned["training"] = tr
        log(json.dumps(learned))
        log(f"plain FP32 split at theta={args.theta_learned} (the method without the network) ...")
        same = run_path(cmhip, soa, cmhip.PREC_FP32_SPLIT, args.theta_learned, args.steps,
                        args.warmup, eps, dt, sample)
        learned["without_network"] = {"value": same["body_steps_per_s"],
                                      "force_err_median": same["force_err_median"],
                                      "ms_per_step": same["ms_per_step"]}

    flops = (FLOPS_PER_INTERACTION * fast["acc_interactions"]
             + FLOPS_PER_FAR_CELL * fast["acc_far_cells"]) / max(1, fast["evaluations"])
    achieved = flops / (fast["ms_force_kernel_avg"] * 1e-3) / 1e12
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
        "config": {"workload": "plummer_bh_verlet", "n_bodies": n, "theta": args.theta,
                   "softening": eps, "dt": dt, "morton_bits": 30, "leaf_capacity": LEAF_CAPACITY,
                   "group_size": GROUP_SIZE, "plummer_mass": 1.0, "plummer_a_au": 1.0, "mass_fraction_cut": 0.999,
                   "far_field_ratio": far, "seed": 42},
        "force_err_median": fast["force_err_median"],
        "force_err_max": fast["force_err_max"],
        "interactions_per_body": fast["interactions_per_body"],
        "far_cells_per_eval": fast["far_cells_per_eval"],
        "cell_fraction": fast["cell_fraction"],
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_VALU_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_VALU_TFLOPS,
                     "traffic": traffic_from_profile(), "kernel": "k_traverse<FP32_SPLIT>",
                     "ms_kernel_avg": fast["ms_force_kernel_avg"],
                     "flops_per_launch": flops},
    }
    if plain:
        line["without_far_field"] = {"value": plain["body_steps_per_s"],
                                     "ms_per_step": plain["ms_per_step"],
                                     "force_err_median": plain["force_err_median"],
                                     "interactions_per_body": plain["interactions_per_body"],
                                     "ms_force_kernel_avg": plain["ms_force_kernel_avg"]}
    if roundtrip:
        line["host_roundtrip_pcie"] = roundtrip
    if base:
        line["fp64_baseline"] = {"value": base["body_steps_per_s"],
                                 "ms_per_step": base["ms_per_step"],
                                 "force_err_median": base["force_err_median"],
                                 "interactions_per_body": base["interactions_per_body"],
                                 "ms_force_kernel_avg": base["ms_force_kernel_avg"]}
        line["speedup_vs_fp64_gpu"] = fast["body_steps_per_s"] / base["body_steps_per_s"]
    for key, rec, th in (("theta_tuned", tuned, args.theta_tuned),
                         ("learned", learned, args.theta_learned)):
        if not rec:
            continue
        line[key] = {"theta": th, "value": rec["body_steps_per_s"],
                     "force_err_median": rec["force_err_median"],
                     "ms_per_step": rec["ms_per_step"],
                     "interactions_per_body": rec["interactions_per_body"]}
        if base:
            line[key]["speedup_vs_fp64_gpu"] = rec["body_steps_per_s"] / base["body_steps_per_s"]
        if "training" in rec:
            line[key]["training"] = rec["training"]
        if "without_network" in rec:
            line[key]["without_network"] = rec["without_network"]
    if not args.no_cpu_baseline:
        cpu_host = host = host_cpu()
        log(f"cpu baseline on {host['threads']} threads ({host['model']}) ...")
        cb = cpu_baseline(soa, eps, dt, cpu_host)
        line["cpu_baseline"] = cb
    if not args.fast_only:
        log("config 1: N=1000 direct FP64, 1000 steps ...")
        line["config1_direct"] = config1_direct(cmhip)
        log(json.dumps(line["config1_direct"]))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
