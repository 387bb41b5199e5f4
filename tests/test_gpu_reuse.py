"""One engine reused across uploads of different body counts (larger,
smaller, with and without inactive bodies) gives, for every upload, the same
bits as a fresh engine on the same input: no per-evaluation buffer (keys,
sort histograms, node arrays, group counters, target-group marks, statistics)
carries state from one evaluation to the next.  Covers the FP64 Barnes-Hut
path, the FP32 split path with the far field and the direct sum, and
device-resident Verlet runs."""
import numpy as np
import pytest

import cmhip

pytestmark = pytest.mark.gpu

EPS = 1e-4
CAP = 90000


def _state(n, seed, inactive):
    soa = cmhip.generate_plummer(n, 1.0, 1.0, seed=seed, mass_fraction_cut=0.999, com_frame=True)
    if inactive:
        rng = np.random.default_rng(seed)
        soa.is_active[:n][rng.random(n) < inactive] = 0
    return soa


def _acc(soa):
    n = soa.count
    return np.stack([soa.acc_x[:n], soa.acc_y[:n], soa.acc_z[:n]]).copy()


KW = {
    "fp64": dict(precision=cmhip.PREC_FP64, theta=0.5, leaf_capacity=8),
    "fp32_far": dict(precision=cmhip.PREC_FP32_SPLIT, theta=0.5, leaf_capacity=8,
                     far_field_ratio=0.25),
    "direct": dict(kind=cmhip.BACKEND_DIRECT, precision=cmhip.PREC_FP64),
}
SEQ = [(60000, 0.0), (85000, 0.1), (3000, 0.0), (90000, 0.05), (1, 0.0), (40000, 0.2)]


@pytest.mark.parametrize("path", ["fp64", "fp32_far", "direct"])
def test_reused_engine_equals_fresh_engines(path):
    print(f"engine reuse: {path}", flush=True)
    be = cmhip.HipPhysicsBackend(CAP, **KW[path])
    for k, (n, inactive) in enumerate(SEQ):
        if path == "direct" and n > 40000:
            n = n // 4   # the O(n^2) path at a quarter of the size
        soa = _state(n, seed=100 + k, inactive=inactive)
        be.compute_forces(soa, EPS)
        got = _acc(soa)
        fresh = cmhip.HipPhysicsBackend(n, **KW[path])
        ref = _state(n, seed=100 + k, inactive=inactive)
        fresh.compute_forces(ref, EPS)
        np.testing.assert_array_equal(got, _acc(ref), err_msg=f"upload {k}: n={n}")
        if path != "direct":
            assert be.stats()["n_groups"] == fresh.stats()["n_groups"]
        fresh.close()
    be.close()


def test_reused_engine_verlet_runs_equal_fresh_engines():
    print("engine reuse: verlet runs", flush=True)
    kw = KW["fp32_far"]
    be = cmhip.HipPhysicsBackend(CAP, **kw)
    for k, (n, inactive) in enumerate(SEQ[:4]):
        soa = _state(n, seed=200 + k, inactive=inactive)
        be.upload(soa)
        be.step(EPS, 1e-3, 3)
        a = cmhip.BodySoA(n)
        be.download(a)
        fresh = cmhip.HipPhysicsBackend(n, **kw)
        fresh.upload(_state(n, seed=200 + k, inactive=inactive))
        fresh.step(EPS, 1e-3, 3)
        b = cmhip.BodySoA(n)
        fresh.download(b)
        for f in ("pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "acc_x", "old_acc_z"):
            np.testing.assert_array_equal(getattr(a, f)[:n], getattr(b, f)[:n],
                                          err_msg=f"run {k}: {f}")
        fresh.close()
    be.close()


@pytest.mark.parametrize("path", ["fp64", "fp32_far"])
def test_force_evaluations_before_steps_leave_the_trajectory_unchanged(path):
    """bench.py evaluates the forces of the resident state for a while before
    stepping (to bring the GPU's clocks up): a step's first kick reads the
    previous accelerations (OldAcc), which only a step rotates in, so the
    state after the steps is bit-identical with or without those evaluations."""
    n, steps = 30000, 3
    soa = _state(n, 31, 0.05)
    runs = []
    for extra in (0, 4):
        be = cmhip.HipPhysicsBackend(n, **KW[path])
        be.upload(soa)
        for _ in range(extra):
            be.compute_forces_resident(EPS)
        be.step(EPS, 1e-3, steps)
        out = cmhip.BodySoA(n)
        be.download(out)
        runs.append(out)
        be.close()
    for f in ("pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "acc_x", "acc_y", "acc_z",
              "old_acc_x", "old_acc_y", "old_acc_z"):
        np.testing.assert_array_equal(getattr(runs[0], f)[:n], getattr(runs[1], f)[:n],
                                      err_msg=f"{path} {f}")
