"""Long runs: the direct-sum configuration over 1000 Verlet steps equals the
CPU restatement of CpuSingleThreadBackend + SoAVerletIntegrator bit for bit;
the 10k-step energy drift of the GPU paths matches the reference's stability
claims and the FP32 split path drifts no more than the FP64 path that
reproduces the reference's Barnes-Hut arithmetic."""
import numpy as np
import pytest

import cmhip
import oracle as O
from conftest import soa_to_state

pytestmark = pytest.mark.gpu


def _energy(be, eps):
    e = be.energy(eps)
    return e["kinetic"] + e["potential"]


def test_direct_plummer_1000_steps_bit_exact():
    # N = 1000 Plummer (M = 1, a = 1, centre-of-mass frame), eps = 1e-4,
    # dt = 1e-3, 1000 steps, CpuSingleThreadBackend's Newton-3 pair order
    eps, dt = 1e-4, 1e-3
    soa = cmhip.generate_plummer(1000, seed=42)
    be = cmhip.HipPhysicsBackend(1000, kind=cmhip.BACKEND_DIRECT_PAIRWISE,
                                 precision=cmhip.PREC_FP64)
    be.upload(soa)
    e0 = _energy(be, eps)
    be.step(eps, dt, 1000)
    e1 = _energy(be, eps)
    out = cmhip.BodySoA(1000)
    be.download(out)
    st = soa_to_state(soa)
    for _ in range(1000):
        O.verlet_step(st, dt, lambda s: O.direct_n3(s, eps))
    for f, k in (("pos_x", "px"), ("pos_y", "py"), ("pos_z", "pz"), ("vel_x", "vx"),
                 ("vel_y", "vy"), ("vel_z", "vz"), ("old_acc_z", "oaz")):
        np.testing.assert_array_equal(getattr(out, f)[:1000], st[k])
    print(f"N=1000 direct FP64, 1000 steps: |dE/E| = {abs((e1 - e0) / e0):.3e}")


@pytest.mark.parametrize("kind,prec", [
    (cmhip.BACKEND_DIRECT_PAIRWISE, cmhip.PREC_FP64),
    (cmhip.BACKEND_BARNES_HUT, cmhip.PREC_FP64),
    (cmhip.BACKEND_BARNES_HUT, cmhip.PREC_FP32_SPLIT),
])
def test_two_body_circular_orbit_10k_steps(kind, prec):
    # SoAStabilityTests.cs:69-91: circular binary, 10k steps, |dE/E| < 1e-4
    # (the README's < 0.01 % claim)
    m = 0.5
    r = 1.0
    v = np.sqrt(m / (2 * r))        # each body orbits the centre at radius r/2
    soa = cmhip.BodySoA.from_arrays(
        pos_x=np.array([-r / 2, r / 2]), vel_y=np.array([-v, v]),
        mass=np.array([m, m]), is_active=np.ones(2, np.uint8))
    be = cmhip.HipPhysicsBackend(2, kind=kind, precision=prec)
    be.upload(soa)
    eps = 1e-4
    e0 = _energy(be, eps)
    be.step(eps, 1e-3, 10000)
    e1 = _energy(be, eps)
    assert abs((e1 - e0) / e0) < 1e-4


def test_plummer_energy_drift_10k_fp32_vs_fp64():
    # N = 4096 Plummer with softening 1e-2 (encounters resolved by dt = 1e-3):
    # over 10k steps the FP32 split traversal, with and without the group
    # far-field expansion (kappa = 0.15, the benchmark's), drifts no more than
    # twice the FP64 traversal (the reference's arithmetic), plus rounding noise
    eps, dt, n = 1e-2, 1e-3, 4096
    soa = cmhip.generate_plummer(n, seed=7, mass_fraction_cut=0.999)
    drift = {}
    for key, prec, far in (("fp64", cmhip.PREC_FP64, 0.0), ("fp32", cmhip.PREC_FP32_SPLIT, 0.0),
                           ("fp32_far", cmhip.PREC_FP32_SPLIT, 0.15)):
        be = cmhip.HipPhysicsBackend(n, precision=prec, theta=0.5, far_field_ratio=far)
        be.upload(soa)
        e0 = _energy(be, eps)
        be.step(eps, dt, 10000)
        drift[key] = abs((_energy(be, eps) - e0) / e0)
    print(f"N=4096 Plummer, 10k steps: |dE/E| FP64 {drift['fp64']:.3e}, "
          f"FP32 split {drift['fp32']:.3e}, FP32 split + far field {drift['fp32_far']:.3e}")
    assert drift["fp32"] <= 2.0 * drift["fp64"] + 1e-6
    assert drift["fp32_far"] <= 2.0 * drift["fp64"] + 1e-6
    assert drift["fp64"] < 1e-3


@pytest.mark.parametrize("name,prec,leaf,far", [
    ("fp64", cmhip.PREC_FP64, 16, 0.0),
    ("fp32_split", cmhip.PREC_FP32_SPLIT, 8, 0.0),
    ("fp32_split_far", cmhip.PREC_FP32_SPLIT, 8, 0.15),
])
def test_energy_drift_no_worse_than_reference_path(name, prec, leaf, far):
    """10k-step energy drift judged against the reference's own path, after the
    pattern of SoAStabilityTests.cs:177-218 (drift compared with a baseline
    run): tests/golden/drift_bh_csharp_plummer2048.json holds |dE/E| of the
    oracle's restatement of SoAVerletIntegrator.Step + BarnesHutBackend
    (serial insertion octree, theta 0.5) every 1000 steps on the same state
    (N = 2048 Plummer, seed 42, softening 1e-2, dt = 1e-3), written by
    tests/golden/make_drift.py.  Each GPU path's largest |dE/E| over the same
    checkpoints must not exceed the reference's (slack 1.0: no worse)."""
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden",
                           "drift_bh_csharp_plummer2048.json")) as fh:
        ref = json.load(fh)
    n, eps, dt = ref["n"], ref["softening"], ref["dt"]
    soa = cmhip.generate_plummer(n, 1.0, 1.0, seed=ref["seed"], mass_fraction_cut=0.999,
                                 com_frame=True)
    be = cmhip.HipPhysicsBackend(n, precision=prec, theta=ref["theta"], leaf_capacity=leaf,
                                 far_field_ratio=far)
    be.upload(soa)
    e0 = _energy(be, eps)
    assert abs(e0 - ref["e0"]) <= 1e-13 * abs(ref["e0"])   # same initial state
    worst = 0.0
    for k in range(10):
        be.step(eps, dt, 1000)
        worst = max(worst, abs((_energy(be, eps) - e0) / e0))
    print(f"{name}: max |dE/E| over 10k steps {worst:.3e}, reference path "
          f"{ref['drift_max_abs']:.3e}", flush=True)
    assert worst <= 1.0 * ref["drift_max_abs"], (worst, ref["drift_max_abs"])
