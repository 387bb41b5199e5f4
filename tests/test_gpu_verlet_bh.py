"""Device-resident Verlet runs with the Barnes-Hut backend
(cm_verlet_step: kick-drift, forces, then the closing kick fused with the
next step's opening kick-drift, k_kick_rotate_drift) against
SoAVerletIntegrator.Step (Integrators/SoAVerletIntegrator.cs:48-115), every
array bit for bit:
  * FP64: the oracle's Verlet step around its Barnes-Hut restatement, over
    several steps with inactive bodies;
  * FP32 split (with the far field): the oracle's Verlet step around the
    library's own force evaluations (cm_compute_forces), so that the kicks,
    rotation and drift are compared exactly while the forces are the
    library's."""
import numpy as np
import pytest

import cmhip
import oracle as O
from conftest import soa_to_state

pytestmark = pytest.mark.gpu

EPS = 1e-4
DT = 1e-3
FIELDS = (("pos_x", "px"), ("pos_y", "py"), ("pos_z", "pz"), ("vel_x", "vx"),
          ("vel_y", "vy"), ("vel_z", "vz"), ("acc_x", "ax"), ("acc_y", "ay"),
          ("acc_z", "az"), ("old_acc_x", "oax"), ("old_acc_y", "oay"), ("old_acc_z", "oaz"))


def _state(n, seed=11):
    soa = cmhip.generate_plummer(n, 1.0, 1.0, seed=seed, mass_fraction_cut=0.999,
                                 com_frame=True)
    rng = np.random.default_rng(seed)
    soa.is_active[:n][rng.random(n) < 0.1] = 0
    return soa


def _compare(be, s, n):
    out = cmhip.BodySoA(n)
    be.download(out)
    for f, k in FIELDS:
        np.testing.assert_array_equal(getattr(out, f)[:n], s[k], err_msg=f)


@pytest.mark.parametrize("steps", [1, 4])
def test_verlet_barnes_hut_fp64_equals_oracle(steps):
    n = 20000
    print(f"verlet bh fp64: n={n} steps={steps}", flush=True)
    soa = _state(n)
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP64, theta=0.5, leaf_capacity=16,
                                 group_size=64)
    be.upload(soa)
    be.step(EPS, DT, steps)
    for _ in range(steps):
        O.verlet_step(s, DT, lambda st: O.bh_tree(st, EPS, 0.5, levels=10, leaf_cap=16,
                                                   group_size=64, nthreads=8)[0])
    _compare(be, s, n)
    be.close()


@pytest.mark.parametrize("steps", [1, 3])
def test_verlet_fp32_split_kicks_equal_oracle_kicks(steps):
    n = 30000
    print(f"verlet fp32 split: n={n} steps={steps}", flush=True)
    soa = _state(n, seed=12)
    s = soa_to_state(soa)
    kw = dict(precision=cmhip.PREC_FP32_SPLIT, theta=0.5, leaf_capacity=8, group_size=64,
              far_field_ratio=0.25)
    be = cmhip.HipPhysicsBackend(n, **kw)
    be.upload(soa)
    be.step(EPS, DT, steps)
    # the same steps with the library's forces and the oracle's kicks
    ref = cmhip.HipPhysicsBackend(n, **kw)
    work = cmhip.BodySoA(n)
    work.count = n

    def forces(st):
        for f, k in (("pos_x", "px"), ("pos_y", "py"), ("pos_z", "pz"), ("mass", "mass"),
                     ("radius", "radius"), ("is_active", "active")):
            getattr(work, f)[:n] = st[k]
        ref.compute_forces(work, EPS)
        return (work.acc_x[:n].copy(), work.acc_y[:n].copy(), work.acc_z[:n].copy())

    for _ in range(steps):
        O.verlet_step(s, DT, forces)
    _compare(be, s, n)
    be.close()
    ref.close()
