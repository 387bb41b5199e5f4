"""Centre of mass (cm_energy's center_of_mass, the engine's gpu_reduce_com)
and maximum acceleration (cm_max_acceleration, Engine::compute_max_acceleration)
on the GPU against the oracle."""
import numpy as np
import pytest

import cmhip
import oracle as O
from conftest import soa_to_state

pytestmark = pytest.mark.gpu

EPS = 1e-4


def _com_close(got, want, s):
    """Compensated device sums against the oracle's near-exact ones, each
    quotient rounded once: within 1e-14 of the largest |coordinate|."""
    scale = max(np.abs(s[k]).max() for k in ("px", "py", "pz")) if len(s["px"]) else 1.0
    for k in range(3):
        assert abs(got[k] - want[k]) <= 1e-14 * scale, (k, got[k], want[k])


@pytest.mark.parametrize("n,inactive", [(1, 0.0), (1000, 0.0), (200000, 0.1), (1 << 20, 0.02)])
def test_center_of_mass(plummer, n, inactive):
    soa = plummer(n, cut=0.999, com=False)
    rng = np.random.default_rng(n)
    soa.pos_x[:n] += 3.0   # move the COM away from the origin
    soa.is_active[:n][rng.random(n) < inactive] = 0
    s = soa_to_state(soa)
    print(f"centre of mass: n={n} inactive={inactive}", flush=True)
    be = cmhip.HipPhysicsBackend(n)
    be.upload(soa)
    e = be.energy(EPS, False)
    want = O.center_of_mass(s)
    _com_close(e["center_of_mass"], want[:3], s)
    assert abs(e["total_mass"] - want[3]) <= 4.5e-16 * want[3]


def test_center_of_mass_empty_and_all_inactive():
    be = cmhip.HipPhysicsBackend(16)
    assert be.energy(EPS, False)["center_of_mass"] == (0.0, 0.0, 0.0)
    soa = cmhip.BodySoA.from_arrays(pos_x=[1.0, 2.0], pos_y=[0.0, 5.0], pos_z=[0.0, 0.0],
                                    mass=[1.0, 3.0], is_active=[0, 0])
    be.upload(soa)
    e = be.energy(EPS, False)
    assert e["center_of_mass"] == (0.0, 0.0, 0.0) and e["total_mass"] == 0.0
    soa.is_active[:2] = 1
    be.upload(soa)
    assert be.energy(EPS, False)["center_of_mass"] == (1.75, 3.75, 0.0)


@pytest.mark.parametrize("n", [1, 777, 300000])
def test_max_acceleration_of_uploaded_accelerations(n):
    """Accelerations as uploaded (spanning 12 decades), a tenth of the bodies
    inactive, one inactive body holding the largest value: bit-exact (a
    maximum of correctly rounded magnitudes does not depend on order)."""
    rng = np.random.default_rng(n)
    a = [rng.normal(0, 1, n) * 10.0 ** rng.uniform(-6, 6, n) for _ in range(3)]
    act = (rng.random(n) >= 0.1).astype(np.uint8)
    act[0] = 1
    if n > 1:
        act[n // 2] = 0
        a[0][n // 2] = 1e30
    p = rng.normal(0, 1, (3, n))
    soa = cmhip.BodySoA.from_arrays(pos_x=p[0], pos_y=p[1], pos_z=p[2], mass=np.ones(n),
                                    acc_x=a[0], acc_y=a[1], acc_z=a[2], is_active=act)
    print(f"max acceleration: n={n}", flush=True)
    be = cmhip.HipPhysicsBackend(max(n, 16))
    be.upload(soa)
    assert be.max_acceleration() == O.max_acceleration(a, act)


def test_max_acceleration_after_forces_and_steps(plummer):
    """After a Barnes-Hut evaluation and after Verlet steps the engine's own
    accelerations are reduced: equal to the oracle's maximum over the
    downloaded arrays; an empty engine gives 0."""
    n = 50000
    soa = plummer(n, cut=0.999)
    be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP32_SPLIT, theta=0.5,
                                 leaf_capacity=8, far_field_ratio=0.15)
    be.upload(soa)
    for step in (False, True):
        if step:
            be.step(EPS, 1e-3, 3)
        else:
            be.compute_forces_resident(EPS)
        out = cmhip.BodySoA(n)
        be.download(out)
        s = soa_to_state(out)
        want = O.max_acceleration((s["ax"], s["ay"], s["az"]), s["active"])
        assert want > 0.0
        assert be.max_acceleration() == want
    assert cmhip.HipPhysicsBackend(16).max_acceleration() == 0.0
