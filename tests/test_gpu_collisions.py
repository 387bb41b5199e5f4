"""GPU collision detection, MergeOnly resolution and compaction against the
oracle (bit-exact: same event set and order, same merged state)."""
import numpy as np
import pytest

import cmhip
import oracle as O
from conftest import soa_to_state

pytestmark = pytest.mark.gpu

MAP = {"pos_x": "px", "pos_y": "py", "pos_z": "pz", "vel_x": "vx", "vel_y": "vy",
       "vel_z": "vz", "acc_x": "ax", "acc_y": "ay", "acc_z": "az", "old_acc_x": "oax",
       "old_acc_y": "oay", "old_acc_z": "oaz", "mass": "mass", "radius": "radius",
       "density": "density", "is_active": "active", "is_collidable": "collidable",
       "body_type": "body_type"}


def crowded_disk(n, seed=5):
    # a narrow annulus so that many bodies overlap
    return cmhip.generate_disk(n, seed=seed, a_min=0.5, a_max=0.7, sigma_i=0.002)


def assert_state_equal(soa, s):
    n = soa.count
    for k_soa, k_or in MAP.items():
        np.testing.assert_array_equal(getattr(soa, k_soa)[:n], s[k_or], err_msg=k_soa)


@pytest.mark.parametrize("n", [20000, 3000])
def test_detect_bit_exact(n):
    """Event set and order equal the oracle's; 20,000 bodies give thousands of
    events (radix-sorted), 3,000 give ~100 (ordered in one workgroup)."""
    soa = crowded_disk(n)
    s = soa_to_state(soa)
    s["collidable"][::97] = 0
    soa.is_collidable[:soa.count] = s["collidable"]
    be = cmhip.HipPhysicsBackend(soa.count)
    be.upload(soa)
    ev, total = be.detect_collisions(1 << 16)
    ref = O.detect(s)
    assert total == len(ref) and total > 20
    assert (total > 2048) == (n == 20000)
    np.testing.assert_array_equal(ev["body_a"], ref["a"])
    np.testing.assert_array_equal(ev["body_b"], ref["b"])
    for f_gpu, f_ref in (("overlap_depth", "overlap"), ("distance", "dist"),
                         ("normal_x", "nx"), ("normal_z", "nz"),
                         ("relative_normal_speed", "rel_normal_speed")):
        np.testing.assert_array_equal(ev[f_gpu], ref[f_ref])


def test_resolve_and_compact_bit_exact():
    soa = crowded_disk(20000, seed=6)
    soa.body_type[5] = 7            # one black hole: type promotion path
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(soa.count)
    be.upload(soa)
    _, total = be.detect_collisions(0)
    merged = be.resolve_merges()
    ref = O.detect(s)
    ref_merged = O.resolve_merge(ref, s)
    assert merged == ref_merged and merged > 10
    out = cmhip.BodySoA(soa.count)
    be.download(out)
    assert_state_equal(out, s)
    n_new, old = be.compact()
    cs, idx = O.compact(s)
    assert n_new == len(idx)
    np.testing.assert_array_equal(old, idx)
    out2 = cmhip.BodySoA(soa.count)
    be.download(out2)
    assert_state_equal(out2, cs)


def test_collision_cycle_with_barnes_hut_bit_exact():
    """Config-3-like cycle: Verlet with FP64 Barnes-Hut, detect, merge, compact,
    repeated; the final state equals the oracle's sequence bit for bit (every
    other compaction leaves the survivors' old indices on the device)."""
    soa = crowded_disk(8000, seed=7)
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(soa.count, precision=cmhip.PREC_FP64, theta=0.5)
    be.upload(soa)
    eps, dt = 1e-4, 1e-3
    total_merged = 0
    for k in range(4):
        be.step(eps, dt, 1)
        be.detect_collisions(0)
        m = be.resolve_merges()
        total_merged += m
        n_new, old = be.compact(with_old_index=k % 2 == 0)
        assert (old is None) == (k % 2 == 1)
        O.verlet_step(s, dt, lambda st: O.bh_tree(st, eps, 0.5, levels=10, leaf_cap=16,
                                                  group_size=64)[0])
        assert O.resolve_merge(O.detect(s), s) == m
        s, idx = O.compact(s)
        assert n_new == len(idx)
        if old is not None:
            np.testing.assert_array_equal(old, idx)
    out = cmhip.BodySoA(soa.count)
    be.download(out)
    assert out.count == len(s["px"]) and total_merged > 0
    assert_state_equal(out, s)


def test_no_collidable_bodies():
    soa = crowded_disk(2000)
    soa.is_collidable[:] = 0
    be = cmhip.HipPhysicsBackend(soa.count)
    be.upload(soa)
    ev, total = be.detect_collisions(16)
    assert total == 0 and be.resolve_merges() == 0


def mixed_outcome_disk(n, seed):
    """A crowded annulus with random body types (a fifth compact) and velocity
    kicks of 1e-3 - 1e-1 AU/TU: in Realistic mode its ~6000 events split into
    merges, fragmentations, catastrophic disruptions and accretions."""
    soa = crowded_disk(n, seed=seed)
    rng = np.random.default_rng(seed)
    m = soa.count
    for f in ("vel_x", "vel_y", "vel_z"):
        getattr(soa, f)[:m] += rng.normal(0, 1, m) * 10.0 ** rng.uniform(-3, -1, m)
    soa.body_type[:m] = rng.integers(0, 10, m)
    return soa


@pytest.mark.parametrize("mode,promote,frame_cap,body_cap", [
    (cmhip.COLLISION_BOUNCE_ONLY, False, 0, 0),
    (cmhip.COLLISION_REALISTIC, False, 0, 0),
    (cmhip.COLLISION_REALISTIC, True, 0, 0),
    (cmhip.COLLISION_MERGE_ONLY, True, 0, 0),
    (cmhip.COLLISION_MERGE_ONLY, False, 64, 2),
    (cmhip.COLLISION_MERGE_ONLY, False, 0, 1),
])
def test_resolve_outcomes_bit_exact(mode, promote, frame_cap, body_cap):
    """CollisionResolver.Resolve in every mode (and the engine's merge caps)
    on the GPU equals the oracle's sequential loop bit for bit."""
    soa = mixed_outcome_disk(20000, seed=8)
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(soa.count)
    be.upload(soa)
    _, total = be.detect_collisions(0)
    be.configure_collisions(mode=mode, promote_compact_remnants=promote,
                            max_merges_per_frame=frame_cap, max_merges_per_body=body_cap)
    k = be.resolve_collisions()
    ref = O.detect(s)
    assert total == len(ref) and total > 1000
    kr = O.resolve_events(ref, s, mode=mode, promote=promote, max_per_frame=frame_cap,
                          max_per_body=body_cap)
    assert k == kr and k > 0
    out = cmhip.BodySoA(soa.count)
    be.download(out)
    assert_state_equal(out, s)
    if frame_cap:
        assert k == frame_cap


@pytest.mark.parametrize("mode", [cmhip.COLLISION_MERGE_ONLY, cmhip.COLLISION_REALISTIC])
def test_nbody_step_with_collisions_bit_exact(mode):
    """cm_nbody_step = n x NBodySolver.StepSoA (Verlet with FP64 Barnes-Hut,
    detect, resolve in the configured mode, no compaction) against the
    oracle's sequence, bit for bit."""
    soa = mixed_outcome_disk(8000, seed=9)
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(soa.count, precision=cmhip.PREC_FP64, theta=0.5)
    be.configure_collisions(mode=mode, promote_compact_remnants=True)
    be.upload(soa)
    eps, dt = 1e-4, 1e-3
    k = be.nbody_step(eps, dt, 3, collisions=True)
    kr = 0
    for _ in range(3):
        O.verlet_step(s, dt, lambda st: O.bh_tree(st, eps, 0.5, levels=10, leaf_cap=16,
                                                  group_size=64)[0])
        kr += O.resolve_events(O.detect(s), s, mode=mode, promote=True)
    assert k == kr and k > 0
    out = cmhip.BodySoA(soa.count)
    be.download(out)
    assert_state_equal(out, s)


def test_disk_configuration_cycles_at_full_size():
    """The collision configuration at its own size: N = 262,144 cold thin disk
    (1 Msun star, a log-uniform 0.5-5 AU, Rayleigh e 0.02 / i 0.01, masses
    log-uniform 1e-9 - 1e-6 Msun, radii from 3 g/cm^3 inflated 100x, seed 42,
    the benchmark's generator); two cycles of Verlet with FP64 Barnes-Hut
    (theta 0.5), detection, MergeOnly resolution and compaction equal the
    oracle's sequence bit for bit (event counts, survivor map, every array).
    The oracle's detection is the brute-force pair test over all 3.4e10 pairs."""
    n = 262144
    print(f"disk cycles: n={n}", flush=True)
    soa = cmhip.generate_disk(n, seed=42, a_min=0.5, a_max=5.0, sigma_e=0.02, sigma_i=0.01,
                              m_min=1e-9, m_max=1e-6, density=cmhip.ROCK_DENSITY_INFLATED)
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP64, theta=0.5)
    be.upload(soa)
    eps, dt = 1e-4, 1e-3
    for cycle in range(2):
        be.step(eps, dt, 1)
        _, ne = be.detect_collisions(0)
        merged = be.resolve_merges()
        n_new, old = be.compact()
        O.verlet_step(s, dt, lambda st: O.bh_tree(st, eps, 0.5, levels=10, leaf_cap=16,
                                                  group_size=64, nthreads=0)[0])
        ev = O.detect(s)
        ref_merged = O.resolve_merge(ev, s)
        s, idx = O.compact(s)
        print(f"cycle {cycle}: events {ne}, merged {merged}, bodies {n_new}", flush=True)
        assert ne == len(ev) and merged == ref_merged and merged > 0
        assert n_new == len(idx)
        np.testing.assert_array_equal(old, idx)
    out = cmhip.BodySoA(n)
    be.download(out)
    assert_state_equal(out, s)
