"""GPU checks of the order in which the traversal takes its target groups
(k_group_order in csrc/k_traverse.hip).  Eighth q of the groups, in body
order, is XCD q's queue; within it the groups whose work in the previous
traversal (read at each group's first body) is at least twice the mean come
first, and the rest follow, each class in body order.  For the bit-exact FP64
path, groups with at least four times the mean work come first of all, as
parts that each compute a slice of the group's targets.  The order must be
exactly that, and it must leave every acceleration bit for bit unchanged: it
decides which wavefront computes a target and when, not what it computes."""
import ctypes as C

import numpy as np
import pytest

import cmhip

pytestmark = pytest.mark.gpu

EPS = 1e-4
XCDS = 8


def export_groups(be, n, with_bounds=False):
    """cm_test_groups (csrc/cm_testing.h); the engine's capacity is n."""
    fn = be._L.cm_test_groups
    fn.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                   C.c_void_p, C.POINTER(C.c_int32)]
    fn.restype = C.c_int32
    groups = np.zeros((n, 2), np.int32)
    queue = np.zeros((2 * n, 4), np.int32)
    qb = np.zeros(16, np.int32)
    cost = np.zeros(n, np.int32)
    order = np.zeros(n, np.int32)
    ng = C.c_int32()
    be._check(fn(be._h, 2 * n, groups.ctypes.data, queue.ctypes.data, qb.ctypes.data,
                 cost.ctypes.data, order.ctypes.data, C.byref(ng)), "cm_test_groups")
    ng = ng.value
    if with_bounds:
        return groups[:ng], queue, qb, cost, order
    return groups[:ng], queue, cost, order


def expected_queue(groups, cost, order, cap, split):
    """Queue entries and bounds that k_group_order must produce: per eighth,
    the parts of its fat groups (split only; unless they need more than the
    eighth's spare entries), its heavy groups, the rest, each in body order."""
    ng = len(groups)
    work = np.array([cost[order[s]] for s, _ in groups], np.int64)
    mean = np.maximum(1, work).sum() // ng
    spare = max(0, (2 * cap - ng) // XCDS) if split else 0
    out, bounds = {}, []
    for q in range(XCDS):
        lo, hi = q * ng // XCDS, (q + 1) * ng // XCDS
        parts, heavy, light = [], [], []
        fat_groups = []
        for g in range(lo, hi):
            s, e = groups[g]
            nm, w = e - s, work[g]
            P = 1
            if split and mean > 0 and w >= 4 * mean and nm >= 2:
                P = int(min(8, nm, w // (2 * mean)))
            if P > 1:
                fat_groups.append((g, P))
                parts += [(s, e, s + p * nm // P, s + (p + 1) * nm // P) for p in range(P)]
            elif w >= 2 * mean:
                heavy.append((s, e, s, e))
            else:
                light.append((s, e, s, e))
        if len(parts) - len(fat_groups) > spare:   # no room: fat groups stay whole
            heavy = sorted(heavy + [(groups[g][0], groups[g][1], groups[g][0], groups[g][1])
                                    for g, _ in fat_groups])
            parts = []
        base = lo + q * spare
        ent = parts + heavy + light
        out[q] = np.array(ent, np.int32).reshape(-1, 4)
        bounds.append((base, base + len(ent)))
    return out, bounds


def _check_queue(be, n, precision):
    """The queue the engine builds for its current state equals the expected
    one.  Returns (groups, entries that are parts, heavy groups)."""
    groups, queue, qb, cost, order = export_groups(be, n, with_bounds=True)
    exp, bounds = expected_queue(groups, cost, order, n, precision == cmhip.PREC_FP64)
    nparts = nheavy = 0
    for q in range(XCDS):
        assert (qb[q], qb[8 + q]) == bounds[q], q
        np.testing.assert_array_equal(queue[bounds[q][0]:bounds[q][1]], exp[q])
        whole = (exp[q][:, 2] == exp[q][:, 0]) & (exp[q][:, 3] == exp[q][:, 1])
        nparts += int((~whole).sum())
    work = np.array([cost[order[s]] for s, _ in groups], np.int64)
    nheavy = int((work >= 2 * (np.maximum(1, work).sum() // len(groups))).sum())
    return groups, nparts, nheavy


@pytest.mark.parametrize("precision", [cmhip.PREC_FP32_SPLIT, cmhip.PREC_FP64])
def test_queue_order_and_forces_unchanged(precision):
    n = 200_000
    print(f"group order: n={n} precision={precision}", flush=True)
    # a Plummer core dense enough that deepest-level leaves hold many bodies
    soa = cmhip.generate_plummer(n, 1.0, 1.0, seed=7, mass_fraction_cut=0.999, com_frame=True)
    far = 0.25 if precision == cmhip.PREC_FP32_SPLIT else 0.0
    be = cmhip.HipPhysicsBackend(n, precision=precision, theta=0.5, morton_levels=10,
                                 leaf_capacity=8, group_size=64, far_field_ratio=far)
    be.upload(soa)
    # a fresh engine has no costs: the queues are the groups in body order
    groups, queue, cost, order = export_groups(be, n)
    assert len(groups) > 1000
    assert not cost.any()
    assert _check_queue(be, n, precision)[1:] == (0, 0)
    be.compute_forces(soa, EPS)
    a1 = (soa.acc_x.copy(), soa.acc_y.copy(), soa.acc_z.copy())
    groups, queue, cost, order = export_groups(be, n)
    # every body of a group carries its group's work count
    for s, e in groups[:: max(1, len(groups) // 500)]:
        c = cost[order[s:e]]
        assert c[0] > 0 and np.all(c == c[0])
    w = np.array([cost[order[s]] for s, _ in groups])
    assert w.max() > 2 * np.median(w)   # dense-core groups do several times the median work
    _, nparts, nheavy = _check_queue(be, n, precision)
    assert nheavy > 0
    # the second evaluation runs in that order: same accelerations, bit for bit
    be.compute_forces(soa, EPS)
    for a, b in zip(a1, (soa.acc_x, soa.acc_y, soa.acc_z)):
        np.testing.assert_array_equal(a, b)
    be.close()


def test_fat_groups_split_into_parts_keep_fp64_bit_exact():
    """A thin disk whose sparse inner region packs groups into wide boxes (up
    to ~10x the mean work): the bit-exact FP64 path enters those groups as
    parts; the accelerations equal those of a fresh engine (no costs, no
    parts) bit for bit."""
    n = 262_144
    print(f"fat groups: disk n={n}", flush=True)
    soa = cmhip.generate_disk(n, seed=42, a_min=0.5, a_max=5.0, sigma_e=0.02, sigma_i=0.01,
                              m_min=1e-9, m_max=1e-6, density=cmhip.ROCK_DENSITY_INFLATED)
    be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP64, theta=0.5, leaf_capacity=8)
    be.upload(soa)
    fresh = soa.copy()
    be.compute_forces(fresh, EPS)            # no costs yet: every group whole
    runs = []
    for _ in range(2):
        out = soa.copy()
        be.compute_forces(out, EPS)          # with the previous evaluation's costs
        runs.append(out)
    _, nparts, _ = _check_queue(be, n, cmhip.PREC_FP64)
    assert nparts > 0
    for out in runs:
        for f in ("acc_x", "acc_y", "acc_z"):
            np.testing.assert_array_equal(getattr(out, f)[:n], getattr(fresh, f)[:n])
    # the Barnes-Hut potential rows walk the same queue (parts included)
    rows = be.potential_rows(EPS, method=2)
    clean = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP64, theta=0.5, leaf_capacity=8)
    clean.upload(soa)
    np.testing.assert_array_equal(rows, clean.potential_rows(EPS, method=2))
    clean.close()
    be.close()


def test_stale_costs_still_give_a_permutation_and_the_same_forces():
    """Costs are kept by the caller's body index: after the bodies change, the
    old costs only steer the order, and the forces equal a fresh engine's."""
    n = 60_000
    print(f"group order, stale costs: n={n}", flush=True)
    soa = cmhip.generate_plummer(n, 1.0, 1.0, seed=11, mass_fraction_cut=0.999, com_frame=True)
    be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP32_SPLIT, theta=0.5, morton_levels=10,
                                 leaf_capacity=8, group_size=64, far_field_ratio=0.25)
    be.upload(soa)
    be.compute_forces(soa, EPS)
    # a different state in the same slots: reversed body order
    rev = cmhip.BodySoA.from_arrays(pos_x=soa.pos_x[:n][::-1], pos_y=soa.pos_y[:n][::-1],
                                    pos_z=soa.pos_z[:n][::-1], mass=soa.mass[:n][::-1],
                                    is_active=np.ones(n))
    be.upload(rev)
    _check_queue(be, n, cmhip.PREC_FP32_SPLIT)
    be.compute_forces(rev, EPS)
    r1 = (rev.acc_x.copy(), rev.acc_y.copy(), rev.acc_z.copy())
    fresh = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP32_SPLIT, theta=0.5,
                                    morton_levels=10, leaf_capacity=8, group_size=64,
                                    far_field_ratio=0.25)
    fresh.upload(rev)
    fresh.compute_forces(rev, EPS)
    for a, b in zip(r1, (rev.acc_x, rev.acc_y, rev.acc_z)):
        np.testing.assert_array_equal(a, b)
    fresh.close()
    be.close()
