"""GPU checks of the order in which the traversal takes its target groups
(k_group_order in csrc/k_traverse.hip).  Eighth q of the groups, in body
order, is XCD q's queue; within it the groups whose work in the previous
traversal (read at each group's first body) is at least twice the mean come
first, and the rest follow, each class in body order.  The order must be that
stable partition of every eighth, and it must leave every acceleration bit for
bit unchanged: it decides which wavefront computes a group and when, not what
it computes."""
import ctypes as C

import numpy as np
import pytest

import cmhip

pytestmark = pytest.mark.gpu

EPS = 1e-4
XCDS = 8


def export_groups(be, n):
    """cm_test_groups (csrc/cm_testing.h)."""
    fn = be._L.cm_test_groups
    fn.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                   C.POINTER(C.c_int32)]
    fn.restype = C.c_int32
    groups = np.zeros((n, 2), np.int32)
    queue = np.zeros((n, 2), np.int32)
    cost = np.zeros(n, np.int32)
    order = np.zeros(n, np.int32)
    ng = C.c_int32()
    be._check(fn(be._h, n, groups.ctypes.data, queue.ctypes.data, cost.ctypes.data,
                 order.ctypes.data, C.byref(ng)), "cm_test_groups")
    ng = ng.value
    return groups[:ng], queue[:ng], cost, order


def _check_queue(groups, queue, cost, order):
    """Each eighth of the queue is its eighth of the groups with the heavy ones
    (work >= 2x the mean of max(1, work) over all groups) moved to the front,
    both parts in body order.  Returns the number of heavy groups."""
    ng = len(groups)
    work = np.array([cost[order[s]] for s, _ in groups], np.int64)
    thr = 2 * np.maximum(1, work).sum() // ng
    heavy = work >= thr
    for q in range(XCDS):
        lo, hi = q * ng // XCDS, (q + 1) * ng // XCDS
        h = heavy[lo:hi]
        expect = np.concatenate([groups[lo:hi][h], groups[lo:hi][~h]])
        np.testing.assert_array_equal(queue[lo:hi], expect)
    return int(heavy.sum())


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
    np.testing.assert_array_equal(queue, groups)
    assert _check_queue(groups, queue, cost, order) == 0
    be.compute_forces(soa, EPS)
    a1 = (soa.acc_x.copy(), soa.acc_y.copy(), soa.acc_z.copy())
    groups, queue, cost, order = export_groups(be, n)
    # every body of a group carries its group's work count
    for s, e in groups[:: max(1, len(groups) // 500)]:
        c = cost[order[s:e]]
        assert c[0] > 0 and np.all(c == c[0])
    w = np.array([cost[order[s]] for s, _ in groups])
    assert w.max() > 2 * np.median(w)   # dense-core groups do several times the median work
    assert _check_queue(groups, queue, cost, order) > 0
    assert not np.array_equal(queue, groups)
    # the second evaluation runs in that order: same accelerations, bit for bit
    be.compute_forces(soa, EPS)
    for a, b in zip(a1, (soa.acc_x, soa.acc_y, soa.acc_z)):
        np.testing.assert_array_equal(a, b)
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
    groups, queue, cost, order = export_groups(be, n)
    _check_queue(groups, queue, cost, order)
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
