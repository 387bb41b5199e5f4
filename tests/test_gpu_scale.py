"""GPU parity at the Barnes-Hut benchmark's own size: N = 2^20 equal-mass
Plummer (mass 1, a = 1, truncated at the 0.999 mass fraction, SplitMix64 seed
42, centre-of-mass frame), 30-bit Morton keys, leaves of 8, groups of <= 64.

At this size three code paths run that smaller inputs never reach: the radix
sort's scan over more than one tile per thread (more than 524,288 keys), the
bounding box's grid-stride loop (more than 262,144 bodies) and the traversal's
per-XCD queues over ~24k target groups.  The tests compare, against the CPU
oracle on the same input:
  * Morton keys, the stable sort order and every tree field (first child,
    child count, level, body range, mass, centre of mass) bit for bit -- the
    stages that replace the reference engine's morton_code_kernel and radix
    sort (engine/src/cuda/morton_kernel.cu:11-45,185-230) and its tree build
    (engine/src/cuda/gpu_tree.cu:269-464);
  * one FP64 Barnes-Hut evaluation bit for bit, with equal interaction counts;
  * the benchmark's fast path (FP32 split precision, group far field
    kappa = 0.25, and 0.15): per-body relative difference median <= 2e-6,
    max <= 1e-3, equal cell, body and far-cell counts."""
import ctypes as C

import numpy as np
import pytest

import cmhip
import oracle as O
from conftest import soa_to_state

pytestmark = pytest.mark.gpu

N = 1 << 20
EPS = 1e-4
LEAF = 8


def _bench_state(inactive_frac=0.0):
    soa = cmhip.generate_plummer(N, 1.0, 1.0, seed=42, mass_fraction_cut=0.999,
                                 com_frame=True)
    if inactive_frac > 0.0:
        rng = np.random.default_rng(77)
        soa.is_active[:N][rng.random(N) < inactive_frac] = 0
    return soa


def _acc(soa):
    n = soa.count
    return soa.acc_x[:n].copy(), soa.acc_y[:n].copy(), soa.acc_z[:n].copy()


def _rel(a, b):
    num = np.sqrt((a[0] - b[0]) ** 2 + (a[1] - b[1]) ** 2 + (a[2] - b[2]) ** 2)
    den = np.sqrt(b[0] ** 2 + b[1] ** 2 + b[2] ** 2)
    return num / den


def export_tree(be, n):
    """cm_test_tree (csrc/cm_testing.h): keys, order and node arrays of the
    tree the library builds for its current state."""
    L = be._L
    fn = L.cm_test_tree
    fn.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                   C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_void_p, C.c_void_p,
                   C.c_void_p]
    fn.restype = C.c_int32
    cap = 3 * n + 1024
    keys = np.zeros(n, np.uint64)
    order = np.zeros(n, np.int32)
    box = np.zeros(4)
    info = np.zeros((cap, 5), np.int32)
    cm = np.zeros((cap, 4))
    ls = np.zeros(24, np.int32)
    na, nn = C.c_int32(), C.c_int32()
    be._check(fn(be._h, cap, keys.ctypes.data, order.ctypes.data, box.ctypes.data,
                 C.byref(na), C.byref(nn), info.ctypes.data, cm.ctypes.data, ls.ctypes.data),
              "cm_test_tree")
    nn = nn.value
    return {"keys": keys, "order": order, "box": box, "n_active": na.value,
            "n_nodes": nn, "info": info[:nn], "cm": cm[:nn], "level_start": ls}


@pytest.mark.parametrize("inactive", [0.0, 0.05])
def test_tree_bit_exact_at_bench_size(inactive):
    """Keys, sort permutation and every node field equal the oracle's at
    N = 2^20 (30-bit keys, leaves of 8)."""
    print(f"tree export: n={N} inactive={inactive}", flush=True)
    soa = _bench_state(inactive)
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(N, precision=cmhip.PREC_FP64, theta=0.5, morton_levels=10,
                                 leaf_capacity=LEAF, group_size=64)
    be.upload(soa)
    g = export_tree(be, N)
    ref = O.tree_build(s, 10, LEAF)
    na = ref["n_active"]
    assert g["n_active"] == na == int(s["active"].sum())
    np.testing.assert_array_equal(g["box"], ref["box"])
    # active keys bit for bit (inactive bodies carry each side's own sentinel
    # and sort last, stably, in both)
    np.testing.assert_array_equal(g["keys"][:na], ref["keys"][:na])
    np.testing.assert_array_equal(g["order"], ref["order"])
    assert np.all(np.diff(g["keys"][:na].astype(np.uint64)) >= 0)
    assert g["n_nodes"] == ref["n_nodes"]
    info = g["info"]
    np.testing.assert_array_equal(info[:, 0], ref["first"])
    np.testing.assert_array_equal(info[:, 1], ref["nchild"])
    np.testing.assert_array_equal(info[:, 2], ref["level"])
    np.testing.assert_array_equal(info[:, 3], ref["start"])
    np.testing.assert_array_equal(info[:, 4], ref["count"])
    np.testing.assert_array_equal(g["cm"][:, 3], ref["mass"])
    np.testing.assert_array_equal(g["cm"][:, :3], ref["com"])
    nl = ref["level"].max() + 3
    np.testing.assert_array_equal(g["level_start"][:nl], ref["level_start"][:nl])
    be.close()


def test_barnes_hut_fp64_bit_exact_at_bench_size():
    """One FP64 Barnes-Hut evaluation (theta 0.5, leaves of 8) equals the
    oracle's group traversal bit for bit at N = 2^20, with equal counts."""
    print(f"fp64 barnes-hut: n={N}", flush=True)
    soa = _bench_state()
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(N, precision=cmhip.PREC_FP64, theta=0.5, morton_levels=10,
                                 leaf_capacity=LEAF, group_size=64)
    be.compute_forces(soa, EPS)
    gs = be.stats()
    ref, st = O.bh_tree(s, EPS, 0.5, levels=10, leaf_cap=LEAF, group_size=64,
                        split_depth=-3, lifo=1536, nthreads=0)
    for a, b in zip(_acc(soa), ref):
        np.testing.assert_array_equal(a, b)
    assert gs["flags"] == 0 and gs["n_active"] == N
    assert gs["cell_interactions"] == int(st["n_cells"].astype(np.int64).sum())
    assert gs["body_interactions"] == int(st["n_parts"].astype(np.int64).sum())
    be.close()


@pytest.mark.parametrize("kappa", [0.25, 0.15])
def test_fast_path_matches_oracle_at_bench_size(kappa):
    """The benchmark's fast path (FP32 split, far field kappa = 0.25, theta
    0.5, leaves of 8) against the oracle's FP64 restatement of the same
    lists: median <= 2e-6, max <= 1e-3, equal term counts."""
    print(f"fp32 split + far field {kappa}: n={N}", flush=True)
    soa = _bench_state()
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(N, precision=cmhip.PREC_FP32_SPLIT, theta=0.5,
                                 morton_levels=10, leaf_capacity=LEAF, group_size=64,
                                 far_field_ratio=kappa)
    be.compute_forces(soa, EPS)
    gs = be.stats()
    ref, st = O.bh_tree(s, EPS, 0.5, levels=10, leaf_cap=LEAF, group_size=64,
                        split_depth=-3, lifo=1536, nthreads=0, far_ratio=kappa)
    assert gs["far_cells"] == st["far_cells"] > 0
    assert gs["cell_interactions"] == int(st["n_cells"].astype(np.int64).sum())
    assert gs["body_interactions"] == int(st["n_parts"].astype(np.int64).sum())
    r = _rel(_acc(soa), ref)
    print(f"median {np.median(r):.3e} max {r.max():.3e}", flush=True)
    assert np.median(r) <= 2e-6, np.median(r)
    assert r.max() <= 1e-3, r.max()
    be.close()


def test_group_offsets_from_the_scan_kernel_above_4m_bodies():
    """Above 4096 group tiles (4,194,304 bodies) the target groups' offsets
    come from one scan kernel instead of each block's own sum over the tiles
    before it.  At 4,198,400 bodies (4100 tiles): every active body is in a
    group (non-zero acceleration, every body counted), and the FP32 fast path
    is within the benchmark's error budget against the FP64 direct sum on a
    4096-body sample (median <= 1e-3)."""
    n = 4096 * 1024 + 4096
    print(f"group scan path: n={n}", flush=True)
    soa = cmhip.generate_plummer(n, 1.0, 1.0, seed=7, mass_fraction_cut=0.999, com_frame=True)
    be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP32_SPLIT, theta=0.5,
                                 morton_levels=10, leaf_capacity=LEAF, group_size=64,
                                 far_field_ratio=0.25)
    be.compute_forces(soa, EPS)
    gs = be.stats()
    assert gs["flags"] == 0 and gs["n_active"] == n
    assert 0 < gs["n_groups"] <= n
    acc = _acc(soa)
    mag = np.sqrt(acc[0] ** 2 + acc[1] ** 2 + acc[2] ** 2)
    assert (mag > 0).all()
    targets = np.sort(np.random.default_rng(4096).choice(n, 4096, replace=False))
    ex = be.direct_sample(EPS, targets)
    r = _rel(tuple(a[targets] for a in acc), ex)
    print(f"groups {gs['n_groups']} median {np.median(r):.3e}", flush=True)
    assert np.median(r) <= 1e-3
    be.close()
