"""GPU parity of the tree build at awkward sizes: body counts around the
64-lane wave, the 1024-body tree tile and the 2048-key sort tile, single
bodies, coincident bodies, 63-bit keys and leaves of one body.  Each case
compares, against the CPU oracle on the same input, the sorted keys, the
stable order and every node field (first child, child count, level, body
range, mass, centre of mass) bit for bit, and one FP64 Barnes-Hut evaluation
bit for bit.  These stages replace the reference engine's morton_code_kernel
and radix sort (engine/src/cuda/morton_kernel.cu:11-45,185-230) and its tree
build and bottom-up mass pass (engine/src/cuda/gpu_tree.cu:269-464)."""
import numpy as np
import pytest

import cmhip
import oracle as O
from conftest import soa_to_state
from test_gpu_scale import export_tree

pytestmark = pytest.mark.gpu

EPS = 1e-4


def _state(n, kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        p = rng.uniform(-1.0, 1.0, (n, 3))
    elif kind == "clustered":
        # a dense core far below one 30-bit key cell plus a sparse halo
        core = max(1, n // 2)
        p = np.vstack([rng.normal(0.0, 1e-7, (core, 3)), rng.uniform(-5.0, 5.0, (n - core, 3))])
    elif kind == "coincident":
        p = np.zeros((n, 3))
        p[: n // 3] = 0.25
    else:
        raise ValueError(kind)
    m = rng.uniform(0.5, 2.0, n)
    act = np.ones(n)
    if n > 4:
        act[rng.random(n) < 0.1] = 0
    return cmhip.BodySoA.from_arrays(pos_x=p[:, 0], pos_y=p[:, 1], pos_z=p[:, 2], mass=m,
                                     is_active=act)


CASES = [
    # (n, distribution, levels, leaf capacity)
    (1, "uniform", 10, 8),
    (2, "uniform", 10, 8),
    (63, "uniform", 10, 8),
    (65, "clustered", 10, 8),
    (1023, "uniform", 10, 8),
    (1025, "clustered", 10, 8),
    (2047, "uniform", 10, 1),
    (2049, "clustered", 10, 8),
    (4097, "coincident", 10, 8),
    (5000, "uniform", 21, 8),
    (5000, "clustered", 21, 1),
    (9000, "clustered", 7, 16),
]


@pytest.mark.parametrize("n,kind,levels,cap", CASES)
def test_tree_and_forces_bit_exact_at_awkward_sizes(n, kind, levels, cap):
    print(f"tree sizes: n={n} {kind} levels={levels} cap={cap}", flush=True)
    soa = _state(n, kind, seed=n + levels)
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP64, theta=0.5, morton_levels=levels,
                                 leaf_capacity=cap, group_size=64)
    be.upload(soa)
    g = export_tree(be, n)
    ref = O.tree_build(s, levels, cap)
    na = ref["n_active"]
    assert g["n_active"] == na
    np.testing.assert_array_equal(g["box"], ref["box"])
    np.testing.assert_array_equal(g["keys"][:na], ref["keys"][:na])
    np.testing.assert_array_equal(g["order"], ref["order"])
    assert g["n_nodes"] == ref["n_nodes"]
    info = g["info"]
    np.testing.assert_array_equal(info[:, 0], ref["first"])
    np.testing.assert_array_equal(info[:, 1], ref["nchild"])
    np.testing.assert_array_equal(info[:, 2], ref["level"])
    np.testing.assert_array_equal(info[:, 3], ref["start"])
    np.testing.assert_array_equal(info[:, 4], ref["count"])
    np.testing.assert_array_equal(g["cm"][:, 3], ref["mass"])
    np.testing.assert_array_equal(g["cm"][:, :3], ref["com"])
    # one FP64 evaluation on the same tree definition
    be.compute_forces(soa, EPS)
    acc, _ = O.bh_tree(s, EPS, 0.5, levels=levels, leaf_cap=cap, group_size=64)
    for a, b in zip((soa.acc_x[:n], soa.acc_y[:n], soa.acc_z[:n]), acc):
        np.testing.assert_array_equal(a, b)
    be.close()
