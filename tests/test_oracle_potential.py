"""Pins the oracle's Barnes-Hut potential energy (the group traversal's
restatement of the engine's compute_potential_bh_kernel,
engine/src/cuda/potential_energy_kernel.cu:29-111) before it judges the GPU:
closed forms, the reference's per-body traversal in the limit where the two
walks coincide, convergence to the exact pair sum, inactive bodies."""
import math

import numpy as np

import oracle as O
from test_oracle import random_state, state


def test_two_bodies_closed_form():
    """GravityTests.cs:79-90: unit masses 2 apart, no softening, PE = -0.5;
    with softening, -m1 m2 / sqrt(d^2 + eps^2)."""
    s = state([0.0, 2.0], [0.0, 0.0], [0.0, 0.0], [1.0, 1.0])
    rows, _ = O.bh_potential(s, 0.0, 0.5)
    assert rows.sum() == -0.5
    np.testing.assert_array_equal(rows, [-0.25, -0.25])
    s = state([0.0, 3.0], [0.0, 0.0], [0.0, 0.0], [2.0, 0.5])
    rows, _ = O.bh_potential(s, 0.1, 0.5)
    want = -2.0 * 0.5 / math.sqrt(9.0 + 0.01)
    assert abs(rows.sum() - want) <= 1e-15 * abs(want)
    ref = O.bh_potential_reference(s, 0.1, 0.5)
    assert abs(ref.sum() - want) <= 1e-15 * abs(want)


def test_group_potential_reduces_to_reference_kernel():
    """With one body per group and per leaf, 63-bit keys and well-separated
    bodies the group walk is the reference kernel's per-body walk on the same
    geometric octree: per-body shares agree with its restatement."""
    rng = np.random.default_rng(12)
    g = rng.permutation(np.array(np.meshgrid(*(np.arange(12),) * 3)).reshape(3, -1).T)[:800]
    p = g + rng.uniform(0.2, 0.8, g.shape)
    s = state(p[:, 0], p[:, 1], p[:, 2], rng.uniform(0.5, 2, len(p)))
    ref = O.bh_potential_reference(s, 1e-3, 0.5)
    got, _ = O.bh_potential(s, 1e-3, 0.5, levels=21, leaf_cap=1, group_size=1, split_depth=0)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0)


def test_potential_converges_to_the_exact_pair_sum():
    s = random_state(3000, seed=4)
    ex = O.energy_exact(s, 1e-3, True, nthreads=8)["pe"]
    errs = []
    for theta in (0.9, 0.5, 0.2):
        rows, _ = O.bh_potential(s, 1e-3, theta, nthreads=8)
        errs.append(abs(math.fsum(rows) - ex) / abs(ex))
    assert errs[0] > errs[1] > errs[2], errs
    assert errs[1] < 1e-4 and errs[2] < 1e-6, errs
    ref = O.bh_potential_reference(s, 1e-3, 0.5, nthreads=8)
    assert abs(math.fsum(ref) - ex) / abs(ex) < 1e-4


def test_inactive_bodies_neither_feel_nor_source_the_potential():
    s = random_state(2000, seed=8)
    s["active"][::3] = 0
    rows, _ = O.bh_potential(s, 1e-3, 0.5, nthreads=8)
    assert np.all(rows[s["active"] == 0] == 0.0)
    live = {k: (v[s["active"] != 0] if isinstance(v, np.ndarray) and len(v) == 2000 else v)
            for k, v in s.items()}
    rows2, _ = O.bh_potential(live, 1e-3, 0.5, nthreads=8)
    assert abs(math.fsum(rows) - math.fsum(rows2)) <= 1e-6 * abs(math.fsum(rows2))
