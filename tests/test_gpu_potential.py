"""Barnes-Hut potential energy on the GPU (cm_energy with_potential = 2,
cm_potential_rows) against the oracle's FP64 restatement."""
import math

import numpy as np
import pytest

import cmhip
import oracle as O
from conftest import soa_to_state

pytestmark = pytest.mark.gpu

EPS = 1e-4


def _with_inactive(soa, frac=0.1, seed=5):
    rng = np.random.default_rng(seed)
    n = soa.count
    soa.is_active[:n][rng.random(n) < frac] = 0
    return soa


@pytest.mark.parametrize("levels,leaf,gsize,theta", [(10, 16, 64, 0.5), (21, 8, 64, 0.65),
                                                     (10, 1, 32, 0.5), (10, 8, 1, 0.2)])
def test_bh_potential_rows_bit_exact(plummer, levels, leaf, gsize, theta):
    """Per-body shares 1/2 m_i phi_i equal the oracle's bit for bit (FP64
    group traversal, same tree and lists); inactive bodies get 0."""
    n = 20000
    soa = _with_inactive(plummer(n, cut=0.999))
    s = soa_to_state(soa)
    print(f"potential rows: levels={levels} leaf={leaf} group={gsize} theta={theta}", flush=True)
    be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP32_SPLIT, theta=theta,
                                 morton_levels=levels, leaf_capacity=leaf, group_size=gsize,
                                 far_field_ratio=0.15)
    be.upload(soa)
    rows = be.potential_rows(EPS, 2)
    ref, _ = O.bh_potential(s, EPS, theta, levels=levels, leaf_cap=leaf, group_size=gsize,
                            split_depth=-3, lifo=1536, nthreads=8)
    np.testing.assert_array_equal(rows, ref)
    assert np.all(rows[s["active"] == 0] == 0.0)


def test_energy_with_bh_potential(plummer):
    """cm_energy(with_potential = 2): PE = sum of the shares (compensated sums,
    within 1e-13 of an exactly rounded sum of the oracle's shares), close to
    the exact PE (1e-4 at theta = 0.5), and the force statistics of the last
    evaluation are left alone."""
    n = 30000
    soa = plummer(n, cut=0.999)
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP32_SPLIT, theta=0.5, leaf_capacity=8)
    be.upload(soa)
    be.compute_forces_resident(EPS)
    before = be.stats()
    e_bh = be.energy(EPS, 2)
    after = be.stats()
    for k in ("cell_interactions", "body_interactions", "n_nodes", "n_groups",
              "acc_interactions"):
        assert before[k] == after[k], k
    ref, _ = O.bh_potential(s, EPS, 0.5, levels=10, leaf_cap=8, group_size=64, nthreads=8)
    want = math.fsum(ref)
    assert abs(e_bh["potential"] - want) <= 1e-13 * abs(want)
    e_ex = be.energy(EPS, 1)
    assert abs(e_bh["potential"] - e_ex["potential"]) <= 1e-4 * abs(e_ex["potential"])
    for k in ("kinetic", "total_mass", "active_count"):
        assert e_bh[k] == e_ex[k]
    exact_rows = be.potential_rows(EPS, 1)
    assert abs(math.fsum(exact_rows) - e_ex["potential"]) <= 1e-12 * abs(e_ex["potential"])


def test_potential_argument_checks_and_empty():
    be = cmhip.HipPhysicsBackend(16)
    with pytest.raises(cmhip.CmError):
        be.energy(EPS, 3)
    with pytest.raises(cmhip.CmError):
        be.potential_rows(EPS, 0)
    assert be.energy(EPS, 2)["potential"] == 0.0
    soa = cmhip.BodySoA.from_arrays(pos_x=[0.0, 2.0], pos_y=[0.0, 0.0], pos_z=[0.0, 0.0],
                                    mass=[1.0, 1.0], is_active=[1, 1])
    be.upload(soa)
    assert be.energy(0.0, 2)["potential"] == -0.5
    np.testing.assert_array_equal(be.potential_rows(0.0, 2), [-0.25, -0.25])
