"""GPU parity of the group far-field expansion (Barnes-Hut FP32 split with
far_field_ratio > 0) against the oracle's FP64 restatement of it."""
import numpy as np
import pytest

import cmhip
import oracle as O
from conftest import soa_to_state

pytestmark = pytest.mark.gpu

EPS = 1e-4


def _acc(soa):
    n = soa.count
    return soa.acc_x[:n].copy(), soa.acc_y[:n].copy(), soa.acc_z[:n].copy()


def _rel(a, b):
    num = np.sqrt((a[0] - b[0]) ** 2 + (a[1] - b[1]) ** 2 + (a[2] - b[2]) ** 2)
    den = np.sqrt(b[0] ** 2 + b[1] ** 2 + b[2] ** 2)
    return num / den


@pytest.mark.parametrize("group,leaf,kappa", [(64, 8, 0.15), (16, 4, 0.15), (64, 16, 0.3),
                                              (1, 1, 0.1)])
def test_far_field_matches_oracle(plummer, group, leaf, kappa):
    """Same tree, lists and far-cell split as the oracle (the classification is
    FP64 in both, so the counts are equal); accelerations within the FP32
    split tolerance: per-body relative difference median <= 2e-6, max <= 1e-3."""
    n = 40000
    soa = _with_inactive(plummer(n, cut=0.999))
    s = soa_to_state(soa)
    print(f"far field: n={n} group={group} leaf={leaf} kappa={kappa}", flush=True)
    be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP32_SPLIT, theta=0.5,
                                 leaf_capacity=leaf, group_size=group, far_field_ratio=kappa)
    be.compute_forces(soa, EPS)
    st = be.stats()
    ref, rst = O.bh_tree(s, EPS, 0.5, levels=10, leaf_cap=leaf, group_size=group,
                         split_depth=-3, lifo=1536, nthreads=8, far_ratio=kappa)
    act = s["active"] != 0
    assert st["far_cells"] == rst["far_cells"] > 0
    assert st["cell_interactions"] == int(rst["n_cells"].astype(np.int64)[act].sum())
    assert st["body_interactions"] == int(rst["n_parts"].astype(np.int64)[act].sum())
    a = _acc(soa)
    r = _rel([x[act] for x in a], [x[act] for x in ref])
    assert np.median(r) <= 2e-6, np.median(r)
    assert r.max() <= 1e-3, r.max()
    for x in a:   # inactive bodies get 0
        assert np.all(x[~act] == 0.0)


def _with_inactive(soa, frac=0.05, seed=5):
    rng = np.random.default_rng(seed)
    n = soa.count
    soa.is_active[:n][rng.random(n) < frac] = 0
    return soa


def test_far_ratio_zero_is_plain_bit_for_bit(plummer):
    n = 30000
    soa = plummer(n, cut=0.999)
    outs = []
    for kw in ({}, {"far_field_ratio": 0.0}):
        be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP32_SPLIT, theta=0.5,
                                     leaf_capacity=8, **kw)
        be.compute_forces(soa, EPS)
        assert be.stats()["far_cells"] == 0
        outs.append(_acc(soa))
        be.close()
    for x, y in zip(*outs):
        np.testing.assert_array_equal(x, y)


def test_far_field_accuracy_against_direct_sum(plummer):
    """At 200k bodies, theta = 0.5, kappa = 0.15: median force error against
    the FP64 direct sum within 15 % of the plain traversal's, with at least
    a quarter of the per-body cell terms moved into group expansions."""
    n = 200000
    soa = plummer(n, cut=0.999)
    sample = np.sort(np.random.default_rng(1).choice(n, 2048, replace=False)).astype(np.int32)
    res = {}
    for kappa in (0.0, 0.15):
        be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP32_SPLIT, theta=0.5,
                                     leaf_capacity=8, far_field_ratio=kappa)
        be.upload(soa)
        be.compute_forces_resident(EPS)
        out = cmhip.BodySoA(n)
        be.download(out, fields=("acc_x", "acc_y", "acc_z"))
        ex = be.direct_sample(EPS, sample)
        a = (out.acc_x[sample], out.acc_y[sample], out.acc_z[sample])
        res[kappa] = (float(np.median(_rel(a, ex))), be.stats())
        be.close()
    e0, s0 = res[0.0]
    e1, s1 = res[0.15]
    assert e1 <= 1.15 * e0, (e0, e1)
    assert s1["cell_interactions"] < 0.75 * s0["cell_interactions"]
    assert s1["body_interactions"] == s0["body_interactions"]


def test_far_field_ratio_validation():
    be = cmhip.HipPhysicsBackend(16, precision=cmhip.PREC_FP32_SPLIT)
    for bad in (-0.1, 0.6, float("nan")):
        with pytest.raises(cmhip.CmError):
            be.configure(far_field_ratio=bad)
    be.configure(far_field_ratio=0.5)


def _state_of(soa):
    return soa_to_state(soa)


@pytest.mark.parametrize("case", ["n1", "n2", "n3", "coincident", "inactive", "empty",
                                  "theta_clamp_max_kappa", "no_softening", "clusters"])
def test_far_field_edge_cases(case):
    """Small, degenerate and clustered inputs through the far-field path,
    against the oracle's FP64 restatement (same counts, relative difference
    median <= 2e-6, max <= 1e-3; inactive bodies get exactly 0)."""
    rng = np.random.default_rng(17)
    theta, kappa, eps, group, leaf = 0.5, 0.15, EPS, 64, 8
    if case == "n1":
        p = np.zeros((1, 3)); m = np.ones(1)
    elif case == "n2":
        p = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]]); m = np.array([1.0, 1e-3])
    elif case == "n3":
        p = rng.normal(0, 1, (3, 3)); m = rng.uniform(0.1, 1, 3)
    elif case == "coincident":
        # many bodies at identical positions: zero-extent groups and leaves
        p = np.repeat(rng.normal(0, 1, (40, 3)), 50, axis=0); m = rng.uniform(0.1, 1, 2000)
    elif case == "inactive" or case == "empty":
        p = rng.normal(0, 1, (3000, 3)); m = rng.uniform(0.1, 1, 3000)
    elif case == "theta_clamp_max_kappa":
        p = rng.normal(0, 1, (5000, 3)); m = rng.uniform(0.1, 1, 5000)
        theta, kappa = 0.0, 0.5
    elif case == "no_softening":
        p = rng.normal(0, 1, (5000, 3)); m = rng.uniform(0.1, 1, 5000)
        eps, group = 0.0, 16
    else:   # a few dense clusters far apart: far cells at every level
        c = rng.normal(0, 100, (8, 3))
        p = np.concatenate([c[k] + rng.normal(0, 0.01, (600, 3)) for k in range(8)])
        m = rng.uniform(0.1, 1, len(p))
    n = len(p)
    act = np.ones(n, np.uint8)
    if case == "inactive":
        act[rng.random(n) < 0.5] = 0
    if case == "empty":
        act[:] = 0
    soa = cmhip.BodySoA.from_arrays(pos_x=p[:, 0], pos_y=p[:, 1], pos_z=p[:, 2], mass=m,
                                    acc_x=np.full(n, 7.0), is_active=act)
    s = _state_of(soa)
    print(f"far-field edge case {case}: n={n}", flush=True)
    be = cmhip.HipPhysicsBackend(max(n, 16), precision=cmhip.PREC_FP32_SPLIT, theta=theta,
                                 leaf_capacity=leaf, group_size=group, far_field_ratio=kappa)
    be.compute_forces(soa, eps)
    st = be.stats()
    a = _acc(soa)
    live = act != 0
    for x in a:
        assert np.all(x[~live] == 0.0)
    if not live.any():
        return
    ref, rst = O.bh_tree(s, eps, theta, levels=10, leaf_cap=leaf, group_size=group,
                         split_depth=-3, lifo=1536, nthreads=8, far_ratio=kappa)
    assert st["far_cells"] == rst["far_cells"]
    assert st["cell_interactions"] == int(rst["n_cells"].astype(np.int64)[live].sum())
    assert st["body_interactions"] == int(rst["n_parts"].astype(np.int64)[live].sum())
    ra = [x[live] for x in a]
    rr = [x[live] for x in ref]
    den = np.sqrt(rr[0] ** 2 + rr[1] ** 2 + rr[2] ** 2)
    num = np.sqrt((ra[0] - rr[0]) ** 2 + (ra[1] - rr[1]) ** 2 + (ra[2] - rr[2]) ** 2)
    nz = den > 0
    if case == "n1":
        assert np.all(num == 0.0) and np.all(den == 0.0)
        return
    r = num[nz] / den[nz]
    assert np.all(num[~nz] == 0.0)
    assert np.median(r) <= 2e-6, np.median(r)
    assert r.max() <= 1e-3, r.max()
