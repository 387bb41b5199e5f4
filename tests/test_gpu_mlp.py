"""Learned far-field correction on the GPU: the MFMA forward and backward
kernels against the CPU oracle (same bf16 rounding points), the collected
inputs / targets against FP64 sums over the oracle's tree, training, and the
ablation: the trained network lowers the Barnes-Hut force error at equal
theta (plain Barnes-Hut is the method without it)."""
import numpy as np
import pytest

import cmhip
import mlp_oracle as MO
import oracle as O

pytestmark = pytest.mark.gpu

EPS = 1e-4


def _backend(n, theta, learned=False):
    return cmhip.HipPhysicsBackend(n, kind=cmhip.BACKEND_BARNES_HUT,
                                   precision=cmhip.PREC_FP32_SPLIT, theta=theta,
                                   learned_correction=learned)


def test_forward_matches_oracle():
    rng = np.random.default_rng(5)
    be = _backend(64, 0.5)
    p = (rng.normal(size=MO.N_PARAMS) * 0.3).astype(np.float32)
    be.mlp_set_params(p)
    np.testing.assert_array_equal(be.mlp_params(), p)
    x = rng.normal(size=(200, MO.IN)).astype(np.float32)
    y = be.mlp_forward(x)
    yo, _ = MO.forward(p, x, emulate=True)
    err = np.abs(y - yo)
    scale = np.abs(yo).max()
    # fp32 vs fp64 accumulation can move a hidden value across a bf16
    # rounding boundary: allow a few bf16 ulps on the outputs, median tight
    assert err.max() <= 2e-2 * scale
    assert np.median(err) <= 1e-4 * scale


def test_gradient_matches_oracle():
    rng = np.random.default_rng(6)
    be = _backend(64, 0.5)
    p = (rng.normal(size=MO.N_PARAMS) * 0.3).astype(np.float32)
    be.mlp_set_params(p)
    x = rng.normal(size=(300, MO.IN)).astype(np.float32)
    t = rng.normal(size=(300, MO.USED)).astype(np.float32)
    g, L = be.mlp_gradient(x, t)
    go, Lo = MO.gradient(p, x, t, emulate=True)
    assert abs(L - Lo) <= 1e-4 * Lo
    assert np.linalg.norm(g - go) <= 1e-2 * np.linalg.norm(go)
    # every layer's block agrees, not just the total
    k = 0
    for o, i in MO.SHAPES:
        for blk in (slice(k, k + o * i), slice(k + o * i, k + o * i + o)):
            gb, gob = g[blk], go[blk]
            assert np.linalg.norm(gb - gob) <= 2e-2 * np.linalg.norm(gob) + 1e-7
        k += o * i + o


def test_collected_pairs_match_fp64_sums():
    soa = cmhip.generate_merger(8192, seed=3)
    be = _backend(soa.count, 0.7)
    be.upload(soa)
    be.mlp_init(1)
    n = be.mlp_collect(EPS, max_groups=24, seed=5)
    assert n > 100
    x, t, sn, centre = be.mlp_pairs(n)
    # the tree of the oracle is the GPU's (bit-exact build, same numbering)
    st = {"px": soa.pos_x[:soa.count], "py": soa.pos_y[:soa.count],
          "pz": soa.pos_z[:soa.count], "mass": soa.mass[:soa.count],
          "active": soa.is_active[:soa.count]}
    tr = O.tree_build(st, 10, 16)
    order = tr["order"]
    bx, by, bz, bm = (a[order] for a in (st["px"], st["py"], st["pz"], st["mass"]))
    h = tr["box"][3]
    rng = np.random.default_rng(0)
    for k in rng.choice(n, 150, replace=False):
        node = sn[k, 1]
        s0, c = tr["start"][node], tr["count"][node]
        sl = slice(s0, s0 + c)
        assert c >= 2
        com = tr["com"][node]
        M = tr["mass"][node]
        hl = np.ldexp(h, -int(tr["level"][node]))
        Q = MO.second_moments(bx[sl], by[sl], bz[sl], bm[sl])
        xo, to = MO.pair_inputs_targets(bx[sl], by[sl], bz[sl], bm[sl], com, M, hl, Q,
                                        centre[k], EPS * EPS)
        np.testing.assert_allclose(x[k], xo, rtol=1e-4, atol=1e-5 * np.abs(xo).max())
        np.testing.assert_allclose(t[k], to, rtol=1e-3, atol=1e-4 * np.abs(to).max() + 1e-6)
    # ordered by sampled group, then node
    key = sn[:, 0].astype(np.int64) * (1 << 31) + sn[:, 1]
    assert np.all(np.diff(key) > 0)
    # deterministic: the same sample again gives the same pairs
    n2 = be.mlp_collect(EPS, max_groups=24, seed=5)
    x2, t2, sn2, _ = be.mlp_pairs(n2)
    assert n2 == n
    np.testing.assert_array_equal(sn2, sn)
    np.testing.assert_array_equal(t2, t)


def _forces(be, soa):
    be.upload(soa)
    be.compute_forces_resident(EPS)
    out = cmhip.BodySoA(soa.count)
    be.download(out, fields=("acc_x", "acc_y", "acc_z"))
    return np.stack([out.acc_x[:soa.count], out.acc_y[:soa.count], out.acc_z[:soa.count]], 1)


def test_learned_traversal_matches_reconstruction():
    # the traversal's correction equals sum over its (group, cell) pairs of
    # residual + gradient . (x - centre) rebuilt on the CPU from the oracle
    # network's outputs for the same inputs
    soa = cmhip.generate_merger(8192, seed=3)
    be = _backend(soa.count, 0.7)
    be.mlp_init(2)
    plain = _forces(be, soa)
    n = be.mlp_collect(EPS, max_groups=1 << 20, seed=1)
    x, _, info, centre = be.mlp_pairs(n)
    be.configure(learned_correction=1)
    learned = _forces(be, soa)
    y, _ = MO.forward(be.mlp_params(), x, emulate=True)
    st = {"px": soa.pos_x[:soa.count], "py": soa.pos_y[:soa.count],
          "pz": soa.pos_z[:soa.count], "mass": soa.mass[:soa.count],
          "active": soa.is_active[:soa.count]}
    tr = O.tree_build(st, 10, 16)
    order = tr["order"]
    P = np.stack([st["px"][order], st["py"][order], st["pz"][order]], 1)
    corr = np.zeros_like(P)
    for k in range(n):
        node = info[k, 1]
        hl = np.ldexp(tr["box"][3], -int(tr["level"][node]))
        R, J = MO.correction_from_outputs(y[k], tr["com"][node], tr["mass"][node], hl,
                                          centre[k])
        s, e = info[k, 2], info[k, 3]
        corr[s:e] += R + (P[s:e] - centre[k]) @ J.T
    expect = np.zeros_like(corr)
    expect[order] = corr
    got = learned - plain
    scale = np.linalg.norm(expect, axis=1)
    rel = np.linalg.norm(got - expect, axis=1) / np.maximum(scale, 1e-300)
    assert np.median(rel) < 1e-2
    assert np.linalg.norm(got - expect) < 2e-2 * np.linalg.norm(expect)


def _median_error(be, soa, targets, exact):
    be.upload(soa)
    be.compute_forces_resident(EPS)
    out = cmhip.BodySoA(soa.count)
    be.download(out, fields=("acc_x", "acc_y", "acc_z"))
    err = O.force_error(exact, (out.acc_x[targets], out.acc_y[targets], out.acc_z[targets]))
    return err["median"]


def test_training_and_ablation():
    soa = cmhip.generate_merger(65536, seed=11)
    theta = 0.7
    be = _backend(soa.count, theta)
    be.upload(soa)
    rng = np.random.default_rng(9)
    targets = np.sort(rng.choice(soa.count, 2048, replace=False)).astype(np.int32)
    exact = be.direct_sample(EPS, targets)
    plain = _median_error(be, soa, targets, exact)

    be.mlp_init(3)
    n = be.mlp_collect(EPS, max_groups=512, seed=21)
    assert n > 10000
    _, t, _, _ = be.mlp_pairs(n)
    losses = be.mlp_train(epochs=30, batch=1024, lr=2e-3, seed=4)
    assert np.all(np.isfinite(losses))
    # the network explains more than 90% of the targets' mean square
    assert losses[-1] < 0.1 * np.mean(t.astype(np.float64) ** 2)

    be.configure(learned_correction=1)
    learned = _median_error(be, soa, targets, exact)
    print(f"theta {theta}: median rel force error plain {plain:.3e}, "
          f"with MLP {learned:.3e} (x{plain / learned:.2f}); loss {losses[0]:.3e} -> "
          f"{losses[-1]:.3e}")
    assert learned < plain / 1.3


def test_learned_needs_network_and_fp32():
    be = _backend(64, 0.5)
    soa = cmhip.generate_plummer(64, seed=1)
    be.upload(soa)
    be.configure(learned_correction=1)
    with pytest.raises(cmhip.CmError):
        be.compute_forces_resident(EPS)
    be.mlp_init(1)
    be.compute_forces_resident(EPS)
    be.configure(precision=cmhip.PREC_FP64)
    with pytest.raises(cmhip.CmError):
        be.compute_forces_resident(EPS)


@pytest.mark.parametrize("theta", [0.5, 0.9])
def test_ablation_at_other_thetas(theta):
    """The learned-correction configuration's other opening angles on the
    65,536-body merger: the network is trained the same way at each theta and
    the median error with it is reported beside the plain traversal's.  At
    theta 0.5 the monopole residual of accepted cells is small and the
    correction is bounded by the group-level expansion; at 0.9 it is large.
    Measured: 4.06e-4 -> 2.79e-4 at 0.5 (1.45x), 2.26e-3 -> 1.85e-3 at 0.9
    (1.22x); the check asks for a 1.1x gain at each."""
    soa = cmhip.generate_merger(65536, seed=11)
    be = _backend(soa.count, theta)
    be.upload(soa)
    rng = np.random.default_rng(9)
    targets = np.sort(rng.choice(soa.count, 2048, replace=False)).astype(np.int32)
    exact = be.direct_sample(EPS, targets)
    plain = _median_error(be, soa, targets, exact)
    be.mlp_init(3)
    be.mlp_collect(EPS, max_groups=512, seed=21)
    losses = be.mlp_train(epochs=30, batch=1024, lr=2e-3, seed=4)
    assert np.all(np.isfinite(losses))
    be.configure(learned_correction=1)
    learned = _median_error(be, soa, targets, exact)
    print(f"theta {theta}: median rel force error plain {plain:.3e}, with MLP {learned:.3e} "
          f"(x{plain / learned:.2f})", flush=True)
    assert learned < plain / 1.1
