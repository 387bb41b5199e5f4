"""GPU parity: libcmhip.so against the CPU oracle on the same seeded inputs.

FP64 paths are bit-exact (same operation order, no FMA, IEEE divide and
square root); FP32 split-precision is judged by relative error with the
tolerance written in each test."""
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


def _eq(a, b):
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def _rel(a, b):
    num = np.sqrt((a[0] - b[0]) ** 2 + (a[1] - b[1]) ** 2 + (a[2] - b[2]) ** 2)
    den = np.sqrt(b[0] ** 2 + b[1] ** 2 + b[2] ** 2)
    return num / den


def _with_inactive(soa, frac=0.1, seed=5):
    rng = np.random.default_rng(seed)
    n = soa.count
    soa.is_active[:n][rng.random(n) < frac] = 0
    return soa


@pytest.mark.parametrize("n", [1, 2, 257, 1000])
def test_direct_pairwise_bit_exact(plummer, n):
    """CpuSingleThreadBackend Newton-3 order (CpuSingleThreadBackend.cs:79-118)."""
    soa = _with_inactive(plummer(n))
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(n, kind=cmhip.BACKEND_DIRECT_PAIRWISE)
    be.compute_forces(soa, EPS)
    _eq(_acc(soa), O.direct_n3(s, EPS))


@pytest.mark.parametrize("shell", [False, True])
def test_direct_sweep_bit_exact(plummer, shell):
    """CpuParallelBackend sweep / shell-theorem GravityKernel (GravityKernel.cs:9-33)."""
    soa = _with_inactive(plummer(1500))
    soa.radius[:soa.count] = 0.05     # large radii so the interior branch is hit
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(1500, kind=cmhip.BACKEND_DIRECT, enable_shell_theorem=shell)
    be.compute_forces(soa, EPS)
    _eq(_acc(soa), O.direct_gk(s, EPS, shell))


def test_verlet_steps_bit_exact(plummer):
    """SoAVerletIntegrator.Step x 20 with the pairwise backend: every array bit-exact."""
    n = 1000
    soa = _with_inactive(plummer(n))
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(n, kind=cmhip.BACKEND_DIRECT_PAIRWISE)
    be.upload(soa)
    be.step(EPS, 1e-3, 20)
    out = soa.copy()
    be.download(out)
    for _ in range(20):
        O.verlet_step(s, 1e-3, lambda st: O.direct_n3(st, EPS))
    for k_soa, k_or in (("pos_x", "px"), ("pos_y", "py"), ("pos_z", "pz"), ("vel_x", "vx"),
                        ("vel_y", "vy"), ("vel_z", "vz"), ("acc_x", "ax"),
                        ("old_acc_z", "oaz")):
        np.testing.assert_array_equal(getattr(out, k_soa)[:n], s[k_or])


@pytest.mark.parametrize("levels,leaf,gsize", [(10, 16, 64), (21, 8, 64), (10, 1, 32), (21, 64, 16)])
def test_barnes_hut_fp64_bit_exact(plummer, levels, leaf, gsize):
    """Build-defined octree + group-MAC traversal in FP64 equals the oracle bit for bit,
    including the interaction counts."""
    n = 20000
    soa = _with_inactive(plummer(n, cut=0.999))
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP64, theta=0.5, morton_levels=levels,
                                 leaf_capacity=leaf, group_size=gsize)
    be.compute_forces(soa, EPS)
    ref, st = O.bh_tree(s, EPS, 0.5, levels=levels, leaf_cap=leaf, group_size=gsize,
                        split_depth=-3, lifo=1536, nthreads=8)
    _eq(_acc(soa), ref)
    gs = be.stats()
    assert gs["n_active"] == int(s["active"].sum())
    assert gs["cell_interactions"] == int(st["n_cells"].astype(np.int64).sum())
    assert gs["body_interactions"] == int(st["n_parts"].astype(np.int64).sum())
    assert gs["flags"] == 0


def _with_radii(soa, seed=9):
    """Source radii large enough that many close pairs fall inside the source
    (GravityKernel's interior branch), zero for a tenth of the bodies."""
    rng = np.random.default_rng(seed)
    n = soa.count
    r = rng.uniform(0.0, 0.02, n)
    r[rng.random(n) < 0.1] = 0.0
    soa.radius[:n] = r
    return soa


def test_barnes_hut_fp64_shell_theorem_bit_exact(plummer):
    """Shell theorem on (BarnesHutBackend.cs:415-421): leaves are always opened and
    body terms use the source radius; FP64 equals the oracle bit for bit."""
    n = 20000
    soa = _with_radii(_with_inactive(plummer(n, cut=0.999)))
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP64, theta=0.5, leaf_capacity=8,
                                 enable_shell_theorem=True)
    be.compute_forces(soa, EPS)
    ref, st = O.bh_tree(s, EPS, 0.5, levels=10, leaf_cap=8, group_size=64, shell=True,
                        split_depth=-3, lifo=1536, nthreads=8)
    _eq(_acc(soa), ref)
    gs = be.stats()
    assert gs["cell_interactions"] == int(st["n_cells"].astype(np.int64).sum())
    assert gs["body_interactions"] == int(st["n_parts"].astype(np.int64).sum())
    # the interior branch is exercised: radii change the result
    plain, _ = O.bh_tree(s, EPS, 0.5, levels=10, leaf_cap=8, group_size=64, shell=False,
                         split_depth=-3, lifo=1536, nthreads=8)
    assert np.count_nonzero(ref[0] != plain[0]) > n // 10


def test_barnes_hut_fp32_split_shell_theorem(plummer):
    """FP32 split with the shell theorem (and the far field, which only takes
    accepted internal nodes) against the oracle: median <= 2e-6, max <= 1e-3."""
    n = 20000
    soa = _with_radii(plummer(n, cut=0.999))
    s = soa_to_state(soa)
    for far in (0.0, 0.15):
        work = soa.copy()
        be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP32_SPLIT, theta=0.5,
                                     leaf_capacity=8, enable_shell_theorem=True,
                                     far_field_ratio=far)
        be.compute_forces(work, EPS)
        ref, _ = O.bh_tree(s, EPS, 0.5, levels=10, leaf_cap=8, group_size=64, shell=True,
                           split_depth=-3, lifo=1536, nthreads=8, far_ratio=far)
        r = _rel(_acc(work), ref)
        assert np.median(r) <= 2e-6, (far, np.median(r))
        assert r.max() <= 1e-3, (far, r.max())
        be.close()


def test_barnes_hut_fp32_split_vs_fp64(plummer):
    """FP32 split precision on the same tree and lists as the FP64 path:
    per-body relative difference median <= 1e-6, max <= 1e-3."""
    n = 50000
    soa = plummer(n, cut=0.999)
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP32_SPLIT, theta=0.5)
    be.compute_forces(soa, EPS)
    ref, _ = O.bh_tree(s, EPS, 0.5, levels=10, leaf_cap=16, group_size=64, split_depth=-3,
                       lifo=1536, nthreads=8)
    r = _rel(_acc(soa), ref)
    assert np.median(r) <= 1e-6, np.median(r)
    assert r.max() <= 1e-3, r.max()


def test_barnes_hut_fp64_fast_vs_exact(plummer):
    """Hardware rsqrt + Newton (the FP64 baseline) within 1e-13 of the IEEE path."""
    n = 30000
    soa = plummer(n, cut=0.999)
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP64_FAST, theta=0.5)
    be.compute_forces(soa, EPS)
    ref, _ = O.bh_tree(s, EPS, 0.5, levels=10, leaf_cap=16, group_size=64, split_depth=-3,
                       lifo=1536, nthreads=8)
    assert _rel(_acc(soa), ref).max() <= 1e-13


def test_barnes_hut_two_body_equals_direct():
    """BarnesHutTests.cs:65-97: two bodies, BH equals brute force to 10 digits."""
    soa = cmhip.BodySoA.from_arrays(pos_x=[0.0, 1.0], pos_y=[0.0, 0.0], pos_z=[0.0, 0.0],
                                    mass=[1.0, 1e-3], is_active=[1, 1])
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(2)
    be.compute_forces(soa, EPS)
    ref = O.direct_gk(s, EPS)
    np.testing.assert_allclose(np.array(_acc(soa)), np.array(ref), rtol=1e-10, atol=0)


def test_theta_clamp_bit_identical(plummer):
    """BarnesHutAccuracyTests.cs:165-186: theta=0 equals theta=0.2 bit for bit."""
    soa = plummer(4000, cut=0.999)
    a = soa.copy()
    b = soa.copy()
    be = cmhip.HipPhysicsBackend(4000, theta=0.0)
    be.compute_forces(a, EPS)
    be.theta = 0.2
    be.compute_forces(b, EPS)
    _eq(_acc(a), _acc(b))


def test_barnes_hut_deterministic_rerun(plummer):
    """BarnesHutTests.cs:156-187: two runs are bit-identical (here: 5 Verlet steps, FP32)."""
    soa = plummer(30000, cut=0.999)
    outs = []
    for _ in range(2):
        be = cmhip.HipPhysicsBackend(30000, precision=cmhip.PREC_FP32_SPLIT, theta=0.6)
        be.upload(soa)
        be.step(EPS, 1e-3, 5)
        o = soa.copy()
        be.download(o)
        outs.append(o)
    for f in ("pos_x", "vel_y", "acc_z"):
        np.testing.assert_array_equal(getattr(outs[0], f), getattr(outs[1], f))


def test_all_inactive_and_empty():
    soa = cmhip.BodySoA.from_arrays(pos_x=[0.0, 1.0, 2.0], pos_y=[0.0] * 3, pos_z=[0.0] * 3,
                                    mass=[1.0] * 3, acc_x=[5.0] * 3, is_active=[0, 0, 0])
    be = cmhip.HipPhysicsBackend(8)
    be.compute_forces(soa, EPS)
    assert (soa.acc_x[:3] == 0).all()
    empty = cmhip.BodySoA(4)
    be.compute_forces(empty, EPS)


def test_energy_compensated_vs_exact(plummer):
    soa = _with_inactive(plummer(3000))
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(3000)
    be.upload(soa)
    e = be.energy(EPS, True)
    x = O.energy_exact(s, EPS, True, nthreads=8)
    assert e["active_count"] == int(s["active"].sum())
    assert abs(e["kinetic"] - x["ke"]) <= 1e-14 * abs(x["ke"])
    assert abs(e["potential"] - x["pe"]) <= 1e-13 * abs(x["pe"])
    for k, ref in enumerate((x["px"], x["py"], x["pz"])):
        assert abs(e["momentum"][k] - ref) <= 1e-15


def test_direct_sample_matches_oracle(plummer):
    soa = plummer(12000, cut=0.999)
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(12000)
    be.upload(soa)
    t = np.random.default_rng(3).choice(12000, 256, replace=False).astype(np.int32)
    _eq(be.direct_sample(EPS, t), O.direct_targets(s, EPS, t))


def test_barnes_hut_oversized_leaves_bit_exact():
    """30-bit keys with a cluster far below one key cell: a deepest-level leaf
    holds thousands of bodies and is split into group-sized slices."""
    rng = np.random.default_rng(23)
    core = rng.uniform(-1e-3, 1e-3, (3000, 3))
    far = rng.uniform(-100, 100, (10, 3))
    p = np.vstack([core, far])
    soa = cmhip.BodySoA.from_arrays(pos_x=p[:, 0], pos_y=p[:, 1], pos_z=p[:, 2],
                                    mass=rng.uniform(0.5, 2, len(p)), is_active=np.ones(len(p)))
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(len(p), precision=cmhip.PREC_FP64, theta=0.5)
    be.compute_forces(soa, EPS)
    ref, _ = O.bh_tree(s, EPS, 0.5, levels=10, leaf_cap=16, group_size=64)
    _eq(_acc(soa), ref)
    assert (np.abs(soa.acc_x[:len(p)]) > 0).all()


def test_pool_overflow_in_an_earlier_step_is_reported():
    """A run whose first force evaluation outgrows the node pool fails with
    CM_ERR_POOL_EXHAUSTED although its last evaluation fits (the state's node
    counts are pinned on the CPU by
    test_pair_collapse_state_overflows_only_after_the_first_drift); the next
    run, whose trees all fit, succeeds."""
    from conftest import pair_collapse_state
    soa = pair_collapse_state()
    n = soa.count
    print(f"sticky overflow: n={n}", flush=True)
    be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP64, theta=0.5, morton_levels=21,
                                 leaf_capacity=1, group_size=64)
    be.upload(soa)
    with pytest.raises(cmhip.CmError) as err:
        be.step(EPS, 1e-3, 2)
    assert err.value.code == cmhip.CM_ERR_POOL_EXHAUSTED
    be.step(EPS, 1e-3, 1)
    assert be.stats()["flags"] == 0


@pytest.mark.parametrize("eps", [1e-14, 0.0])
def test_fp32_split_tiny_softening_is_finite(plummer, eps):
    """Softening below 1e-6 (eps^2 < 1e-12) takes the self-skipping loop in
    FP32 split, where rsq(eps^2)^3 would overflow for the body's own term:
    the result stays finite and within the FP32 tolerance of the FP64 path."""
    n = 20000
    soa = plummer(n, cut=0.999)
    a = soa.copy()
    b = soa.copy()
    for work, prec in ((a, cmhip.PREC_FP64), (b, cmhip.PREC_FP32_SPLIT)):
        be = cmhip.HipPhysicsBackend(n, precision=prec, theta=0.5, leaf_capacity=8)
        be.compute_forces(work, eps)
        be.close()
    got, ref = _acc(b), _acc(a)
    assert all(np.isfinite(x).all() for x in got)
    r = _rel(got, ref)
    assert np.median(r) <= 2e-6 and r.max() <= 1e-3, (np.median(r), r.max())


def test_compare_forces_matches_oracle(plummer):
    """ForceErrorAnalyzer.CompareForces (ForceErrorAnalyzer.cs:67-135) on the GPU:
    exact FP64 sum vs FP64 Barnes-Hut, per-body relative error over active
    bodies with the 1e-30 skip; mean, max, RMS and median equal the oracle's
    statistics of the oracle's forces bit for bit (every body, and a sample
    in list order).  The Barnes-Hut forces are left in acc."""
    n = 20000
    soa = _with_inactive(plummer(n, cut=0.999))
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP64, theta=0.3, leaf_capacity=8)
    be.upload(soa)
    got = be.compare_forces(EPS, 0.6)
    exact = O.direct_gk(s, EPS)
    bh, _ = O.bh_tree(s, EPS, 0.6, levels=10, leaf_cap=8, group_size=64, nthreads=8)
    ref = O.force_error(exact, bh, s["active"])
    assert got["body_count"] == ref["count"] == int(s["active"].sum())
    for k in ("mean", "max", "rms", "median"):
        assert got[f"{k}_relative_error"] == ref[k], k
    assert got["theta"] == 0.6
    out = soa.copy()
    be.download(out, fields=("acc_x", "acc_y", "acc_z"))
    _eq(_acc(out), bh)
    # a sample in list order; the sample judge folds its sums over chunks of
    # 16,384 sources, which rounds differently from one sequential sweep
    t = np.random.default_rng(8).choice(n, 700, replace=False).astype(np.int32)
    got = be.compare_forces(EPS, 0.6, t)
    ex = O.direct_targets(s, EPS, t)
    ref = O.force_error(ex, tuple(b[t] for b in bh), s["active"][t])
    assert got["body_count"] == ref["count"]
    for k in ("mean", "max", "rms", "median"):
        assert abs(got[f"{k}_relative_error"] - ref[k]) <= 1e-12 * ref[k], k


def test_compare_forces_skips_zero_exact_force():
    """Bodies whose exact force is below 1e-30 are skipped (ForceErrorAnalyzer.cs:108):
    with one massive body among massless ones, that body feels nothing."""
    rng = np.random.default_rng(4)
    p = rng.normal(size=(500, 3))
    m = np.zeros(500)
    m[7] = 1.0
    soa = cmhip.BodySoA.from_arrays(pos_x=p[:, 0], pos_y=p[:, 1], pos_z=p[:, 2], mass=m,
                                    is_active=np.ones(500))
    be = cmhip.HipPhysicsBackend(500, precision=cmhip.PREC_FP32_SPLIT)
    be.upload(soa)
    got = be.compare_forces(EPS, 0.5)
    assert got["body_count"] == 499
    assert 0.0 <= got["median_relative_error"] <= got["max_relative_error"] < 1e-5
