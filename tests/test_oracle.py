"""Pins the CPU oracle before it is trusted as the GPU's judge.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a sign, a transposed index, a wrong association) fails: closed forms
from the reference's own tests, brute force in extended precision, invariants,
and cross-checks between independent restatements (C# insertion tree vs the
build's level-order tree)."""
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def state(px, py, pz, m, vx=None, vy=None, vz=None, radius=None, active=None):
    n = len(px)
    z = np.zeros(n)
    s = {"px": np.array(px, float), "py": np.array(py, float), "pz": np.array(pz, float),
         "vx": np.array(vx if vx is not None else z, float),
         "vy": np.array(vy if vy is not None else z, float),
         "vz": np.array(vz if vz is not None else z, float),
         "ax": z.copy(), "ay": z.copy(), "az": z.copy(),
         "oax": z.copy(), "oay": z.copy(), "oaz": z.copy(),
         "mass": np.array(m, float),
         "radius": np.array(radius if radius is not None else z, float),
         "density": z.copy(),
         "active": np.array(active if active is not None else np.ones(n), np.uint8),
         "collidable": np.ones(n, np.uint8), "body_type": np.zeros(n, np.int32)}
    return s


def random_state(n, seed=0, spread=1.0):
    rng = np.random.default_rng(seed)
    return state(*(rng.normal(0, spread, n) for _ in range(3)), rng.uniform(0.5, 2.0, n),
                 *(rng.normal(0, 0.1, n) for _ in range(3)))


def brute_acc_ld(s, eps):
    """Direct sum in long double, straight from the formula a_i = -sum m_j r_ij / |r_ij|^3."""
    x = s["px"].astype(np.longdouble)
    y = s["py"].astype(np.longdouble)
    z = s["pz"].astype(np.longdouble)
    m = s["mass"].astype(np.longdouble)
    act = s["active"].astype(bool)
    n = len(x)
    out = np.zeros((3, n), dtype=np.longdouble)
    for i in range(n):
        if not act[i]:
            continue
        dx, dy, dz = x[i] - x, y[i] - y, z[i] - z
        r2 = dx * dx + dy * dy + dz * dz + np.longdouble(eps) ** 2
        w = m / r2 ** np.longdouble(1.5)
        w[i] = 0
        w[~act] = 0
        out[:, i] = [-(w * dx).sum(), -(w * dy).sum(), -(w * dz).sum()]
    return out


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.sqrt(((a - b) ** 2).sum(0)) / np.sqrt((b ** 2).sum(0))


# ---------------------------------------------------------------- gravity ---
def test_force_is_g_at_unit_distance():
    """GravityTests.cs:22-34: two unit masses 1 apart feel |a| = G = 1."""
    s = state([0, 1], [0, 0], [0, 0], [1, 1])
    ax, ay, az = O.direct_n3(s, 0.0)
    assert ax[0] == 1.0 and ax[1] == -1.0 and ay[0] == 0.0


def test_potential_at_distance_two():
    """GravityTests.cs:79-90: PE of two unit masses 2 apart is -0.5."""
    s = state([0, 2], [0, 0], [0, 0], [1, 1])
    assert O.energy_naive(s, 0.0)["pe"] == -0.5


def test_inverse_square_and_softening():
    s1 = state([0, 1], [0, 0], [0, 0], [1, 3])
    s2 = state([0, 2], [0, 0], [0, 0], [1, 3])
    a1 = O.direct_gk(s1, 0.0)[0][0]
    a2 = O.direct_gk(s2, 0.0)[0][0]
    assert a1 == 3.0 and a2 == 0.75
    e = 0.5
    a_soft = O.direct_gk(s1, e)[0][0]
    assert a_soft == pytest.approx(3.0 / (1 + e * e) ** 1.5, rel=1e-15)


def test_shell_theorem_interior_branch():
    """GravityKernel.cs:20-27: inside the source radius the field is linear."""
    R, e2 = 2.0, 0.01
    inner = O.gravity_coeff(1.0, 5.0, R, e2, True)
    assert inner == pytest.approx(5.0 / (R * R + e2) ** 1.5, rel=1e-15)
    outer = O.gravity_coeff(9.0, 5.0, R, e2, True)
    assert outer == pytest.approx(5.0 / (9.0 + e2) ** 1.5, rel=1e-15)
    assert O.gravity_coeff(1.0, 5.0, R, e2, False) == pytest.approx(5.0 / (1 + e2) ** 1.5,
                                                                    rel=1e-15)


@pytest.mark.parametrize("fn", ["n3", "gk"])
def test_direct_sums_match_extended_precision_brute_force(fn):
    s = random_state(40, seed=3)
    s["active"][[5, 17]] = 0
    ref = brute_acc_ld(s, 1e-3)
    got = O.direct_n3(s, 1e-3) if fn == "n3" else O.direct_gk(s, 1e-3)
    r = rel(got, ref)
    act = s["active"].astype(bool)
    assert r[act].max() < 1e-13
    assert all((np.asarray(g)[~act] == 0).all() for g in got)


def test_direct_targets_is_the_sampled_direct_sum():
    """The force judge (ForceErrorAnalyzer's exact forces, ForceErrorAnalyzer.cs:
    67-92, on a sample of targets): its rows equal the pinned full sweep
    (CpuParallelBackend order) at the sampled indices bit for bit, with
    inactive targets and sources, repeated and unsorted targets; and they
    agree with the long-double brute force.  A dropped source, a wrong index
    mapping or a self term fails one of these."""
    s = random_state(300, seed=11)
    s["active"][[4, 50, 51, 299]] = 0
    full = O.direct_gk(s, 1e-3)
    idx = np.array([299, 0, 7, 50, 7, 123, 4, 298, 1], dtype=np.int32)
    got = O.direct_targets(s, 1e-3, idx, nthreads=3)
    for g, f in zip(got, full):
        np.testing.assert_array_equal(g, f[idx])
    ref = brute_acc_ld(s, 1e-3)[:, idx]
    act = s["active"][idx].astype(bool)
    assert rel(np.array(got)[:, act], ref[:, act]).max() < 1e-13
    assert all((np.asarray(g)[~act] == 0).all() for g in got)
    # the Plummer case the benchmarks judge with (softening 1e-4, 2000 bodies)
    import cmhip
    from conftest import soa_to_state
    p = soa_to_state(cmhip.generate_plummer(2000, 1.0, 1.0, seed=5, mass_fraction_cut=0.999))
    t = np.random.default_rng(4).choice(2000, 64, replace=False).astype(np.int32)
    for g, f in zip(O.direct_targets(p, 1e-4, t), O.direct_gk(p, 1e-4)):
        np.testing.assert_array_equal(g, f[t])


def test_newton_third_law_momentum():
    s = random_state(200, seed=4)
    ax, ay, az = O.direct_n3(s, 1e-3)
    m = s["mass"]
    scale = np.abs(m * ax).sum()
    assert abs((m * ax).sum()) < 1e-13 * scale and abs((m * az).sum()) < 1e-13 * scale


# ------------------------------------------------------------------ Verlet ---
def test_verlet_two_body_circular_orbit_energy():
    """SoAStabilityTests.cs:69-91: circular two-body orbit, |dE/E| < 1e-4 over 10k steps."""
    M, m, r = 1.0, 1e-3, 1.0
    v = np.sqrt((M + m) / r)
    s = state([0, r], [0, 0], [0, 0], [M, m], vx=[0, 0], vy=[-v * m / (M + m), v * M / (M + m)])
    s["px"] -= m * r / (M + m)
    e0 = O.energy_naive(s, 0.0)
    e0 = e0["ke"] + e0["pe"]
    acc = O.direct_n3(s, 0.0)
    for k, a in zip(("oax", "oay", "oaz"), acc):
        s[k][:] = a
    for _ in range(10000):
        O.verlet_step(s, 1e-3, lambda st: O.direct_n3(st, 0.0))
    e1 = O.energy_naive(s, 0.0)
    assert abs((e1["ke"] + e1["pe"] - e0) / e0) < 1e-4


def test_verlet_three_body_momentum():
    """SoAStabilityTests.cs:139-172: total momentum conserved to 1e-8 relative."""
    s = state([0, 1, -1], [0, 0.2, 0.1], [0, 0, 0.3], [1.0, 0.5, 0.7],
              vx=[0, 0.1, -0.2], vy=[0.3, -0.1, 0.0], vz=[0.0, 0.0, 0.1])
    p0 = np.array([np.sum(s["mass"] * s[k]) for k in ("vx", "vy", "vz")])
    for _ in range(2000):
        O.verlet_step(s, 1e-3, lambda st: O.direct_n3(st, 1e-2))
    p1 = np.array([np.sum(s["mass"] * s[k]) for k in ("vx", "vy", "vz")])
    assert np.abs(p1 - p0).max() < 1e-8 * np.abs(s["mass"]).sum()


def test_verlet_inactive_bodies_frozen_but_rotated():
    s = random_state(10, seed=5)
    s["active"][3] = 0
    x3, v3 = s["px"][3], s["vx"][3]
    O.verlet_step(s, 1e-3, lambda st: O.direct_n3(st, 1e-3))
    assert s["px"][3] == x3 and s["vx"][3] == v3 and s["oax"][3] == 0.0


# ------------------------------------------------- reference Barnes-Hut ---
def cluster128():
    a = np.fromfile(os.path.join(HERE, "golden", "cluster128.bin")).reshape(8, 128)
    return state(a[0], a[1], a[2], a[6], a[3], a[4], a[5], radius=a[7])


def test_cluster128_fixture_first_body():
    """Values of std::mt19937_64(12345) + uniform_real_distribution (libstdc++)."""
    s = cluster128()
    assert s["px"][0] == pytest.approx(-1.42370277, abs=1e-8)
    assert s["mass"][0] == pytest.approx(0.628982147, abs=1e-9)


def test_bh_rms_error_cluster128():
    """engine/tests/test_phase18_19.cpp:74-137: N=128 cluster, theta=0.5,
    eps=1e-3.  That test asserts RMS < 1%, but the reference engine's own
    BarnesHutSolver, compiled from its sources (oracle/_ref, test_reference.py),
    gives 1.2375% on exactly this cluster, as does this restatement: the
    reference's bound does not hold for its own code (the engine tests never
    run in its CI).  Pinned here at the value both produce."""
    s = cluster128()
    ref = O.direct_gk(s, 1e-3)
    bh, _ = O.bh_csharp(s, 1e-3, 0.5)
    e = O.force_error(ref, bh)
    assert e["count"] == 128
    assert e["rms"] == pytest.approx(0.0123747693648392, rel=1e-9)


def test_bh_two_body_equals_direct():
    """BarnesHutTests.cs:65-97."""
    s = state([0, 1.3], [0, 0.2], [0, -0.1], [1.0, 2e-3])
    bh, _ = O.bh_csharp(s, 1e-4, 0.5)
    np.testing.assert_allclose(np.array(bh), np.array(O.direct_gk(s, 1e-4)), rtol=1e-10)


def test_bh_theta_clamp_bit_identical():
    """BarnesHutAccuracyTests.cs:165-186."""
    s = random_state(500, seed=6)
    a, _ = O.bh_csharp(s, 1e-3, 0.0)
    b, _ = O.bh_csharp(s, 1e-3, 0.2)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_bh_error_monotone_in_theta():
    """BarnesHutTests.cs:196-222."""
    s = random_state(2000, seed=7)
    ref = O.direct_gk(s, 1e-3)
    errs = [O.force_error(ref, O.bh_csharp(s, 1e-3, t)[0])["mean"] for t in (0.3, 0.6, 0.9)]
    assert errs[0] < errs[1] < errs[2]


def test_bh_coincident_bodies_aggregate():
    """BarnesHutBackend.cs:279-292: coincident bodies share one leaf; no infinite split."""
    s = state([0.5, 0.5, 2.0], [0, 0, 0], [0, 0, 0], [1.0, 1.0, 1.0])
    (ax, ay, az), st = O.bh_csharp(s, 1e-2, 0.5)
    assert np.isfinite(ax).all() and st.node_count < 200


# -------------------------------------------- build-defined octree pins ---
def test_morton_top_bits_are_csharp_octants():
    """Top key group = OctreeNode.OctantFor relative to the root centre
    (x->4, y->2, z->1), away from the centre planes."""
    s = random_state(3000, seed=8)
    box = O.root_box(s)
    keys = O.morton_keys(s, box, 10)
    dx, dy, dz = s["px"] - box[0], s["py"] - box[1], s["pz"] - box[2]
    far = (np.abs(dx) > 1e-6) & (np.abs(dy) > 1e-6) & (np.abs(dz) > 1e-6)
    oct_cs = (dx >= 0) * 4 + (dy >= 0) * 2 + (dz >= 0) * 1
    assert ((keys[far] >> 27) == oct_cs[far]).all()


def test_root_box_matches_reference_formula():
    s = random_state(100, seed=9)
    box = O.root_box(s)
    for k, c in zip(range(3), ("px", "py", "pz")):
        assert box[k] == (s[c].min() + s[c].max()) * 0.5
    ext = max(s[c].max() - s[c].min() for c in ("px", "py", "pz"))
    assert box[3] == ext * 0.5 * 1.001


def test_stable_sort_matches_numpy_stable_argsort():
    rng = np.random.default_rng(10)
    keys = rng.integers(0, 50, 5000).astype(np.uint64) << np.uint64(20)
    ks, order = O.stable_sort_keys(keys)
    ref = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(order, ref)
    np.testing.assert_array_equal(ks, keys[ref])


@pytest.mark.parametrize("levels,leaf", [(10, 16), (21, 1), (21, 8)])
def test_tree_invariants(levels, leaf):
    s = random_state(4000, seed=11)
    s["active"][::17] = 0
    t = O.tree_build(s, levels, leaf)
    na = int(s["active"].sum())
    assert t["n_active"] == na and t["count"][0] == na
    x = s["px"][t["order"]].astype(np.longdouble)
    m = s["mass"][t["order"]].astype(np.longdouble)
    for k in range(t["n_nodes"]):
        st, c, f, nc = t["start"][k], t["count"][k], t["first"][k], t["nchild"][k]
        if f >= 0:
            ch = np.arange(f, f + nc)
            assert (t["level"][ch] == t["level"][k] + 1).all()
            assert t["start"][f] == st and t["count"][ch].sum() == c
            assert (np.diff(t["start"][ch]) == t["count"][ch][:-1]).all()
            assert c > leaf
        else:
            assert c <= leaf or t["level"][k] == levels
        keys = t["keys"][st:st + c]
        sh = np.uint64(3 * (levels - t["level"][k]))
        assert len(set((keys >> sh).tolist())) <= 1
        M = m[st:st + c].sum()
        assert abs(float(M) - t["mass"][k]) <= 1e-15 * float(M)
        cx = float((m[st:st + c] * x[st:st + c]).sum() / M)
        assert abs(cx - t["com"][k][0]) <= 1e-14 * max(1.0, abs(cx))


def test_group_traversal_reduces_to_reference_bh():
    """With one body per group and per leaf, 63-bit keys and well-separated
    bodies, the group-MAC traversal is the reference's per-body traversal on
    the same (geometric) octree: forces agree with the C# restatement."""
    rng = np.random.default_rng(12)
    g = rng.permutation(np.array(np.meshgrid(*(np.arange(12),) * 3)).reshape(3, -1).T)[:800]
    p = g + rng.uniform(0.2, 0.8, g.shape)       # one body per unit cell: no near ties
    s = state(p[:, 0], p[:, 1], p[:, 2], rng.uniform(0.5, 2, len(p)))
    ref, _ = O.bh_csharp(s, 1e-3, 0.5)
    got, st = O.bh_tree(s, 1e-3, 0.5, levels=21, leaf_cap=1, group_size=1, split_depth=0,
                        lifo=1536)
    assert rel(got, ref).max() < 1e-12


def test_group_traversal_converges_to_direct():
    s = random_state(3000, seed=13)
    ref = O.direct_gk(s, 1e-3)
    errs = []
    for th in (0.2, 0.5, 0.9):
        a, _ = O.bh_tree(s, 1e-3, th, levels=21, leaf_cap=8, group_size=64, split_depth=-3,
                         lifo=1536)
        errs.append(np.median(rel(a, ref)))
    assert errs[0] < 1e-4 and errs[0] < errs[1] < errs[2]


def test_group_traversal_stack_bounded():
    s = random_state(20000, seed=14, spread=3.0)
    _, st = O.bh_tree(s, 1e-3, 0.5, levels=10, leaf_cap=16, group_size=64, split_depth=-3,
                      lifo=1536)
    assert st["max_queue"] <= 2048


# ------------------------------------------------------------ diagnostics ---
def test_energy_naive_vs_extended():
    s = random_state(300, seed=15)
    a = O.energy_naive(s, 1e-3)
    b = O.energy_exact(s, 1e-3)
    for k in ("ke", "pe", "px", "l_mag"):
        assert abs(a[k] - b[k]) <= 1e-12 * max(1.0, abs(b[k]))
    # brute force in long double, straight from the definitions
    m = s["mass"].astype(np.longdouble)
    ke = 0.5 * (m * (s["vx"] ** 2 + s["vy"] ** 2 + s["vz"] ** 2)).sum()
    assert abs(float(ke) - b["ke"]) <= 1e-15 * float(ke)


def test_force_error_statistics():
    ex = (np.array([1.0, 0.0, 2.0]), np.zeros(3), np.zeros(3))
    te = (np.array([1.1, 5.0, 2.0]), np.zeros(3), np.zeros(3))
    r = O.force_error(ex, te)
    # body 1 has |a_exact| = 0 < 1e-30 and is skipped (ForceErrorAnalyzer.cs:108)
    assert r["count"] == 2
    assert r["max"] == pytest.approx(0.1) and r["mean"] == pytest.approx(0.05)
    assert r["median"] == pytest.approx(0.05)


# ------------------------------------------------------------- collisions ---
def test_cbrt_exact_cubes_and_libm_agreement():
    for k in (1, 2, 3, 10, 123, 4096):
        assert O.cbrt(float(k) ** 3) == float(k)
    rng = np.random.default_rng(16)
    xs = np.exp(rng.uniform(-60, 60, 2000))
    got = np.array([O.cbrt(x) for x in xs])
    ref = np.cbrt(xs)
    ulp = np.spacing(ref)
    assert (np.abs(got - ref) <= ulp).all()
    # compared with an extended-precision root, never worse than libm
    ld = np.cbrt(xs.astype(np.longdouble))
    assert (np.abs(got - ld) <= np.abs(ref - ld) + 1e-30).all()


def spatial_hash_pairs(s):
    """SpatialHashBroadPhase.GetCandidatePairs (SpatialHashBroadPhase.cs:15-94)
    restated in Python, then the narrow phase of TryAddEvent."""
    act = s["active"].astype(bool)
    r = s["radius"]
    cell = max(2.0 * r[act].max(), 1e-6)
    inv = 1.0 / cell
    cells = {}
    for i in np.nonzero(act)[0]:
        key = tuple(int(np.floor(s[c][i] * inv)) for c in ("px", "py", "pz"))
        cells.setdefault(key, []).append(i)
    pairs = set()
    for (cx, cy, cz), home in cells.items():
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for a in home:
                        for b in cells.get((cx + dx, cy + dy, cz + dz), ()):
                            if a < b:
                                pairs.add((a, b))
    out = []
    for a, b in sorted(pairs):
        if not (s["collidable"][a] and s["collidable"][b]):
            continue
        d2 = sum((s[c][a] - s[c][b]) ** 2 for c in ("px", "py", "pz"))
        if d2 < (r[a] + r[b]) ** 2:
            out.append((a, b))
    return out


def test_detect_equals_spatial_hash():
    """CollisionTests.cs:420-453: the spatial-hash candidate set, after the
    narrow phase, equals brute force (160 bodies)."""
    rng = np.random.default_rng(42)
    s = state(*(rng.uniform(-1, 1, 160) for _ in range(3)), rng.uniform(1, 2, 160),
              radius=rng.uniform(0.02, 0.12, 160))
    s["active"][7] = 0
    ev = O.detect(s)
    got = sorted((min(a, b), max(a, b)) for a, b in zip(ev["a"], ev["b"]))
    assert got == spatial_hash_pairs(s) and len(got) > 10
    assert (np.diff(ev["overlap"]) <= 0).all()
    heavier = s["mass"][ev["a"]] >= s["mass"][ev["b"]]
    assert heavier.all()


def test_merge_conservation_and_survivor():
    """CollisionTests.cs:159-322 / MergePolicy.cs:31-80."""
    rng = np.random.default_rng(17)
    n = 300
    s = state(*(rng.uniform(-1, 1, n) for _ in range(3)), rng.uniform(1, 2, n),
              *(rng.normal(0, 1, n) for _ in range(3)), radius=np.full(n, 0.08))
    s["density"][:] = 1000.0
    m0 = s["mass"].sum()
    p0 = np.array([(s["mass"] * s[k]).sum() for k in ("vx", "vy", "vz")])
    ev = O.detect(s)
    k = O.resolve_merge(ev, s)
    act = s["active"].astype(bool)
    assert k > 0 and act.sum() == n - k
    assert abs(s["mass"][act].sum() - m0) <= 1e-13 * m0
    p1 = np.array([(s["mass"][act] * s[c][act]).sum() for c in ("vx", "vy", "vz")])
    assert np.abs(p1 - p0).max() < 1e-12 * np.abs(s["mass"]).sum()
    # absorbed bodies are zeroed and non-collidable
    assert (s["oax"][~act] == 0).all() and (s["collidable"][~act] == 0).all()
    # survivors' radii follow the density model
    i = np.nonzero(act & (s["mass"] > 2.0))[0][0]
    assert s["radius"][i] == O.body_radius(s["mass"][i], s["density"][i], 0)


def test_merge_black_hole_promotion():
    """CollisionTests.cs:107-123: absorbing a black hole makes the survivor one."""
    s = state([0, 0.01], [0, 0], [0, 0], [10.0, 5.0], radius=[0.1, 0.1])
    s["body_type"][:] = [0, 7]
    s["density"][:] = 1000.0
    ev = O.detect(s)
    assert O.resolve_merge(ev, s) == 1
    assert s["body_type"][0] == 7 and s["active"][1] == 0
    assert s["radius"][0] == O.body_radius(15.0, s["density"][0], 7)


def test_compaction_order():
    """engine/tests/test_phase14_15.cpp:172-196 (Compact.Basic)."""
    s = state([1.0, 2.0, 3.0, 4.0], [0] * 4, [0] * 4, [1.0, 2.0, 3.0, 4.0],
              radius=[0.1, 0.2, 0.3, 0.4], active=[1, 0, 1, 0])
    out, idx = O.compact(s)
    assert list(out["px"]) == [1.0, 3.0] and list(out["radius"]) == [0.1, 0.3]
    assert list(idx) == [0, 2]


def dense_leaf_state(n_core=3000, seed=23):
    """A tight cluster far smaller than one 30-bit key cell of a root box
    stretched by a few distant bodies: one deepest-level leaf holds n_core bodies."""
    rng = np.random.default_rng(seed)
    core = rng.uniform(-1e-3, 1e-3, (n_core, 3))
    far = rng.uniform(-100, 100, (10, 3))
    p = np.vstack([core, far])
    return state(p[:, 0], p[:, 1], p[:, 2], rng.uniform(0.5, 2, len(p)))


@pytest.mark.parametrize("policy", [3, -1, -2, -3])
def test_groups_cover_bodies_in_oversized_leaves(policy):
    """Every active body gets a force when a deepest-level leaf holds more
    bodies than a group: the groups partition the bodies (the oracle raises
    otherwise) and the result matches the direct sum."""
    s = dense_leaf_state()
    t = O.tree_build(s, 10, 16)
    leaf_counts = t["count"][t["first"] < 0]
    assert leaf_counts.max() > 64
    a, st = O.bh_tree(s, 1e-4, 0.5, levels=10, leaf_cap=16, group_size=64,
                      split_depth=policy)
    ref = O.direct_gk(s, 1e-4)
    assert rel(a, ref).max() < 1e-2
    assert (st["n_parts"] + st["n_cells"] > 0).all()


def test_pair_collapse_state_overflows_only_after_the_first_drift():
    """The state of the sticky pool-overflow GPU test: with leaves of one body
    and 21-bit keys, the tree after the first drift needs more nodes than the
    library's pool (3 n + 1024), the tree after the second drift fewer."""
    from conftest import pair_collapse_state, soa_to_state
    soa = pair_collapse_state()
    s = soa_to_state(soa)
    n = soa.count
    pool = 3 * n + 1024
    counts = []
    for _ in range(2):
        O.verlet_step(s, 1e-3, lambda st: (np.zeros(n), np.zeros(n), np.zeros(n)))
        counts.append(O.tree_build(s, 21, 1, node_cap=40 * n)["n_nodes"])
    assert counts[0] > pool > counts[1], (counts, pool)
