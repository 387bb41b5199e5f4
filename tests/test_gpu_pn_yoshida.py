"""GPU parity of the widening rows, through the C ABI: the 1PN decorator
(PostNewtonianBackend + PostNewtonian1Correction), the engine's per-type and
adaptive softening of the direct sum, and the Yoshida-4 integrator.  All FP64
and bit-exact against the oracle; then the reference's PostNewtonianTests
(precession, drift, toggle) and the integrator's convergence order."""
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


def relativistic_cluster(n, seed=11, massless=True):
    """Plummer positions with velocities of ~1 % of c, some inactive and
    (optionally) massless bodies, one body beyond the 0.3 c cut and one
    strong-field pair."""
    soa = cmhip.generate_plummer(n, 1.0, 1.0, seed=seed, mass_fraction_cut=0.99)
    rng = np.random.default_rng(seed)
    c = O.c_sim()
    for f in ("vel_x", "vel_y", "vel_z"):
        getattr(soa, f)[:n] = rng.normal(0, 0.01 * c, n)
    soa.mass[:n] = rng.uniform(0.5, 20.0, n)
    soa.mass[:n][::19] = 0.0 if massless else 1e-3
    soa.is_active[:n][3::13] = 0
    soa.vel_x[5] = 0.5 * c
    for f in ("pos_x", "pos_y", "pos_z"):
        getattr(soa, f)[9] = getattr(soa, f)[8] + 1e-6
    soa.mass[8] = soa.mass[9] = 3000.0
    return soa


def _pn_ref(s, base):
    (ax, ay, az), warns, nw = O.pn_apply(s, EPS, *base)
    return (ax, ay, az), warns, nw


def test_pn_direct_bit_exact_with_warnings():
    n = 700
    soa = relativistic_cluster(n)
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(n, kind=cmhip.BACKEND_DIRECT)
    be.set_post_newtonian(True)
    be.compute_forces(soa, EPS)
    ref, warns, nw = _pn_ref(s, O.direct_gk(s, EPS))
    _eq(_acc(soa), ref)
    got, total = be.pn_warnings()
    assert total == nw and nw >= 2
    assert [(w[0], w[1]) for w in got] == [(w[0], w[1]) for w in warns]
    np.testing.assert_array_equal([w[2] for w in got], [w[2] for w in warns])
    np.testing.assert_array_equal([w[3] for w in got], [w[3] for w in warns])


def test_pn_warnings_past_the_cap_are_the_first_in_loop_order():
    """More strong-field pairs than the 4096 kept: the kept ones are the first
    4096 of the reference's (i, j) loop (PostNewtonian1Correction.cs:124-129
    raises them in that order), with the oracle's distances and radii bit for
    bit, on every evaluation; the total counts them all."""
    n = 400
    soa = relativistic_cluster(n, seed=21)
    rng = np.random.default_rng(21)
    core = np.arange(100, 230)   # 130 heavy bodies inside 1e-4 AU: every live pair is strong
    for f in ("pos_x", "pos_y", "pos_z"):
        getattr(soa, f)[core] = 0.2 + rng.uniform(0.0, 1e-4, core.size)
    soa.mass[core] = 3000.0
    soa.mass[core[::17]] = 0.0           # massless sources are skipped before the test
    soa.is_active[core[5::29]] = 0
    s = soa_to_state(soa)
    ref, warns, nw = _pn_ref(s, O.direct_gk(s, EPS))
    assert nw > 3 * 4096
    be = cmhip.HipPhysicsBackend(n, kind=cmhip.BACKEND_DIRECT)
    be.set_post_newtonian(True)
    for rep in range(2):
        out = soa.copy()
        be.compute_forces(out, EPS)
        _eq(_acc(out), ref)
        got, total = be.pn_warnings(cap=4096)
        assert total == nw, rep
        assert [(w[0], w[1]) for w in got] == [(w[0], w[1]) for w in warns[:4096]], rep
        np.testing.assert_array_equal([w[2] for w in got], [w[2] for w in warns[:4096]])
        np.testing.assert_array_equal([w[3] for w in got], [w[3] for w in warns[:4096]])
    be.close()


def test_pn_barnes_hut_fp64_bit_exact():
    n = 3000
    soa = relativistic_cluster(n, seed=12, massless=False)
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP64, theta=0.5)
    be.set_post_newtonian(True)
    be.compute_forces(soa, EPS)
    newton, _ = O.bh_tree(s, EPS, 0.5, nthreads=8)
    ref, _, _ = _pn_ref(s, newton)
    _eq(_acc(soa), ref)


def test_pn_verlet_and_yoshida_steps_bit_exact():
    n = 300
    soa = relativistic_cluster(n, seed=13)
    soa.pos_x[:n] *= 3.0       # keep the strong-field pair, spread the rest
    s0 = soa_to_state(soa)
    force = lambda st: O.pn_apply(st, EPS, *O.direct_n3(st, EPS))[0]
    for integ in ("verlet", "yoshida"):
        be = cmhip.HipPhysicsBackend(n, kind=cmhip.BACKEND_DIRECT_PAIRWISE)
        be.set_post_newtonian(True)
        be.upload(soa)
        st = {k: v.copy() for k, v in s0.items()}
        steps, dt = 20, 1e-5
        if integ == "verlet":
            be.step(EPS, dt, steps)
            for _ in range(steps):
                O.verlet_step(st, dt, force)
        else:
            be.yoshida_step(EPS, dt, steps)
            for _ in range(steps):
                O.yoshida_step(st, dt, force)
        out = cmhip.BodySoA(n)
        be.download(out)
        for f, k in (("pos_x", "px"), ("pos_y", "py"), ("pos_z", "pz"), ("vel_x", "vx"),
                     ("vel_y", "vy"), ("vel_z", "vz"), ("old_acc_x", "oax")):
            np.testing.assert_array_equal(getattr(out, f)[:n], st[k], err_msg=f"{integ} {f}")
        be.close()


def _binary(mass, r, v, plane="y"):
    vel = {"vel_y": np.array([v, -v])} if plane == "y" else {"vel_z": np.array([v, -v])}
    return cmhip.BodySoA.from_arrays(pos_x=np.array([r, -r]), mass=np.array([mass, mass]),
                                     is_active=np.ones(2, np.uint8), **vel)


def test_pn_precession_reference_case():
    """PostNewtonianTests.MercuryPerihelionPrecession_PNEnabled_ShowsPrecession:
    two 50 Msun bodies at +-0.5 AU, v = sqrt(m / (4 r)), dt = 1e-4, 20,000
    steps: the 1PN run's angle differs from the Newtonian run's by > 1e-10."""
    mass, r = 50.0, 0.5
    v = np.sqrt(mass / (4.0 * r))
    ang = {}
    for pn in (True, False):
        be = cmhip.HipPhysicsBackend(2, kind=cmhip.BACKEND_DIRECT_PAIRWISE)
        be.set_post_newtonian(pn)
        soa = _binary(mass, r, v)
        be.upload(soa)
        be.step(1e-4, 1e-4, 20000)
        out = cmhip.BodySoA(2)
        be.download(out)
        ang[pn] = np.arctan2(out.pos_y[0], out.pos_x[0])
    diff = abs(ang[True] - ang[False])
    print(f"1PN vs Newton angle after 20k steps: {diff:.4e} rad")
    assert diff > 1e-10


def test_pn_energy_drift_reference_case():
    """PostNewtonianTests.EnergyDrift_NewtonVs1PN_Comparable: 10 Msun binary,
    r = 1, dt = 1e-3, 10,000 steps with 1PN on: Newtonian energy drift < 1 %."""
    mass, r = 10.0, 1.0
    v = np.sqrt(mass / (4.0 * r))
    be = cmhip.HipPhysicsBackend(2, kind=cmhip.BACKEND_DIRECT_PAIRWISE)
    be.set_post_newtonian(True)
    be.upload(_binary(mass, r, v))
    e = be.energy(1e-4)
    e0 = e["kinetic"] + e["potential"]
    be.step(1e-4, 1e-3, 10000)
    e = be.energy(1e-4)
    drift = abs((e["kinetic"] + e["potential"] - e0) / e0)
    print(f"1PN binary, 10k steps: |dE/E| = {drift:.3e}")
    assert drift < 0.01


def test_pn_toggle_off_is_newton_bit_for_bit():
    """PostNewtonianTests.ToggleRegression_PNDisabled_MatchesNewton."""
    v = np.sqrt(1.0 / 4.0)
    out = []
    for mode in ("never", "off", "on_then_off"):
        be = cmhip.HipPhysicsBackend(2, kind=cmhip.BACKEND_DIRECT_PAIRWISE)
        if mode == "off":
            be.set_post_newtonian(False)
        if mode == "on_then_off":
            be.set_post_newtonian(True)
            be.set_post_newtonian(False)
        be.upload(_binary(1.0, 1.0, v, plane="z"))
        be.step(1e-4, 1e-3, 1000)
        o = cmhip.BodySoA(2)
        be.download(o)
        out.append(np.concatenate([o.pos_x[:2], o.pos_y[:2], o.pos_z[:2]]))
    np.testing.assert_array_equal(out[0], out[1])
    np.testing.assert_array_equal(out[0], out[2])


@pytest.mark.parametrize("mode", [cmhip.SOFTENING_PER_TYPE, cmhip.SOFTENING_ADAPTIVE])
@pytest.mark.parametrize("kind", [cmhip.BACKEND_DIRECT, cmhip.BACKEND_DIRECT_PAIRWISE])
def test_softening_modes_bit_exact(mode, kind):
    n = 1100
    soa = cmhip.generate_plummer(n, 1.0, 1.0, seed=21)
    rng = np.random.default_rng(21)
    soa.mass[:n] = 10.0 ** rng.uniform(-9, 0, n)
    soa.body_type[:n] = rng.integers(0, 40, n)      # types index the table mod 16
    soa.is_active[:n][rng.random(n) < 0.1] = 0
    table = rng.uniform(1e-4, 5e-2, 16)
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(n, kind=kind)
    be.set_softening(mode, type_softening=table, adaptive_scale=0.02, adaptive_min=1e-5)
    be.compute_forces(soa, EPS)
    _eq(_acc(soa), O.direct_softened(s, mode, table, scale=0.02, mn=1e-5, global_eps=EPS))


def test_softening_mode_leaves_barnes_hut_alone():
    """The engine applies softening modes to the brute force only; its
    Barnes-Hut uses the global softening."""
    n = 2000
    soa = cmhip.generate_plummer(n, 1.0, 1.0, seed=22)
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP64)
    be.set_softening(cmhip.SOFTENING_ADAPTIVE, adaptive_scale=0.5)
    be.compute_forces(soa, EPS)
    ref, _ = O.bh_tree(s, EPS, 0.5, nthreads=8)
    _eq(_acc(soa), ref)


def test_yoshida_barnes_hut_fp64_bit_exact():
    n = 2000
    soa = cmhip.generate_plummer(n, 1.0, 1.0, seed=23, mass_fraction_cut=0.999)
    s = soa_to_state(soa)
    be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP64)
    be.upload(soa)
    be.yoshida_step(EPS, 1e-3, 3)
    out = cmhip.BodySoA(n)
    be.download(out)
    for _ in range(3):
        O.yoshida_step(s, 1e-3, lambda st: O.bh_tree(st, EPS, 0.5, nthreads=8)[0])
    for f, k in (("pos_x", "px"), ("vel_y", "vy"), ("acc_z", "az"), ("old_acc_x", "oax")):
        np.testing.assert_array_equal(getattr(out, f)[:n], s[k])


def test_yoshida_fourth_order_on_gpu():
    """Circular equal-mass binary, one period: halving dt divides the Yoshida
    error by ~16 and the Verlet error by ~4."""
    def err(steps, yoshida):
        v = 0.5
        soa = cmhip.BodySoA.from_arrays(pos_x=np.array([0.5, -0.5]), vel_y=np.array([v, -v]),
                                        mass=np.array([0.5, 0.5]),
                                        is_active=np.ones(2, np.uint8))
        be = cmhip.HipPhysicsBackend(2, kind=cmhip.BACKEND_DIRECT)
        be.upload(soa)
        be.compute_forces_resident(0.0)
        o = cmhip.BodySoA(2)
        be.download(o)
        o.old_acc_x[:2], o.old_acc_y[:2] = o.acc_x[:2], o.acc_y[:2]
        be.upload(o)
        dt = 2 * np.pi / steps
        (be.yoshida_step if yoshida else be.step)(0.0, dt, steps)
        be.download(o)
        return np.hypot(o.pos_x[0] - 0.5, o.pos_y[0])
    ry = err(40, True) / err(80, True)
    rv = err(40, False) / err(80, False)
    print(f"error ratio on halving dt: Yoshida {ry:.2f}, Verlet {rv:.2f}")
    assert 12.0 < ry < 20.0 and 3.0 < rv < 5.0


def test_widening_entry_points_edge_cases():
    """Empty and one-body states, zero step counts and the argument checks of
    the widening entry points (error codes, no state change)."""
    be = cmhip.HipPhysicsBackend(16, kind=cmhip.BACKEND_DIRECT)
    one = cmhip.BodySoA.from_arrays(pos_x=np.array([0.3]), vel_y=np.array([0.1]),
                                    mass=np.array([2.0]), is_active=np.ones(1, np.uint8))
    be.set_post_newtonian(True)
    be.upload(one)
    be.compute_forces_resident(EPS)
    out = cmhip.BodySoA(1)
    be.download(out)
    assert out.acc_x[0] == 0.0 and out.acc_y[0] == 0.0
    assert be.pn_warnings() == ([], 0)
    be.yoshida_step(EPS, 1e-3, 0)
    assert be.nbody_step(EPS, 1e-3, 0) == 0
    be.set_post_newtonian(False)
    assert be.pn_warnings() == ([], 0)
    with pytest.raises(cmhip.CmError):
        be.set_softening(7)
    with pytest.raises(cmhip.CmError):
        be.set_softening(cmhip.SOFTENING_PER_TYPE, type_softening=[-1.0] * 16)
    with pytest.raises(cmhip.CmError):
        be.configure_collisions(mode=cmhip.COLLISION_REALISTIC, max_merges_per_frame=4)
    with pytest.raises(cmhip.CmError):
        be.configure_collisions(fragmentation_mass_loss_cap=0.01)
    with pytest.raises(cmhip.CmError):
        be.configure_collisions(mode=5)
    # Configure's clamps: restitution 3 -> 1 is accepted
    be.configure_collisions(mode=cmhip.COLLISION_BOUNCE_ONLY, default_restitution=3.0)
    empty = cmhip.BodySoA.from_arrays(pos_x=np.zeros(0), mass=np.zeros(0),
                                      is_active=np.zeros(0, np.uint8))
    be.upload(empty)
    be.compute_forces_resident(EPS)
    assert be.detect_collisions(0)[1] == 0
    assert be.resolve_collisions() == 0
    assert be.nbody_step(EPS, 1e-3, 3) == 0
    be.close()
