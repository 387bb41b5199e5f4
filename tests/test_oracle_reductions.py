"""Pins the oracle's centre of mass (the engine's gpu_reduce_com,
engine/src/cuda/gpu_reduction.cu:232-262 and :481-517) and maximum
acceleration (Engine::compute_max_acceleration, engine/src/sim/engine.cpp:
468-500) before they judge the GPU: closed forms, exclusion of inactive
bodies, the empty case, and invariants that a dropped mass weight, a swapped
axis or a missing square root would break."""
import math

import numpy as np

import oracle as O
from test_oracle import random_state, state


def test_center_of_mass_closed_form():
    """Masses 1 and 3 at x = 0 and x = 4: COM at x = 3 (the unweighted mean,
    2, would fail); each axis from its own coordinate."""
    s = state([0.0, 4.0], [1.0, -1.0], [2.0, 6.0], [1.0, 3.0])
    assert O.center_of_mass(s) == (3.0, -0.5, 5.0, 4.0)


def test_center_of_mass_excludes_inactive_and_handles_empty():
    s = state([0.0, 4.0, 100.0], [0.0, 0.0, 100.0], [0.0, 0.0, 100.0], [1.0, 1.0, 1e6],
              active=[1, 1, 0])
    assert O.center_of_mass(s) == (2.0, 0.0, 0.0, 2.0)
    s["active"][:] = 0
    assert O.center_of_mass(s) == (0.0, 0.0, 0.0, 0.0)
    e = state([], [], [], [])
    assert O.center_of_mass(e) == (0.0, 0.0, 0.0, 0.0)


def test_center_of_mass_translates_and_matches_exact_sums():
    s = random_state(5000, seed=3)
    c = O.center_of_mass(s)
    m = s["mass"]
    for k, ax in enumerate(("px", "py", "pz")):
        # exactly rounded sums of the exactly rounded products, then one quotient
        want = math.fsum(m * s[ax]) / math.fsum(m)
        assert abs(c[k] - want) <= 4e-16 * np.abs(s[ax]).max()
    assert c[3] == math.fsum(m)
    # a million equal masses: a plain long double running sum would drift by
    # ~1e-14 relative; the sum must stay exactly rounded
    k = 1 << 20
    big = state(np.zeros(k), np.zeros(k), np.zeros(k), np.full(k, 0.999 / k * 1.1))
    assert O.center_of_mass(big)[3] == math.fsum(big["mass"])
    d = (10.0, -7.5, 3.25)
    t = dict(s, px=s["px"] + d[0], py=s["py"] + d[1], pz=s["pz"] + d[2])
    c2 = O.center_of_mass(t)
    for k in range(3):
        assert abs(c2[k] - (c[k] + d[k])) <= 1e-13
    assert c2[3] == c[3]


def test_max_acceleration_closed_form_and_inactive():
    """|(3, 4, 0)| = 5 and |(0, 0, -13)| = 13 (a missing square root would give
    25 and 169, a dropped component 3 or 0); an inactive body with a larger
    acceleration is skipped; none active gives 0."""
    ax = np.array([3.0, 0.0, 1e9])
    ay = np.array([4.0, 0.0, 0.0])
    az = np.array([0.0, -13.0, 0.0])
    act = np.array([1, 1, 0], np.uint8)
    assert O.max_acceleration((ax, ay, az), act) == 13.0
    act[1] = 0
    assert O.max_acceleration((ax, ay, az), act) == 5.0
    act[:] = 0
    assert O.max_acceleration((ax, ay, az), act) == 0.0
    assert O.max_acceleration((ax[:0], ay[:0], az[:0]), act[:0]) == 0.0


def test_max_acceleration_is_the_largest_magnitude():
    """Against the magnitude of every active body, one at a time, computed as
    math.hypot (a different, correctly rounded-to-within-1-ulp route)."""
    rng = np.random.default_rng(9)
    a = [rng.normal(0, 1, 4000) * 10.0 ** rng.uniform(-3, 3, 4000) for _ in range(3)]
    act = (rng.random(4000) < 0.8).astype(np.uint8)
    got = O.max_acceleration(a, act)
    mags = [math.hypot(a[0][i], a[1][i], a[2][i]) for i in range(4000) if act[i]]
    assert abs(got - max(mags)) <= 2e-16 * max(mags)
