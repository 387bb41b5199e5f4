"""The oracle against the reference's own engine code.

oracle/ref.mk compiles the reference engine's CPU Barnes-Hut, octree builder
and ParticleSystem (engine/src/physics, engine/src/job, engine/src/core) from
where they lie into oracle/_ref/libcmref.so.  That code is an independent
implementation of the same algorithm as the C# BarnesHutBackend (octant bits
in the opposite order, so children are visited in a different order): the
oracle's C# restatement must agree with it to rounding.  Skipped when the
reference tree (and so the build) is absent."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from test_oracle import cluster128, random_state, rel, state

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "oracle", "_ref", "libcmref.so")
pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="reference engine not built")


def _lib():
    L = C.CDLL(REF)
    L.ref_bh_forces.argtypes = [C.c_int] + [C.c_void_p] * 5 + [C.c_double, C.c_double] + \
        [C.c_void_p] * 3
    L.ref_compact.argtypes = [C.c_int] + [C.c_void_p] * 4
    L.ref_pn_apply.argtypes = [C.c_int] + [C.c_void_p] * 8 + [C.c_double] + [C.c_void_p] * 3
    return L


def ref_bh(s, eps, theta):
    n = len(s["px"])
    out = [np.zeros(n) for _ in range(3)]
    arrs = [np.ascontiguousarray(s[k]) for k in ("px", "py", "pz", "mass")]
    act = np.ascontiguousarray(s["active"], dtype=np.uint8)
    _lib().ref_bh_forces(n, *(a.ctypes.data for a in arrs), act.ctypes.data, eps, theta,
                         *(o.ctypes.data for o in out))
    return out


def test_cluster128_matches_engine_barnes_hut():
    s = cluster128()
    r = ref_bh(s, 1e-3, 0.5)
    mine, _ = O.bh_csharp(s, 1e-3, 0.5)
    assert rel(mine, r).max() < 1e-13
    e = O.force_error(O.direct_gk(s, 1e-3), r)
    assert e["rms"] == pytest.approx(0.0123747693648392, rel=1e-9)


@pytest.mark.parametrize("theta", [0.3, 0.5, 0.9])
def test_random_systems_match_engine_barnes_hut(theta):
    s = random_state(5000, seed=21)
    s["active"][::13] = 0
    r = ref_bh(s, 1e-3, theta)
    mine, _ = O.bh_csharp(s, 1e-3, theta)
    act = s["active"].astype(bool)
    assert rel([a[act] for a in mine], [b[act] for b in r]).max() < 1e-12
    assert all((a[~act] == 0).all() for a in r)


def test_compaction_matches_engine():
    """ParticleSystem::compact (engine/src/physics/particle_system.cpp:156-184)."""
    rng = np.random.default_rng(22)
    n = 1000
    s = state(rng.normal(size=n), np.zeros(n), np.zeros(n), rng.uniform(1, 2, n),
              radius=rng.uniform(0, 1, n), active=(rng.random(n) < 0.6).astype(np.uint8))
    px, m, r, a = (np.ascontiguousarray(s[k]).copy() for k in ("px", "mass", "radius", "active"))
    k = _lib().ref_compact(n, px.ctypes.data, m.ctypes.data, r.ctypes.data, a.ctypes.data)
    out, idx = O.compact(s)
    assert k == len(idx)
    np.testing.assert_array_equal(px[:k], out["px"])
    np.testing.assert_array_equal(r[:k], out["radius"])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_pn_matches_engine_port_bit_for_bit(seed):
    """PostNewtonianCorrection::apply_corrections (engine/src/physics/
    post_newtonian.cpp:7-96), the engine's port of the C# class, against the
    oracle's restatement of PostNewtonian1Correction: identical arithmetic,
    so identical bits, including skipped zero-mass, fast and strong-field
    pairs and inactive targets."""
    rng = np.random.default_rng(seed)
    n = 300
    c = O.c_sim()
    s = state(*(rng.normal(0, 1, n) for _ in range(3)), rng.uniform(0.5, 40, n),
              *(rng.normal(0, 0.01 * c, n) for _ in range(3)))
    s["mass"][::17] = 0.0
    s["vx"][5] = 0.6 * c                      # beyond the 0.3 c cut for most partners
    for k in ("px", "py", "pz"):              # a strong-field pair
        s[k][9] = s[k][8] + 1e-7
    s["mass"][8] = s["mass"][9] = 2000.0
    s["active"][::11] = 0
    a0 = [rng.normal(0, 1, n) for _ in range(3)]
    (ox, oy, oz), warns, nw = O.pn_apply(s, 1e-4, *a0)
    assert nw >= 2
    rx, ry, rz = (np.array(a, copy=True) for a in a0)
    arrs = [np.ascontiguousarray(s[k]) for k in ("px", "py", "pz", "vx", "vy", "vz", "mass")]
    act = np.ascontiguousarray(s["active"], dtype=np.uint8)
    _lib().ref_pn_apply(n, *(a.ctypes.data for a in arrs), act.ctypes.data, 1e-4,
                        rx.ctypes.data, ry.ctypes.data, rz.ctypes.data)
    for p, q in ((ox, rx), (oy, ry), (oz, rz)):
        np.testing.assert_array_equal(p, q)
