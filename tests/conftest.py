"""Shared test setup: the gpu marker, import paths, built libraries."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "celestial-mechanics-desktop_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcmhip.so)")


def _ensure_built():
    lib = os.path.join(PKG, "libcmhip.so")
    orc = os.path.join(ROOT, "oracle", "liboracle_cm.so")
    if not (os.path.exists(lib) and os.path.exists(orc)):
        import __graft_entry__
        __graft_entry__.build()


_ensure_built()


def soa_to_state(soa):
    """BodySoA -> oracle state dict (copies of the first `count` entries)."""
    n = soa.count
    m = {"pos_x": "px", "pos_y": "py", "pos_z": "pz", "vel_x": "vx", "vel_y": "vy",
         "vel_z": "vz", "acc_x": "ax", "acc_y": "ay", "acc_z": "az",
         "old_acc_x": "oax", "old_acc_y": "oay", "old_acc_z": "oaz",
         "mass": "mass", "radius": "radius", "density": "density",
         "is_active": "active", "is_collidable": "collidable", "body_type": "body_type"}
    return {v: np.array(getattr(soa, k)[:n], copy=True) for k, v in m.items()}


@pytest.fixture
def plummer():
    import cmhip

    def make(n, seed=42, cut=1.0, com=True):
        return cmhip.generate_plummer(n, 1.0, 1.0, seed=seed, mass_fraction_cut=cut,
                                      com_frame=com)
    return make


def pair_collapse_state(n_pairs=10000, v=50.0, dt=1e-3, delta=1e-9, seed=31):
    """Pairs of light bodies that approach each other at speed v: after one
    drift of dt their separation is delta (coincident at 21-bit key
    resolution, so each pair hangs a long single-child chain off the octree),
    after the second drift they have passed each other and are 2 v dt - delta
    apart.  Masses are 1e-12, so the forces change neither drift measurably."""
    import cmhip
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1.0, 1.0, (n_pairs, 3))
    u = rng.normal(size=(n_pairs, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    d0 = 2 * v * dt + delta
    p = np.concatenate([c - 0.5 * d0 * u, c + 0.5 * d0 * u])
    w = np.concatenate([v * u, -v * u])
    n = 2 * n_pairs
    return cmhip.BodySoA.from_arrays(pos_x=p[:, 0], pos_y=p[:, 1], pos_z=p[:, 2],
                                     vel_x=w[:, 0], vel_y=w[:, 1], vel_z=w[:, 2],
                                     mass=np.full(n, 1e-12), is_active=np.ones(n))
