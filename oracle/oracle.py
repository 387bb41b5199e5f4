"""ctypes front end of the CPU oracle (liboracle_cm.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() to
check its result and by bench.py's cpu_baseline leg.  The product (libcmhip.so and the
cmhip package) never imports it.

Each wrapper returns numpy arrays; see cm_oracle.c for the reference citations
of every restated function.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_cm.so")
_lib = None

f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


class BHStats(C.Structure):
    _fields_ = [("node_count", C.c_int64), ("interactions", C.c_int64),
                ("build_ms", C.c_double), ("traverse_ms", C.c_double)]


class Energy(C.Structure):
    _fields_ = [(k, C.c_double) for k in
                ("ke", "pe", "px", "py", "pz", "lx", "ly", "lz", "l_mag")]


class ForceErr(C.Structure):
    _fields_ = [("mean", C.c_double), ("max", C.c_double), ("rms", C.c_double),
                ("median", C.c_double), ("count", C.c_int32)]


class Event(C.Structure):
    _fields_ = [("a", C.c_int32), ("b", C.c_int32), ("overlap", C.c_double),
                ("dist", C.c_double), ("nx", C.c_double), ("ny", C.c_double),
                ("nz", C.c_double), ("rel_normal_speed", C.c_double)]


EVENT_DTYPE = np.dtype([("a", "<i4"), ("b", "<i4"), ("overlap", "<f8"),
                        ("dist", "<f8"), ("nx", "<f8"), ("ny", "<f8"),
                        ("nz", "<f8"), ("rel_normal_speed", "<f8")])


def build():
    subprocess.run(["make", "-B", "-C", _HERE], check=True,
                   stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    I, D, V = C.c_int, C.c_double, None
    L.cmo_gravity_coeff.argtypes = [D, D, D, D, I]
    L.cmo_gravity_coeff.restype = D
    L.cmo_direct_n3.argtypes = [I, f64p, f64p, f64p, f64p, u8p, D, f64p, f64p, f64p]
    L.cmo_direct_n3.restype = V
    L.cmo_direct_gk.argtypes = [I, f64p, f64p, f64p, f64p, f64p, u8p, D, I,
                                f64p, f64p, f64p]
    L.cmo_direct_gk.restype = V
    L.cmo_direct_targets.argtypes = [I, f64p, f64p, f64p, f64p, u8p, D, I, i32p,
                                     f64p, f64p, f64p, I]
    L.cmo_direct_targets.restype = V
    L.cmo_bh_csharp.argtypes = [I, f64p, f64p, f64p, f64p, f64p, u8p, D, D, I, D,
                                I, f64p, f64p, f64p, C.POINTER(BHStats)]
    L.cmo_bh_csharp.restype = I
    L.cmo_verlet_kick_drift.argtypes = [I, u8p, D, f64p, f64p, f64p, f64p, f64p,
                                        f64p, f64p, f64p, f64p]
    L.cmo_verlet_kick_drift.restype = V
    L.cmo_verlet_kick_rotate.argtypes = [I, u8p, D, f64p, f64p, f64p, f64p, f64p,
                                         f64p, f64p, f64p, f64p]
    L.cmo_verlet_kick_rotate.restype = V
    L.cmo_root_box.argtypes = [I, f64p, f64p, f64p, u8p, f64p]
    L.cmo_root_box.restype = V
    L.cmo_morton_keys.argtypes = [I, f64p, f64p, f64p, u8p, f64p, I, u64p]
    L.cmo_morton_keys.restype = V
    L.cmo_stable_sort_keys.argtypes = [I, u64p, u64p, i32p]
    L.cmo_stable_sort_keys.restype = V
    L.cmo_tree_build.argtypes = [I, f64p, f64p, f64p, f64p, u8p, I, I, I, f64p,
                                 C.POINTER(C.c_int32), u64p, i32p, i32p, i32p,
                                 i32p, i32p, i32p, f64p, f64p, i32p]
    L.cmo_tree_build.restype = I
    L.cmo_bh_tree.argtypes = [I, f64p, f64p, f64p, f64p, f64p, u8p, D, D, I, I,
                              I, I, I, I, I, D, f64p, f64p, f64p, i32p, i32p,
                              C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
    L.cmo_bh_potential.argtypes = [I, f64p, f64p, f64p, f64p, u8p, D, D, I, I, I, I, I, I,
                                   f64p, i32p, i32p]
    L.cmo_bh_potential.restype = I
    L.cmo_bh_potential_reference.argtypes = [I, f64p, f64p, f64p, f64p, u8p, D, D, I, f64p]
    L.cmo_bh_potential_reference.restype = I
    L.cmo_far_coeffs.argtypes = [D, D, D, D, D, f64p]
    L.cmo_far_coeffs.restype = V
    L.cmo_far_eval.argtypes = [f64p, D, D, D, f64p]
    L.cmo_far_eval.restype = V
    L.cmo_bh_tree.restype = I
    L.cmo_energy_naive.argtypes = [I, f64p, f64p, f64p, f64p, f64p, f64p, f64p,
                                   u8p, D, I, C.POINTER(Energy)]
    L.cmo_energy_naive.restype = V
    L.cmo_energy_exact.argtypes = [I, f64p, f64p, f64p, f64p, f64p, f64p, f64p,
                                   u8p, D, I, I, C.POINTER(Energy)]
    L.cmo_energy_exact.restype = V
    L.cmo_dotnet_random.argtypes = [C.c_int32, I, i32p]
    L.cmo_dotnet_random.restype = V
    L.cmo_plummer_dotnet.argtypes = [I, D, D, C.c_int32] + [f64p] * 7
    L.cmo_plummer_dotnet.restype = V
    L.cmo_center_of_mass.argtypes = [I, f64p, f64p, f64p, f64p, u8p, f64p]
    L.cmo_center_of_mass.restype = V
    L.cmo_max_acceleration.argtypes = [I, f64p, f64p, f64p, u8p]
    L.cmo_max_acceleration.restype = D
    L.cmo_force_error.argtypes = [I, C.c_void_p, f64p, f64p, f64p, f64p, f64p,
                                  f64p, C.POINTER(ForceErr)]
    L.cmo_force_error.restype = V
    L.cmo_detect.argtypes = [I, f64p, f64p, f64p, f64p, f64p, f64p, f64p, f64p,
                             u8p, u8p, C.c_void_p, I]
    L.cmo_detect.restype = I
    L.cmo_cbrt.argtypes = [D]
    L.cmo_cbrt.restype = D
    L.cmo_body_radius.argtypes = [D, D, C.c_int32]
    L.cmo_body_radius.restype = D
    L.cmo_resolve_merge.argtypes = [I, C.c_void_p, I] + [f64p] * 15 + [u8p, u8p, i32p]
    L.cmo_resolve_merge.restype = I
    L.cmo_c_sim.argtypes = []
    L.cmo_c_sim.restype = D
    L.cmo_pn_apply.argtypes = [I] + [f64p] * 7 + [u8p, D, D, D, f64p, f64p, f64p,
                                                  i32p, i32p, f64p, f64p, I]
    L.cmo_pn_apply.restype = I
    L.cmo_direct_softened.argtypes = [I, f64p, f64p, f64p, f64p, u8p, i32p, I, f64p,
                                      D, D, D, f64p, f64p, f64p]
    L.cmo_direct_softened.restype = V
    L.cmo_yoshida_coefficients.argtypes = [f64p]
    L.cmo_yoshida_coefficients.restype = V
    L.cmo_collision_energy.argtypes = [D, D, D, D, D, C.POINTER(C.c_int32),
                                       C.POINTER(C.c_int32), C.POINTER(D), C.POINTER(D)]
    L.cmo_collision_energy.restype = V
    L.cmo_resolve_events.argtypes = ([I, C.c_void_p, I] + [f64p] * 15 + [u8p, u8p, i32p]
                                     + [I, D, D, I, I, I, i32p])
    L.cmo_resolve_events.restype = I
    _lib = L
    return L


# ---------------------------------------------------------------------------
# state helpers: a "state" is a dict of numpy arrays keyed like BodySoA fields
# ---------------------------------------------------------------------------
F64_FIELDS = ("px", "py", "pz", "vx", "vy", "vz", "ax", "ay", "az",
              "oax", "oay", "oaz", "mass", "radius", "density")
U8_FIELDS = ("active", "collidable")
I32_FIELDS = ("body_type",)


def copy_state(s):
    return {k: np.array(v, copy=True) for k, v in s.items()}


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def gravity_coeff(dist_sq, m, r, eps2, shell):
    return lib().cmo_gravity_coeff(dist_sq, m, r, eps2, int(shell))


def direct_n3(s, eps):
    n = len(s["px"])
    ax, ay, az = (np.zeros(n) for _ in range(3))
    lib().cmo_direct_n3(n, _c(s["px"], np.float64), _c(s["py"], np.float64),
                        _c(s["pz"], np.float64), _c(s["mass"], np.float64),
                        _c(s["active"], np.uint8), eps, ax, ay, az)
    return ax, ay, az


def direct_gk(s, eps, shell=False):
    n = len(s["px"])
    ax, ay, az = (np.zeros(n) for _ in range(3))
    lib().cmo_direct_gk(n, _c(s["px"], np.float64), _c(s["py"], np.float64),
                        _c(s["pz"], np.float64), _c(s["mass"], np.float64),
                        _c(s["radius"], np.float64), _c(s["active"], np.uint8),
                        eps, int(shell), ax, ay, az)
    return ax, ay, az


def direct_targets(s, eps, targets, nthreads=0):
    t = _c(targets, np.int32)
    ax, ay, az = (np.zeros(len(t)) for _ in range(3))
    n = len(s["px"])
    lib().cmo_direct_targets(n, _c(s["px"], np.float64), _c(s["py"], np.float64),
                             _c(s["pz"], np.float64), _c(s["mass"], np.float64),
                             _c(s["active"], np.uint8), eps, len(t), t,
                             ax, ay, az, nthreads)
    return ax, ay, az


def bh_csharp(s, eps, theta, shell=False, min_node_size=1e-12, nthreads=0):
    n = len(s["px"])
    ax, ay, az = (np.zeros(n) for _ in range(3))
    st = BHStats()
    rc = lib().cmo_bh_csharp(n, _c(s["px"], np.float64), _c(s["py"], np.float64),
                             _c(s["pz"], np.float64), _c(s["mass"], np.float64),
                             _c(s["radius"], np.float64), _c(s["active"], np.uint8),
                             eps, theta, int(shell), min_node_size, nthreads,
                             ax, ay, az, C.byref(st))
    if rc != 0:
        raise MemoryError("oracle BH pool allocation failed")
    return (ax, ay, az), st


def verlet_step(s, dt, force_fn):
    """SoAVerletIntegrator.Step (Integrators/SoAVerletIntegrator.cs:48-115) on a
    state dict in place; force_fn(state) -> (ax, ay, az)."""
    L = lib()
    n = len(s["px"])
    L.cmo_verlet_kick_drift(n, s["active"], dt, s["px"], s["py"], s["pz"],
                            s["vx"], s["vy"], s["vz"], s["oax"], s["oay"], s["oaz"])
    ax, ay, az = force_fn(s)
    s["ax"][:] = ax
    s["ay"][:] = ay
    s["az"][:] = az
    L.cmo_verlet_kick_rotate(n, s["active"], dt, s["vx"], s["vy"], s["vz"],
                             s["ax"], s["ay"], s["az"], s["oax"], s["oay"], s["oaz"])


def root_box(s):
    n = len(s["px"])
    box = np.zeros(4)
    lib().cmo_root_box(n, _c(s["px"], np.float64), _c(s["py"], np.float64),
                       _c(s["pz"], np.float64), _c(s["active"], np.uint8), box)
    return box


def morton_keys(s, box, levels):
    n = len(s["px"])
    keys = np.zeros(n, dtype=np.uint64)
    lib().cmo_morton_keys(n, _c(s["px"], np.float64), _c(s["py"], np.float64),
                          _c(s["pz"], np.float64), _c(s["active"], np.uint8),
                          _c(box, np.float64), levels, keys)
    return keys


def stable_sort_keys(keys):
    n = len(keys)
    ks = np.zeros(n, dtype=np.uint64)
    order = np.zeros(n, dtype=np.int32)
    lib().cmo_stable_sort_keys(n, _c(keys, np.uint64), ks, order)
    return ks, order


def tree_build(s, levels, leaf_cap, node_cap=None):
    n = len(s["px"])
    cap = node_cap or 2 * n + 16
    box = np.zeros(4)
    na = C.c_int32(0)
    ks = np.zeros(n, dtype=np.uint64)
    order = np.zeros(n, dtype=np.int32)
    arrs = [np.zeros(cap, dtype=np.int32) for _ in range(5)]
    mass = np.zeros(cap)
    com = np.zeros(3 * cap)
    level_start = np.zeros(levels + 3, dtype=np.int32)
    nn = lib().cmo_tree_build(n, _c(s["px"], np.float64), _c(s["py"], np.float64),
                              _c(s["pz"], np.float64), _c(s["mass"], np.float64),
                              _c(s["active"], np.uint8), levels, leaf_cap, cap, box,
                              C.byref(na), ks, order, *arrs, mass, com, level_start)
    if nn < 0:
        raise RuntimeError("node capacity exceeded")
    return {
        "box": box, "n_active": na.value, "keys": ks, "order": order,
        "level": arrs[0][:nn], "start": arrs[1][:nn], "count": arrs[2][:nn],
        "first": arrs[3][:nn], "nchild": arrs[4][:nn], "mass": mass[:nn],
        "com": com[:3 * nn].reshape(nn, 3), "level_start": level_start,
        "n_nodes": nn,
    }


def bh_tree(s, eps, theta, levels=10, leaf_cap=16, group_size=64, shell=False,
            split_depth=-3, lifo=1536, nthreads=0, far_ratio=0.0):
    n = len(s["px"])
    ax, ay, az = (np.zeros(n) for _ in range(3))
    nc = np.zeros(n, dtype=np.int32)
    npart = np.zeros(n, dtype=np.int32)
    mq = C.c_int32(0)
    nfar = C.c_int64(0)
    rc = lib().cmo_bh_tree(n, _c(s["px"], np.float64), _c(s["py"], np.float64),
                           _c(s["pz"], np.float64), _c(s["mass"], np.float64),
                           _c(s["radius"], np.float64), _c(s["active"], np.uint8),
                           eps, theta, int(shell), levels, leaf_cap, group_size,
                           split_depth, int(lifo), nthreads, float(far_ratio), ax, ay, az,
                           nc, npart, C.byref(mq), C.byref(nfar))
    if rc < 0:
        raise RuntimeError("oracle target groups do not partition the active bodies")
    return (ax, ay, az), {"n_cells": nc, "n_parts": npart, "max_queue": mq.value,
                          "far_cells": nfar.value}


def bh_potential(s, eps, theta, levels=10, leaf_cap=16, group_size=64, split_depth=-3,
                 lifo=1536, nthreads=0):
    """Per-body shares 1/2 m_i phi_i of the Barnes-Hut potential energy through
    the group traversal, with the per-body term counts."""
    n = len(s["px"])
    rows = np.zeros(n)
    nc = np.zeros(n, dtype=np.int32)
    npart = np.zeros(n, dtype=np.int32)
    rc = lib().cmo_bh_potential(n, _c(s["px"], np.float64), _c(s["py"], np.float64),
                                _c(s["pz"], np.float64), _c(s["mass"], np.float64),
                                _c(s["active"], np.uint8), eps, theta, levels, leaf_cap,
                                group_size, split_depth, int(lifo), nthreads, rows, nc, npart)
    if rc < 0:
        raise RuntimeError("oracle target groups do not partition the active bodies")
    return rows, {"n_cells": nc, "n_parts": npart}


def bh_potential_reference(s, eps, theta, nthreads=0):
    """compute_potential_bh_kernel's per-body traversal over the C# octree."""
    n = len(s["px"])
    rows = np.zeros(n)
    lib().cmo_bh_potential_reference(n, _c(s["px"], np.float64), _c(s["py"], np.float64),
                                     _c(s["pz"], np.float64), _c(s["mass"], np.float64),
                                     _c(s["active"], np.uint8), eps, theta, nthreads, rows)
    return rows


def far_coeffs(M, r, eps2, t=None):
    """Adds the far-field Taylor coefficients of one cell to t (19 values)."""
    t = np.zeros(19) if t is None else t
    lib().cmo_far_coeffs(float(M), float(r[0]), float(r[1]), float(r[2]), float(eps2), t)
    return t


def far_eval(t, d):
    out = np.zeros(3)
    lib().cmo_far_eval(_c(t, np.float64), float(d[0]), float(d[1]), float(d[2]), out)
    return out


def energy_naive(s, eps, with_pe=True):
    e = Energy()
    n = len(s["px"])
    lib().cmo_energy_naive(n, *(_c(s[k], np.float64) for k in
                                ("px", "py", "pz", "vx", "vy", "vz", "mass")),
                           _c(s["active"], np.uint8), eps, int(with_pe), C.byref(e))
    return {k: getattr(e, k) for k, _ in Energy._fields_}


def energy_exact(s, eps, with_pe=True, nthreads=0):
    e = Energy()
    n = len(s["px"])
    lib().cmo_energy_exact(n, *(_c(s[k], np.float64) for k in
                                ("px", "py", "pz", "vx", "vy", "vz", "mass")),
                           _c(s["active"], np.uint8), eps, int(with_pe), nthreads,
                           C.byref(e))
    return {k: getattr(e, k) for k, _ in Energy._fields_}


def center_of_mass(s):
    """(cx, cy, cz, total mass) of the active bodies (gpu_reduce_com)."""
    n = len(s["px"])
    out = np.zeros(4)
    lib().cmo_center_of_mass(n, *(_c(s[k], np.float64) for k in ("px", "py", "pz", "mass")),
                             _c(s["active"], np.uint8), out)
    return tuple(float(x) for x in out)


def max_acceleration(acc, active):
    """Largest |a| over active bodies (Engine::compute_max_acceleration)."""
    n = len(acc[0])
    return lib().cmo_max_acceleration(n, *(_c(a, np.float64) for a in acc), _c(active, np.uint8))


def dotnet_random(seed, count):
    """The first `count` raw draws of .NET's System.Random(seed)."""
    out = np.zeros(count, np.int32)
    lib().cmo_dotnet_random(seed, count, out)
    return out


def plummer_dotnet(n, total_mass=100.0, scale_length=2.0, seed=42):
    """PlummerSphere.Generate(n, totalMass, scaleLength, seed): dict of
    px, py, pz, vx, vy, vz, mass."""
    out = {k: np.zeros(n) for k in ("px", "py", "pz", "vx", "vy", "vz", "mass")}
    lib().cmo_plummer_dotnet(n, total_mass, scale_length, seed,
                             *(out[k] for k in ("px", "py", "pz", "vx", "vy", "vz", "mass")))
    return out


def force_error(exact, test, active=None):
    r = ForceErr()
    n = len(exact[0])
    act = None if active is None else _c(active, np.uint8)
    lib().cmo_force_error(n, None if act is None else act.ctypes.data,
                          *(_c(a, np.float64) for a in exact),
                          *(_c(a, np.float64) for a in test), C.byref(r))
    return {"mean": r.mean, "max": r.max, "rms": r.rms, "median": r.median,
            "count": r.count}


def detect(s, cap=1 << 20):
    n = len(s["px"])
    out = np.zeros(cap, dtype=EVENT_DTYPE)
    ne = lib().cmo_detect(n, *(_c(s[k], np.float64) for k in
                               ("px", "py", "pz", "vx", "vy", "vz", "mass", "radius")),
                          _c(s["active"], np.uint8), _c(s["collidable"], np.uint8),
                          out.ctypes.data, cap)
    if ne > cap:
        raise RuntimeError("event capacity exceeded")
    return out[:ne].copy()


def cbrt(x):
    return lib().cmo_cbrt(float(x))


def body_radius(mass, density, body_type):
    return lib().cmo_body_radius(float(mass), float(density), int(body_type))


def resolve_merge(events, s):
    """In-place MergeOnly resolution; returns merge count."""
    ev = np.ascontiguousarray(events, dtype=EVENT_DTYPE)
    n = len(s["px"])
    return lib().cmo_resolve_merge(len(ev), ev.ctypes.data, n,
                                   *(s[k] for k in F64_FIELDS),
                                   s["active"], s["collidable"], s["body_type"])


def compact(s):
    """Order-preserving compaction of a state dict; returns (state, old_index)."""
    act = s["active"].astype(bool)
    idx = np.nonzero(act)[0].astype(np.int32)
    out = {k: np.ascontiguousarray(v[idx]) for k, v in s.items()}
    return out, idx


# ---------------------------------------------------------------------------
# widening rows: 1PN decorator, softening modes, Yoshida-4 (cm_oracle_ext.c)
# ---------------------------------------------------------------------------
def c_sim():
    """PhysicalConstants.C_Sim (src/CelestialMechanics.Math/PhysicalConstants.cs:37)."""
    return lib().cmo_c_sim()


def pn_apply(s, eps, ax, ay, az, max_vel_frac=0.3, warn_factor=3.0, warn_cap=1 << 16):
    """PostNewtonian1Correction.ApplyCorrections on copies of (ax, ay, az);
    returns ((ax, ay, az), warnings) with warnings a list of
    (i, j, dist, rs) in (i, j) order."""
    n = len(s["px"])
    out = [np.array(a, dtype=np.float64, copy=True) for a in (ax, ay, az)]
    wi = np.zeros(warn_cap, dtype=np.int32)
    wj = np.zeros(warn_cap, dtype=np.int32)
    wd = np.zeros(warn_cap)
    wr = np.zeros(warn_cap)
    nw = lib().cmo_pn_apply(n, *(_c(s[k], np.float64) for k in
                                 ("px", "py", "pz", "vx", "vy", "vz", "mass")),
                            _c(s["active"], np.uint8), eps, max_vel_frac, warn_factor,
                            *out, wi, wj, wd, wr, warn_cap)
    k = min(nw, warn_cap)
    warns = [(int(wi[q]), int(wj[q]), float(wd[q]), float(wr[q])) for q in range(k)]
    return tuple(out), warns, nw


def direct_softened(s, mode, table=None, scale=0.01, mn=1e-8, global_eps=1e-4):
    """Engine::compute_cpu_brute_forces with SofteningMode `mode`."""
    n = len(s["px"])
    t = np.zeros(16) if table is None else _c(table, np.float64)
    ax, ay, az = (np.zeros(n) for _ in range(3))
    lib().cmo_direct_softened(n, *(_c(s[k], np.float64) for k in ("px", "py", "pz", "mass")),
                              _c(s["active"], np.uint8), _c(s["body_type"], np.int32),
                              int(mode), t, scale, mn, global_eps, ax, ay, az)
    return ax, ay, az


def yoshida_coefficients():
    w = np.zeros(3)
    lib().cmo_yoshida_coefficients(w)
    return w


def yoshida_step(s, dt, force_fn):
    """One Yoshida-4 step (engine/src/sim/engine.cpp:1057-1066, sub-step
    :685-745 without collisions): three Verlet sub-steps of W1 dt, W0 dt, W1 dt."""
    for w in yoshida_coefficients():
        verlet_step(s, w * dt, force_fn)


def collision_energy(m1, m2, eff_radius_au, rel_speed, max_loss):
    """CollisionEnergyModel.Evaluate: the fields that change the state."""
    im, fr = C.c_int32(), C.c_int32()
    ej, ratio = C.c_double(), C.c_double()
    lib().cmo_collision_energy(float(m1), float(m2), float(eff_radius_au), float(rel_speed),
                               float(max_loss), C.byref(im), C.byref(fr), C.byref(ej),
                               C.byref(ratio))
    return {"is_merge": bool(im.value), "is_fragmentation": bool(fr.value),
            "is_catastrophic": not im.value and not fr.value, "ejecta_kg": ej.value,
            "ratio": ratio.value}


def resolve_events(events, s, mode=0, restitution=0.15, mass_loss_cap=0.3, promote=False,
                   max_per_frame=0, max_per_body=0):
    """CollisionResolver.Resolve (every mode) in place; returns the number of
    events resolved."""
    ev = np.ascontiguousarray(events, dtype=EVENT_DTYPE)
    n = len(s["px"])
    scratch = np.zeros(max(n, 1), dtype=np.int32)
    return lib().cmo_resolve_events(len(ev), ev.ctypes.data, n,
                                    *(s[k] for k in F64_FIELDS),
                                    s["active"], s["collidable"], s["body_type"],
                                    int(mode), float(restitution), float(mass_loss_cap),
                                    int(bool(promote)), int(max_per_frame), int(max_per_body),
                                    scratch)
