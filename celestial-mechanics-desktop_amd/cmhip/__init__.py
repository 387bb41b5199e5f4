"""Python host side of libcmhip.so, mirroring the reference's C# physics surface.

Class and method names follow src/CelestialMechanics.Physics of
mfscpayload-690/celestial-mechanics-desktop so that tests read like the
reference's own:

  BodySoA                 SoA/BodySoA.cs:39-175
  HipPhysicsBackend       IPhysicsComputeBackend (Solvers/IPhysicsComputeBackend.cs:21-33)
                          + IGravityModelAwareBackend (Solvers/IGravityModelAwareBackend.cs:6-9),
                          in the slot of CudaPhysicsBackend (Solvers/CudaPhysicsBackend.cs:35-48)
  SoAVerletIntegrator     Integrators/SoAVerletIntegrator.cs:48-115
  EnergyCalculator        Validation/EnergyCalculator.cs:16-81
  CollisionDetector       Collisions/CollisionDetector.cs:35-135
  CollisionResolver       Collisions/CollisionResolver.cs:58-345 (MergeOnly)

Every computation runs in libcmhip.so on the GPU.  There is no CPU fallback:
if the library is missing or no GPU is present the calls raise.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_DIR, "libcmhip.so")

# ---- constants of include/cmhip.h ----------------------------------------
CM_OK = 0
CM_ERR_HIP, CM_ERR_OUT_OF_MEMORY, CM_ERR_INVALID_ARGUMENT = -1, -2, -3
CM_ERR_POOL_EXHAUSTED, CM_ERR_DEVICE_NOT_FOUND, CM_ERR_NOT_INITIALIZED = -4, -5, -6
ERRORS = {-1: "HipError", -2: "OutOfMemory", -3: "InvalidArgument",
          -4: "PoolExhausted", -5: "DeviceNotFound", -6: "NotInitialized"}
BACKEND_DIRECT_PAIRWISE = 0
BACKEND_DIRECT = 1
BACKEND_BARNES_HUT = 2
PREC_FP64 = 0
PREC_FP64_FAST = 1
PREC_FP32_SPLIT = 2

F64_FIELDS = ("pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z",
              "acc_x", "acc_y", "acc_z", "old_acc_x", "old_acc_y", "old_acc_z",
              "mass", "radius", "density")
U8_FIELDS = ("is_active", "is_collidable")
I32_FIELDS = ("body_type",)
ALL_FIELDS = F64_FIELDS + U8_FIELDS + I32_FIELDS

# Names exported by libcmhip.so, as declared in include/cmhip.h.
EXPORTS = ("cm_create", "cm_destroy", "cm_last_error", "cm_configure",
           "cm_default_config", "cm_upload_soa", "cm_download_soa",
           "cm_get_count", "cm_compute_forces", "cm_verlet_step", "cm_energy",
           "cm_potential_rows", "cm_max_acceleration",
           "cm_direct_sample", "cm_get_force_stats", "cm_detect_collisions",
           "cm_resolve_merges", "cm_compact", "cm_generate_plummer",
           "cm_generate_plummer_dotnet",
           "cm_reset_force_stats", "cm_generate_disk", "cm_ensemble_create",
           "cm_ensemble_destroy", "cm_ensemble_last_error", "cm_ensemble_configure",
           "cm_ensemble_upload", "cm_ensemble_download", "cm_ensemble_step",
           "cm_ensemble_energy", "cm_ensemble_get_stats", "cm_generate_ensemble",
           "cm_mlp_init", "cm_mlp_set_params", "cm_mlp_get_params", "cm_mlp_forward",
           "cm_mlp_gradient", "cm_mlp_collect", "cm_mlp_get_pairs", "cm_mlp_train",
           "cm_default_pn_config", "cm_set_post_newtonian", "cm_pn_warnings",
           "cm_default_softening_config", "cm_set_softening", "cm_yoshida4_step",
           "cm_default_collision_config", "cm_configure_collisions", "cm_resolve_collisions",
           "cm_nbody_step", "cm_energy_noalloc", "cm_compare_forces", "cm_nbody_step_diag",
           "cm_upload_force_inputs")

COLLISION_MERGE_ONLY, COLLISION_BOUNCE_ONLY, COLLISION_REALISTIC = 0, 1, 2

SOFTENING_GLOBAL, SOFTENING_PER_TYPE, SOFTENING_ADAPTIVE = 0, 1, 2

MLP_IN, MLP_OUT, MLP_USED, MLP_PARAMS = 16, 16, 8, 10448

# 3 g/cm^3 in Msun / AU^3 (AU = 1.496e11 m, Msun = 1.989e30 kg), and the same
# density with radii inflated 100x (density / 100^3)
ROCK_DENSITY_SIM = 3000.0 * 1.496e11 ** 3 / 1.989e30
ROCK_DENSITY_INFLATED = ROCK_DENSITY_SIM / 100.0 ** 3


class CmError(RuntimeError):
    def __init__(self, code, message=""):
        self.code = code
        super().__init__(f"{ERRORS.get(code, code)}: {message}")


class BackendConfig(C.Structure):
    _fields_ = [("kind", C.c_int32), ("precision", C.c_int32),
                ("shell_theorem", C.c_int32), ("morton_levels", C.c_int32),
                ("theta", C.c_double), ("leaf_capacity", C.c_int32),
                ("group_size", C.c_int32), ("learned_correction", C.c_int32),
                ("far_field_ratio", C.c_double)]


class PnConfig(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("max_velocity_fraction_c", C.c_double),
                ("schwarzschild_warning_factor", C.c_double)]


class PnWarning(C.Structure):
    _fields_ = [("body", C.c_int32), ("other", C.c_int32), ("distance", C.c_double),
                ("schwarzschild_radius", C.c_double)]


class SofteningConfig(C.Structure):
    _fields_ = [("mode", C.c_int32), ("type_softening", C.c_double * 16),
                ("adaptive_scale", C.c_double), ("adaptive_min", C.c_double)]


class CollisionConfig(C.Structure):
    _fields_ = [("mode", C.c_int32), ("default_restitution", C.c_double),
                ("fragmentation_specific_energy_threshold", C.c_double),
                ("fragmentation_mass_loss_cap", C.c_double),
                ("capture_velocity_factor", C.c_double),
                ("promote_compact_remnants", C.c_int32),
                ("max_merges_per_frame", C.c_int32), ("max_merges_per_body", C.c_int32)]


class SoAPtrs(C.Structure):
    _fields_ = ([(f, C.POINTER(C.c_double)) for f in F64_FIELDS]
                + [(f, C.POINTER(C.c_uint8)) for f in U8_FIELDS]
                + [("body_type", C.POINTER(C.c_int32))])


class ForceStats(C.Structure):
    _fields_ = [("cell_interactions", C.c_int64), ("body_interactions", C.c_int64),
                ("n_active", C.c_int32), ("n_nodes", C.c_int32),
                ("n_groups", C.c_int32), ("flags", C.c_int32),
                ("ms_tree", C.c_float), ("ms_sort", C.c_float),
                ("ms_force", C.c_float), ("ms_total", C.c_float),
                ("acc_interactions", C.c_int64), ("acc_ms_force", C.c_double),
                ("acc_ms_total", C.c_double), ("acc_evaluations", C.c_int32),
                ("far_cells", C.c_int64), ("acc_far_cells", C.c_int64),
                ("acc_timed_evaluations", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ForceComparison(C.Structure):
    """ForceErrorAnalyzer.ComparisonResult (Validation/ForceErrorAnalyzer.cs:67-135)."""
    _fields_ = [("mean_relative_error", C.c_double), ("max_relative_error", C.c_double),
                ("rms_relative_error", C.c_double), ("median_relative_error", C.c_double),
                ("body_count", C.c_int32), ("theta", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SimulationState(C.Structure):
    """Types/SimulationState.cs:9-25, as NBodySolver.Step returns it."""
    _fields_ = [("time", C.c_double), ("body_count", C.c_int32),
                ("active_body_count", C.c_int32), ("kinetic_energy", C.c_double),
                ("potential_energy", C.c_double), ("total_energy", C.c_double),
                ("total_momentum", C.c_double * 3), ("energy_drift", C.c_double),
                ("collision_count", C.c_int32), ("current_dt", C.c_double),
                ("step_energy_delta", C.c_double), ("step_momentum_delta", C.c_double),
                ("step_angular_momentum_delta", C.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["total_momentum"] = tuple(self.total_momentum)
        return d


class EnergyReport(C.Structure):
    _fields_ = [("kinetic", C.c_double), ("potential", C.c_double),
                ("momentum", C.c_double * 3), ("angular_momentum", C.c_double * 3),
                ("angular_momentum_mag", C.c_double), ("total_mass", C.c_double),
                ("active_count", C.c_int32), ("center_of_mass", C.c_double * 3)]


class EnsembleStats(C.Structure):
    _fields_ = [("threads", C.c_int32), ("tuned_threads", C.c_int32),
                ("tuned_cost_order", C.c_int32), ("tuned_ms_per_step", C.c_float),
                ("last_ms", C.c_float), ("last_steps", C.c_int32),
                ("interactions_total", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


EVENT_DTYPE = np.dtype([("body_a", "<i4"), ("body_b", "<i4"),
                        ("overlap_depth", "<f8"), ("distance", "<f8"),
                        ("normal_x", "<f8"), ("normal_y", "<f8"),
                        ("normal_z", "<f8"), ("relative_normal_speed", "<f8")])

_lib = None


def load_library():
    """Load libcmhip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CmError(-6, f"{LIB_PATH} not built; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    vp, i32, f64 = C.c_void_p, C.c_int32, C.c_double
    sig = {
        "cm_create": ([i32, i32, C.POINTER(vp)], i32),
        "cm_destroy": ([vp], i32),
        "cm_last_error": ([vp], C.c_char_p),
        "cm_configure": ([vp, C.POINTER(BackendConfig)], i32),
        "cm_default_config": ([C.POINTER(BackendConfig)], i32),
        "cm_upload_soa": ([vp, C.POINTER(SoAPtrs), i32], i32),
        "cm_download_soa": ([vp, C.POINTER(SoAPtrs), i32], i32),
        "cm_upload_force_inputs": ([vp, C.POINTER(SoAPtrs), i32], i32),
        "cm_get_count": ([vp, C.POINTER(i32)], i32),
        "cm_compute_forces": ([vp, f64], i32),
        "cm_verlet_step": ([vp, f64, f64, i32], i32),
        "cm_energy": ([vp, f64, i32, C.POINTER(EnergyReport)], i32),
        "cm_energy_noalloc": ([vp, f64, i32, C.POINTER(EnergyReport)], i32),
        "cm_potential_rows": ([vp, f64, i32, C.POINTER(C.c_double)], i32),
        "cm_max_acceleration": ([vp, C.POINTER(C.c_double)], i32),
        "cm_direct_sample": ([vp, f64, vp, i32, vp, vp, vp], i32),
        "cm_get_force_stats": ([vp, C.POINTER(ForceStats)], i32),
        "cm_compare_forces": ([vp, f64, f64, vp, i32, C.POINTER(ForceComparison)], i32),
        "cm_reset_force_stats": ([vp], i32),
        "cm_detect_collisions": ([vp, vp, i32, C.POINTER(i32)], i32),
        "cm_resolve_merges": ([vp, C.POINTER(i32)], i32),
        "cm_compact": ([vp, C.POINTER(i32), vp], i32),
        "cm_generate_plummer": ([i32, f64, f64, C.c_uint64, f64, i32,
                                 C.POINTER(SoAPtrs)], i32),
        "cm_generate_plummer_dotnet": ([i32, f64, f64, i32, C.POINTER(SoAPtrs)], i32),
        "cm_generate_disk": ([i32, C.c_uint64, f64, f64, f64, f64, f64, f64, f64, f64,
                              C.POINTER(SoAPtrs)], i32),
        "cm_ensemble_create": ([i32, vp, i32, C.POINTER(vp)], i32),
        "cm_ensemble_destroy": ([vp], i32),
        "cm_ensemble_last_error": ([vp], C.c_char_p),
        "cm_ensemble_configure": ([vp, i32, i32, i32], i32),
        "cm_ensemble_upload": ([vp, C.POINTER(SoAPtrs)], i32),
        "cm_ensemble_download": ([vp, C.POINTER(SoAPtrs)], i32),
        "cm_ensemble_step": ([vp, f64, f64, i32], i32),
        "cm_ensemble_energy": ([vp, f64, vp], i32),
        "cm_ensemble_get_stats": ([vp, C.POINTER(EnsembleStats)], i32),
        "cm_generate_ensemble": ([i32, C.c_uint64, i32, i32, f64, f64, f64, f64, f64, vp,
                                  C.POINTER(SoAPtrs)], i32),
        "cm_mlp_init": ([vp, C.c_uint64], i32),
        "cm_mlp_set_params": ([vp, vp, i32], i32),
        "cm_mlp_get_params": ([vp, vp, i32], i32),
        "cm_mlp_forward": ([vp, vp, i32, vp], i32),
        "cm_mlp_gradient": ([vp, vp, vp, i32, vp, C.POINTER(C.c_float)], i32),
        "cm_mlp_collect": ([vp, f64, i32, C.c_uint64, i32, C.POINTER(i32)], i32),
        "cm_mlp_get_pairs": ([vp, vp, vp, vp, vp, i32], i32),
        "cm_mlp_train": ([vp, i32, i32, C.c_float, C.c_uint64, vp], i32),
        "cm_default_pn_config": ([C.POINTER(PnConfig)], i32),
        "cm_set_post_newtonian": ([vp, C.POINTER(PnConfig)], i32),
        "cm_pn_warnings": ([vp, vp, i32, C.POINTER(i32)], i32),
        "cm_default_softening_config": ([C.POINTER(SofteningConfig)], i32),
        "cm_set_softening": ([vp, C.POINTER(SofteningConfig)], i32),
        "cm_yoshida4_step": ([vp, f64, f64, i32], i32),
        "cm_default_collision_config": ([C.POINTER(CollisionConfig)], i32),
        "cm_configure_collisions": ([vp, C.POINTER(CollisionConfig)], i32),
        "cm_resolve_collisions": ([vp, C.POINTER(i32)], i32),
        "cm_nbody_step": ([vp, f64, f64, i32, i32, C.POINTER(i32)], i32),
        "cm_nbody_step_diag": ([vp, f64, f64, i32, i32, i32, vp], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


class BodySoA:
    """Structure-of-arrays body store (SoA/BodySoA.cs:39-175): 15 float64
    arrays, IsActive / IsCollidable as uint8 and BodyTypeIndex as int32."""

    def __init__(self, capacity):
        if capacity <= 0:
            raise ValueError("capacity")   # BodySoA.cs:83
        self.capacity = int(capacity)
        self.count = 0
        for f in F64_FIELDS:
            setattr(self, f, np.zeros(capacity, dtype=np.float64))
        for f in U8_FIELDS:
            setattr(self, f, np.zeros(capacity, dtype=np.uint8))
        self.body_type = np.zeros(capacity, dtype=np.int32)

    @classmethod
    def from_arrays(cls, count=None, capacity=None, **arrays):
        n = count if count is not None else len(arrays["pos_x"])
        soa = cls(capacity or max(n, 1))
        soa.count = n
        for k, v in arrays.items():
            getattr(soa, k)[:n] = v
        if "old_acc_x" not in arrays:   # CopyFrom seeds OldAcc = Acc (BodySoA.cs:141-144)
            soa.old_acc_x[:n] = soa.acc_x[:n]
            soa.old_acc_y[:n] = soa.acc_y[:n]
            soa.old_acc_z[:n] = soa.acc_z[:n]
        return soa

    def copy(self):
        out = BodySoA(self.capacity)
        out.count = self.count
        for f in ALL_FIELDS:
            getattr(out, f)[:] = getattr(self, f)
        return out

    def ptrs(self):
        p = SoAPtrs()
        for f in F64_FIELDS:
            setattr(p, f, _ptr(getattr(self, f), C.c_double))
        for f in U8_FIELDS:
            setattr(p, f, _ptr(getattr(self, f), C.c_uint8))
        p.body_type = _ptr(self.body_type, C.c_int32)
        return p


def generate_plummer(count, total_mass=1.0, scale_length=1.0, seed=42,
                     mass_fraction_cut=1.0, com_frame=True, capacity=None):
    """PlummerSphere.Generate (ScenarioGenerators/PlummerSphere.cs:41-116) with a
    SplitMix64 stream; optional truncation at a mass fraction; all bodies
    active and non-collidable, radius 0.02."""
    soa = BodySoA(capacity or max(count, 1))
    rc = load_library().cm_generate_plummer(count, total_mass, scale_length,
                                            C.c_uint64(seed), mass_fraction_cut,
                                            int(com_frame), C.byref(soa.ptrs()))
    if rc != CM_OK:
        raise CmError(rc, "cm_generate_plummer")
    soa.count = count
    soa.is_active[:count] = 1
    return soa


def generate_plummer_dotnet(count, total_mass=100.0, scale_length=2.0, seed=42,
                            capacity=None):
    """PlummerSphere.Generate(count, totalMass, scaleLength, seed)
    (ScenarioGenerators/PlummerSphere.cs:41-116), with the reference's
    defaults, body for body: the NextDouble stream of .NET's System.Random(seed).
    All bodies active, radius 0.02; no centre-of-mass shift."""
    soa = BodySoA(capacity or max(count, 1))
    rc = load_library().cm_generate_plummer_dotnet(count, total_mass, scale_length, int(seed),
                                                   C.byref(soa.ptrs()))
    if rc != CM_OK:
        raise CmError(rc, "cm_generate_plummer_dotnet")
    soa.count = count
    soa.is_active[:count] = 1
    return soa


def generate_merger(count, seed=42, separation=10.0, speed_fraction=0.3, mass_each=0.5,
                    scale_length=1.0, mass_fraction_cut=0.999, capacity=None):
    """Two equal Plummer spheres (count / 2 bodies each, SplitMix64 seeds seed
    and seed + 1) with centres `separation` apart on the x axis, approaching
    head-on with relative speed speed_fraction x the escape speed
    sqrt(2 G M_total / separation); centre-of-mass frame.  No reference
    generator exists for it (a synthetic merger workload)."""
    h1 = count // 2
    h2 = count - h1
    a = generate_plummer(h1, mass_each, scale_length, seed, mass_fraction_cut)
    b = generate_plummer(h2, mass_each, scale_length, seed + 1, mass_fraction_cut)
    v = speed_fraction * np.sqrt(2.0 * 2.0 * mass_each / separation)
    soa = BodySoA(capacity or max(count, 1))
    for f in F64_FIELDS + U8_FIELDS + I32_FIELDS:
        getattr(soa, f)[:h1] = getattr(a, f)[:h1]
        getattr(soa, f)[h1:count] = getattr(b, f)[:h2]
    soa.pos_x[:h1] -= 0.5 * separation
    soa.pos_x[h1:count] += 0.5 * separation
    soa.vel_x[:h1] += 0.5 * v
    soa.vel_x[h1:count] -= 0.5 * v
    soa.count = count
    return soa


def generate_disk(count, seed=42, star_radius=0.005, a_min=0.5, a_max=5.0, sigma_e=0.02,
                  sigma_i=0.01, m_min=1e-9, m_max=1e-6, density=ROCK_DENSITY_INFLATED,
                  capacity=None):
    """Cold thin disk of count-1 bodies around a 1 Msun star (see cm_generate_disk)."""
    soa = BodySoA(capacity or max(count, 1))
    rc = load_library().cm_generate_disk(count, C.c_uint64(seed), star_radius, a_min, a_max,
                                         sigma_e, sigma_i, m_min, m_max, density,
                                         C.byref(soa.ptrs()))
    if rc != CM_OK:
        raise CmError(rc, "cm_generate_disk")
    soa.count = count
    return soa


class HipPhysicsBackend:
    """IPhysicsComputeBackend + IGravityModelAwareBackend on the GPU.

    compute_forces(bodies, softening) follows the C# contract exactly: it
    reads positions / masses / radii / active flags of `bodies`, and
    overwrites AccX/Y/Z for every index < Count (inactive bodies get 0).
    For device-resident stepping use upload() / step() / download()."""

    def __init__(self, capacity, kind=BACKEND_BARNES_HUT, precision=PREC_FP64,
                 theta=0.5, morton_levels=10, leaf_capacity=16,
                 group_size=64, enable_shell_theorem=False, learned_correction=False,
                 far_field_ratio=0.0, device=0):
        L = load_library()
        h = C.c_void_p()
        rc = L.cm_create(int(capacity), int(device), C.byref(h))
        if rc != CM_OK:
            raise CmError(rc, "cm_create")
        self._h = h
        self._L = L
        self.capacity = int(capacity)
        self.config = BackendConfig()
        L.cm_default_config(C.byref(self.config))
        self.config.kind = kind
        self.config.precision = precision
        self.config.theta = theta
        self.config.morton_levels = morton_levels
        self.config.leaf_capacity = leaf_capacity
        self.config.group_size = group_size
        self.config.shell_theorem = int(enable_shell_theorem)
        self.config.learned_correction = int(learned_correction)
        self.config.far_field_ratio = float(far_field_ratio)
        self._apply()

    # -- configuration ----------------------------------------------------
    def _check(self, rc, what):
        if rc != CM_OK:
            msg = self._L.cm_last_error(self._h)
            raise CmError(rc, f"{what}: {msg.decode() if msg else ''}")

    def _apply(self):
        self._check(self._L.cm_configure(self._h, C.byref(self.config)), "cm_configure")

    def configure(self, **kw):
        for k, v in kw.items():
            if k == "enable_shell_theorem":
                k, v = "shell_theorem", int(v)
            setattr(self.config, k, v)
        self._apply()

    @property
    def enable_shell_theorem(self):
        return bool(self.config.shell_theorem)

    @enable_shell_theorem.setter
    def enable_shell_theorem(self, v):
        self.configure(shell_theorem=int(bool(v)))

    @property
    def theta(self):
        return self.config.theta

    @theta.setter
    def theta(self, v):
        self.configure(theta=float(v))

    # -- IPhysicsComputeBackend ---------------------------------------------
    def compute_forces(self, bodies, softening):
        self.upload_force_inputs(bodies)
        self._check(self._L.cm_compute_forces(self._h, float(softening)), "cm_compute_forces")
        self.download(bodies, fields=("acc_x", "acc_y", "acc_z"))

    def upload_force_inputs(self, bodies):
        """Only what ComputeForces reads (cm_upload_force_inputs)."""
        self._check(self._L.cm_upload_force_inputs(self._h, C.byref(bodies.ptrs()), bodies.count),
                    "cm_upload_force_inputs")

    # -- device-resident interface ----------------------------------------
    def upload(self, bodies):
        self._check(self._L.cm_upload_soa(self._h, C.byref(bodies.ptrs()), bodies.count),
                    "cm_upload_soa")

    def download(self, bodies, fields=ALL_FIELDS):
        p = SoAPtrs()
        for f in fields:
            a = getattr(bodies, f)
            ct = C.c_double if f in F64_FIELDS else (C.c_uint8 if f in U8_FIELDS else C.c_int32)
            setattr(p, f, _ptr(a, ct))
        n = self.count
        bodies.count = n
        self._check(self._L.cm_download_soa(self._h, C.byref(p), n), "cm_download_soa")

    @property
    def count(self):
        c = C.c_int32()
        self._check(self._L.cm_get_count(self._h, C.byref(c)), "cm_get_count")
        return c.value

    def compute_forces_resident(self, softening):
        self._check(self._L.cm_compute_forces(self._h, float(softening)), "cm_compute_forces")

    def step(self, softening, dt, n_steps=1):
        self._check(self._L.cm_verlet_step(self._h, float(softening), float(dt), int(n_steps)),
                    "cm_verlet_step")

    def nbody_step(self, softening, dt, n_steps=1, collisions=True):
        """n_steps of NBodySolver.StepSoA on the device: Verlet step, then
        detection and resolution with the configured collision mode."""
        k = C.c_int32()
        self._check(self._L.cm_nbody_step(self._h, float(softening), float(dt), int(n_steps),
                                          int(bool(collisions)), C.byref(k)), "cm_nbody_step")
        return k.value

    def nbody_step_diag(self, softening, dt, n_steps=1, collisions=False, with_potential=1):
        """n_steps of NBodySolver.Step: returns the SimulationState of each step."""
        st = (SimulationState * max(1, int(n_steps)))()
        self._check(self._L.cm_nbody_step_diag(self._h, float(softening), float(dt),
                                               int(n_steps), int(bool(collisions)),
                                               int(with_potential), C.cast(st, C.c_void_p)),
                    "cm_nbody_step_diag")
        return [st[k].as_dict() for k in range(int(n_steps))]

    def yoshida_step(self, softening, dt, n_steps=1):
        """n_steps of the engine's Yoshida4 integrator (three Verlet sub-steps)."""
        self._check(self._L.cm_yoshida4_step(self._h, float(softening), float(dt),
                                             int(n_steps)), "cm_yoshida4_step")

    # -- PostNewtonianBackend decorator and the engine's softening modes -----
    def set_post_newtonian(self, enabled=True, max_velocity_fraction_c=0.3,
                           schwarzschild_warning_factor=3.0):
        c = PnConfig()
        self._L.cm_default_pn_config(C.byref(c))
        c.enabled = int(bool(enabled))
        c.max_velocity_fraction_c = float(max_velocity_fraction_c)
        c.schwarzschild_warning_factor = float(schwarzschild_warning_factor)
        self._check(self._L.cm_set_post_newtonian(self._h, C.byref(c)), "cm_set_post_newtonian")

    def pn_warnings(self, cap=4096):
        """Strong-field pairs skipped by the last evaluation: (list of
        (body, other, distance, rs) in (body, other) order, total count)."""
        buf = (PnWarning * max(1, cap))()
        n = C.c_int32()
        self._check(self._L.cm_pn_warnings(self._h, buf, int(cap), C.byref(n)), "cm_pn_warnings")
        k = min(n.value, cap, 4096)
        return [(w.body, w.other, w.distance, w.schwarzschild_radius) for w in buf[:k]], n.value

    def set_softening(self, mode=SOFTENING_GLOBAL, type_softening=None, adaptive_scale=0.01,
                      adaptive_min=1e-8):
        c = SofteningConfig()
        self._L.cm_default_softening_config(C.byref(c))
        c.mode = int(mode)
        if type_softening is not None:
            for k, v in enumerate(type_softening):
                c.type_softening[k] = float(v)
        c.adaptive_scale = float(adaptive_scale)
        c.adaptive_min = float(adaptive_min)
        self._check(self._L.cm_set_softening(self._h, C.byref(c)), "cm_set_softening")

    def stats(self):
        s = ForceStats()
        self._check(self._L.cm_get_force_stats(self._h, C.byref(s)), "cm_get_force_stats")
        return s.as_dict()

    def reset_stats(self):
        self._check(self._L.cm_reset_force_stats(self._h), "cm_reset_force_stats")

    def energy(self, softening, with_potential=True, noalloc=False):
        """with_potential: False/0 none, True/1 exact pair sum, 2 Barnes-Hut.
        noalloc: use cm_energy_noalloc (no device allocation; same results)."""
        r = EnergyReport()
        fn = self._L.cm_energy_noalloc if noalloc else self._L.cm_energy
        self._check(fn(self._h, float(softening), int(with_potential), C.byref(r)),
                    "cm_energy_noalloc" if noalloc else "cm_energy")
        return {"kinetic": r.kinetic, "potential": r.potential,
                "momentum": tuple(r.momentum), "angular_momentum": tuple(r.angular_momentum),
                "angular_momentum_mag": r.angular_momentum_mag,
                "total_mass": r.total_mass, "active_count": r.active_count,
                "center_of_mass": tuple(r.center_of_mass)}

    def max_acceleration(self):
        """Largest |a| over active bodies of the accelerations the engine holds
        (Engine::compute_max_acceleration); 0 when none is active."""
        out = C.c_double(0.0)
        self._check(self._L.cm_max_acceleration(self._h, C.byref(out)), "cm_max_acceleration")
        return out.value

    def potential_rows(self, softening, method=2):
        """Per-body shares 0.5 m_i phi_i of the potential energy (method 1 exact,
        2 Barnes-Hut), in body order."""
        n = self.count
        rows = np.zeros(max(n, 1), dtype=np.float64)
        self._check(self._L.cm_potential_rows(self._h, float(softening), int(method),
                                              rows.ctypes.data_as(C.POINTER(C.c_double))),
                    "cm_potential_rows")
        return rows[:n]

    def direct_sample(self, softening, targets):
        t = np.ascontiguousarray(targets, dtype=np.int32)
        ax, ay, az = (np.zeros(len(t)) for _ in range(3))
        self._check(self._L.cm_direct_sample(self._h, float(softening), t.ctypes.data, len(t),
                                             ax.ctypes.data, ay.ctypes.data, az.ctypes.data),
                    "cm_direct_sample")
        return ax, ay, az

    def compare_forces(self, softening, theta, targets=None):
        """ForceErrorAnalyzer.CompareForces on the device state: exact FP64
        direct sum vs the Barnes-Hut backend at theta (all bodies, or the
        given target indices); leaves the Barnes-Hut forces in acc."""
        r = ForceComparison()
        t = None if targets is None else np.ascontiguousarray(targets, dtype=np.int32)
        self._check(self._L.cm_compare_forces(self._h, float(softening), float(theta),
                                              None if t is None else t.ctypes.data,
                                              0 if t is None else len(t), C.byref(r)),
                    "cm_compare_forces")
        return r.as_dict()

    def detect_collisions(self, cap=1 << 16):
        out = np.zeros(cap, dtype=EVENT_DTYPE)
        ne = C.c_int32()
        self._check(self._L.cm_detect_collisions(self._h, out.ctypes.data, cap, C.byref(ne)),
                    "cm_detect_collisions")
        return out[:min(ne.value, cap)].copy(), ne.value

    def resolve_merges(self):
        n = C.c_int32()
        self._check(self._L.cm_resolve_merges(self._h, C.byref(n)), "cm_resolve_merges")
        return n.value

    def configure_collisions(self, mode=COLLISION_MERGE_ONLY, default_restitution=0.15,
                             fragmentation_specific_energy_threshold=0.6,
                             fragmentation_mass_loss_cap=0.3, capture_velocity_factor=0.9,
                             promote_compact_remnants=False, max_merges_per_frame=0,
                             max_merges_per_body=0):
        c = CollisionConfig()
        self._L.cm_default_collision_config(C.byref(c))
        c.mode = int(mode)
        c.default_restitution = float(default_restitution)
        c.fragmentation_specific_energy_threshold = float(fragmentation_specific_energy_threshold)
        c.fragmentation_mass_loss_cap = float(fragmentation_mass_loss_cap)
        c.capture_velocity_factor = float(capture_velocity_factor)
        c.promote_compact_remnants = int(bool(promote_compact_remnants))
        c.max_merges_per_frame = int(max_merges_per_frame)
        c.max_merges_per_body = int(max_merges_per_body)
        self._check(self._L.cm_configure_collisions(self._h, C.byref(c)),
                    "cm_configure_collisions")

    def resolve_collisions(self):
        n = C.c_int32()
        self._check(self._L.cm_resolve_collisions(self._h, C.byref(n)), "cm_resolve_collisions")
        return n.value

    def compact(self, with_old_index=True):
        """Removes the inactive bodies in place; returns the new count and,
        with with_old_index, each survivor's index before the call (a host
        copy of n int32 values; None without it)."""
        n = C.c_int32()
        if not with_old_index:
            self._check(self._L.cm_compact(self._h, C.byref(n), None), "cm_compact")
            return n.value, None
        old = np.zeros(max(self.count, 1), dtype=np.int32)
        self._check(self._L.cm_compact(self._h, C.byref(n), old.ctypes.data), "cm_compact")
        return n.value, old[:n.value].copy()

    # -- learned far-field correction ----------------------------------------
    def mlp_init(self, seed=1):
        self._check(self._L.cm_mlp_init(self._h, int(seed)), "cm_mlp_init")

    def mlp_params(self):
        p = np.zeros(MLP_PARAMS, dtype=np.float32)
        self._check(self._L.cm_mlp_get_params(self._h, p.ctypes.data, MLP_PARAMS),
                    "cm_mlp_get_params")
        return p

    def mlp_set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        self._check(self._L.cm_mlp_set_params(self._h, p.ctypes.data, len(p)),
                    "cm_mlp_set_params")

    def mlp_forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_IN)
        y = np.zeros((len(x), MLP_OUT), dtype=np.float32)
        self._check(self._L.cm_mlp_forward(self._h, x.ctypes.data, len(x), y.ctypes.data),
                    "cm_mlp_forward")
        return y

    def mlp_gradient(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_IN)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, MLP_USED)
        g = np.zeros(MLP_PARAMS, dtype=np.float32)
        loss = C.c_float()
        self._check(self._L.cm_mlp_gradient(self._h, x.ctypes.data, t.ctypes.data, len(x),
                                             g.ctypes.data, C.byref(loss)), "cm_mlp_gradient")
        return g, loss.value

    def mlp_collect(self, softening, max_groups=256, seed=7, max_pairs=1 << 20):
        n = C.c_int32()
        self._check(self._L.cm_mlp_collect(self._h, float(softening), int(max_groups),
                                           int(seed), int(max_pairs), C.byref(n)),
                    "cm_mlp_collect")
        return n.value

    def mlp_pairs(self, n):
        x = np.zeros((n, MLP_IN), dtype=np.float32)
        t = np.zeros((n, MLP_USED), dtype=np.float32)
        gn = np.zeros((n, 4), dtype=np.int32)
        c = np.zeros((n, 3))
        self._check(self._L.cm_mlp_get_pairs(self._h, x.ctypes.data, t.ctypes.data,
                                             gn.ctypes.data, c.ctypes.data, n),
                    "cm_mlp_get_pairs")
        return x, t, gn, c

    def mlp_train(self, epochs=20, batch=1024, lr=1e-3, seed=11):
        losses = np.zeros(epochs, dtype=np.float32)
        self._check(self._L.cm_mlp_train(self._h, int(epochs), int(batch), float(lr),
                                         int(seed), losses.ctypes.data), "cm_mlp_train")
        return losses

    def close(self):
        if self._h:
            self._L.cm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SoAVerletIntegrator:
    """ISoAIntegrator.Step (Integrators/SoAVerletIntegrator.cs:48-115) with the
    HIP backend: upload, one device-resident KDK step, download."""

    name = "Verlet-SoA"

    def step(self, bodies, backend, softening, dt):
        backend.upload(bodies)
        backend.step(softening, dt, 1)
        backend.download(bodies)


class CollisionDetector:
    """CollisionDetector.Detect (Collisions/CollisionDetector.cs:35-68) on the
    device state of a HipPhysicsBackend: returns the events (a structured array
    with the CollisionEvent fields) sorted deepest overlap first."""

    def detect(self, backend, cap=1 << 16):
        events, total = backend.detect_collisions(cap)
        if total > cap:
            events, total = backend.detect_collisions(total)
        return events


class CollisionResolver:
    """CollisionResolver (Collisions/CollisionResolver.cs:23-247): Configure
    sets the mode and parameters, Resolve applies the events of the last
    detection on the same backend; last_collision_count mirrors
    LastCollisionCount."""

    last_collision_count = 0

    def __init__(self):
        self._cfg = {}

    def configure(self, mode, default_restitution, fragmentation_specific_energy_threshold,
                  fragmentation_mass_loss_cap, capture_velocity_factor):
        self._cfg = dict(mode=mode, default_restitution=default_restitution,
                         fragmentation_specific_energy_threshold=
                         fragmentation_specific_energy_threshold,
                         fragmentation_mass_loss_cap=fragmentation_mass_loss_cap,
                         capture_velocity_factor=capture_velocity_factor)

    def resolve(self, backend, promote_compact_remnants=False):
        if self._cfg or promote_compact_remnants:
            backend.configure_collisions(promote_compact_remnants=promote_compact_remnants,
                                         **self._cfg)
            self.last_collision_count = backend.resolve_collisions()
        else:
            self.last_collision_count = backend.resolve_merges()
        return self.last_collision_count


class DiagnosticThresholds:
    """DiagnosticThresholds (engine/include/celestial/profile/energy_tracker.hpp:14-19)."""

    def __init__(self, max_energy_drift=0.001, max_momentum_drift=1e-8,
                 max_angular_momentum_drift=1e-8, max_com_velocity_drift=1e-10):
        self.max_energy_drift = max_energy_drift
        self.max_momentum_drift = max_momentum_drift
        self.max_angular_momentum_drift = max_angular_momentum_drift
        self.max_com_velocity_drift = max_com_velocity_drift


class EnergyTracker:
    """EnergyTracker (engine/src/profile/energy_tracker.cpp): snapshots of a
    backend's device state and their drift against the first one recorded.

    compute() takes one cm_energy call (compensated sums on the GPU; the
    potential is the exact pair sum, with_potential=1, or the Barnes-Hut
    potential of compute_with_bh, with_potential=2) and derives the snapshot's
    other fields the way energy_tracker.cpp:65-92 does: |p| and |L|, the centre
    of mass, its velocity (the momentum over the total mass: the engine's
    com_mv sums are the momentum sums), the virial ratio 2 KE / |PE|.
    record(), the drifts, check_diagnostics() and the 300-entry rolling window
    follow energy_tracker.cpp:169-262 (host bookkeeping, no device work)."""

    ROLLING_WINDOW = 300

    def __init__(self):
        self.reset()

    def reset(self):
        self.initial = None
        self.current = None
        self.accumulated_error = 0.0
        self._energy = [0.0] * self.ROLLING_WINDOW
        self._drift = [0.0] * self.ROLLING_WINDOW
        self._count = 0
        self._index = 0

    @property
    def has_initial(self):
        return self.initial is not None

    @staticmethod
    def compute(backend, softening, with_potential=1, step_number=0):
        e = backend.energy(softening, with_potential=with_potential)
        ke, pe = e["kinetic"], e["potential"]
        p, lv, m = e["momentum"], e["angular_momentum"], e["total_mass"]
        snap = {"kinetic_energy": ke, "potential_energy": pe, "total_energy": ke + pe,
                "momentum": tuple(p), "momentum_magnitude": math.sqrt(p[0] * p[0] + p[1] * p[1] +
                                                                      p[2] * p[2]),
                "angular_momentum": tuple(lv), "angular_momentum_magnitude":
                    math.sqrt(lv[0] * lv[0] + lv[1] * lv[1] + lv[2] * lv[2]),
                "total_mass": m, "com": (0.0, 0.0, 0.0), "com_velocity": (0.0, 0.0, 0.0),
                "step_number": int(step_number)}
        if m > 0.0:
            snap["com"] = tuple(e["center_of_mass"])
            snap["com_velocity"] = (p[0] / m, p[1] / m, p[2] / m)
        snap["virial_ratio"] = 2.0 * ke / abs(pe) if pe != 0.0 else 0.0
        return snap

    def record(self, snap):
        if self.initial is None:
            self.initial = snap
        self.current = snap
        drift = abs(self.energy_drift())
        self.accumulated_error += drift
        self._energy[self._index] = snap["total_energy"]
        self._drift[self._index] = drift
        self._index = (self._index + 1) % self.ROLLING_WINDOW
        self._count = min(self._count + 1, self.ROLLING_WINDOW)

    def energy_drift(self):
        if self.initial is None or self.initial["total_energy"] == 0.0:
            return 0.0
        e0 = self.initial["total_energy"]
        return (self.current["total_energy"] - e0) / abs(e0)

    def _vec_drift(self, key):
        if self.initial is None:
            return 0.0
        a, b = self.current[key], self.initial[key]
        d = (a[0] - b[0], a[1] - b[1], a[2] - b[2])
        return math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])

    def momentum_drift(self):
        return self._vec_drift("momentum")

    def angular_momentum_drift(self):
        return self._vec_drift("angular_momentum")

    def com_position_drift(self):
        return self._vec_drift("com")

    def com_velocity_drift(self):
        return self._vec_drift("com_velocity")

    def check_diagnostics(self, thresholds=None):
        t = thresholds or DiagnosticThresholds()
        if self.initial is None:
            return True
        return not (abs(self.energy_drift()) > t.max_energy_drift or
                    self.momentum_drift() > t.max_momentum_drift or
                    self.angular_momentum_drift() > t.max_angular_momentum_drift or
                    self.com_velocity_drift() > t.max_com_velocity_drift)

    def rolling_avg_energy(self):
        return sum(self._energy[:self._count]) / self._count if self._count else 0.0

    def rolling_avg_drift(self):
        return sum(self._drift[:self._count]) / self._count if self._count else 0.0


def generate_ensemble(n_systems, seed=42, n_min=8, n_max=2048, m_min=1e-7, m_max=1e-3,
                      a_min=0.1, a_max=30.0, sigma_e=0.05):
    """Planetary systems for the batched ensemble (see cm_generate_ensemble).
    Returns (BodySoA of all bodies, offsets)."""
    L = load_library()
    offsets = np.zeros(n_systems + 1, dtype=np.int32)
    rc = L.cm_generate_ensemble(n_systems, C.c_uint64(seed), n_min, n_max, m_min, m_max,
                                a_min, a_max, sigma_e, offsets.ctypes.data, None)
    if rc != CM_OK:
        raise CmError(rc, "cm_generate_ensemble")
    soa = BodySoA(int(offsets[-1]))
    rc = L.cm_generate_ensemble(n_systems, C.c_uint64(seed), n_min, n_max, m_min, m_max,
                                a_min, a_max, sigma_e, offsets.ctypes.data,
                                C.byref(soa.ptrs()))
    if rc != CM_OK:
        raise CmError(rc, "cm_generate_ensemble")
    soa.count = int(offsets[-1])
    return soa, offsets


class HipEnsemble:
    """Many independent systems stepped together (cm_ensemble_*): each system
    is a BodySoA slice advanced by SoAVerletIntegrator.Step with the
    CpuParallelBackend direct sweep."""

    def __init__(self, offsets, precision=PREC_FP32_SPLIT, threads=0, steps_per_launch=100,
                 device=0):
        L = load_library()
        self._L = L
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int32)
        self.n_systems = len(self.offsets) - 1
        h = C.c_void_p()
        rc = L.cm_ensemble_create(self.n_systems, self.offsets.ctypes.data, device, C.byref(h))
        if rc != CM_OK:
            raise CmError(rc, "cm_ensemble_create")
        self._h = h
        self._check(L.cm_ensemble_configure(h, precision, threads, steps_per_launch),
                    "cm_ensemble_configure")

    def _check(self, rc, what):
        if rc != CM_OK:
            msg = self._L.cm_ensemble_last_error(self._h)
            raise CmError(rc, f"{what}: {msg.decode() if msg else ''}")

    def upload(self, bodies):
        self._check(self._L.cm_ensemble_upload(self._h, C.byref(bodies.ptrs())),
                    "cm_ensemble_upload")

    def download(self, bodies):
        self._check(self._L.cm_ensemble_download(self._h, C.byref(bodies.ptrs())),
                    "cm_ensemble_download")

    def step(self, softening, dt, n_steps):
        self._check(self._L.cm_ensemble_step(self._h, float(softening), float(dt), int(n_steps)),
                    "cm_ensemble_step")

    def energy(self, softening):
        out = np.zeros(self.n_systems)
        self._check(self._L.cm_ensemble_energy(self._h, float(softening), out.ctypes.data),
                    "cm_ensemble_energy")
        return out

    def stats(self):
        s = EnsembleStats()
        self._check(self._L.cm_ensemble_get_stats(self._h, C.byref(s)), "cm_ensemble_get_stats")
        return s.as_dict()

    def close(self):
        if self._h:
            self._L.cm_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
