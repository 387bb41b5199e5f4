/*
 * cmhip.h -- C ABI of libcmhip.so, the MI355X (gfx950) N-body backend.
 *
 * The library is a drop-in compute backend for the reference's C# physics
 * surface (mfscpayload-690/celestial-mechanics-desktop,
 * src/CelestialMechanics.Physics).  A C# class implementing
 * IPhysicsComputeBackend / IGravityModelAwareBackend binds these entry points
 * with [DllImport("cmhip")] (celestial-mechanics-desktop_amd/dotnet/
 * HipPhysicsBackend.cs).
 *
 * Conventions (modelled on the reference's native ABI,
 * engine/include/celestial/interop/native_api.h:30-365, and its ErrorCode,
 * engine/include/celestial/core/error.hpp:9-18):
 *   - every function returns int32_t: 0 on success, a negative CM_ERR_* code
 *     otherwise; no exception ever crosses the boundary;
 *   - state lives in an opaque engine handle (not a global singleton); the
 *     last error message is kept per handle (cm_last_error);
 *   - the caller owns host memory, the library owns device memory; host arrays
 *     are copied in (cm_upload_soa) and out (cm_download_soa);
 *   - bool arrays travel as uint8_t (0/1);
 *   - calls on one handle must be serialised by the caller; different handles
 *     are independent.
 */
#ifndef CMHIP_H
#define CMHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes (values of the reference's ErrorCode) ------------------ */
#define CM_OK 0
#define CM_ERR_HIP (-1)              /* ErrorCode::CudaError */
#define CM_ERR_OUT_OF_MEMORY (-2)    /* ErrorCode::OutOfMemory */
#define CM_ERR_INVALID_ARGUMENT (-3) /* ErrorCode::InvalidArgument */
#define CM_ERR_POOL_EXHAUSTED (-4)   /* ErrorCode::PoolExhausted */
#define CM_ERR_DEVICE_NOT_FOUND (-5) /* ErrorCode::DeviceNotFound */
#define CM_ERR_NOT_INITIALIZED (-6)  /* ErrorCode::NotInitialized */

/* ---- force backends ----------------------------------------------------- */
/* CM_BACKEND_DIRECT_PAIRWISE: CpuSingleThreadBackend.ComputeForces
 *   (Solvers/CpuSingleThreadBackend.cs:50-158): Newton-3 pair order
 *   (shell theorem off) or the full GravityKernel sweep (shell on).
 * CM_BACKEND_DIRECT: CpuParallelBackend.ComputeForces
 *   (Solvers/CpuParallelBackend.cs:69-142): full j sweep per body.
 * CM_BACKEND_BARNES_HUT: BarnesHutBackend.ComputeForces
 *   (Solvers/BarnesHutBackend.cs:118-236) re-designed for the GPU: Morton
 *   keys, stable radix sort, level-order bucket octree, bottom-up centre of
 *   mass, group-MAC traversal with the reference's (2h)^2 < theta^2 d^2 test. */
#define CM_BACKEND_DIRECT_PAIRWISE 0
#define CM_BACKEND_DIRECT 1
#define CM_BACKEND_BARNES_HUT 2

/* FP64: every interaction in double, IEEE 1/sqrt (bit-reproducible).
 * FP64_FAST: double with a hardware reciprocal square root refined by Newton
 *   steps (the plain FP64 GPU path used as the performance baseline).
 * FP32_SPLIT: pair terms in float on offsets taken relative to the target
 *   group's centre (computed in double), partial sums folded into double. */
#define CM_PREC_FP64 0
#define CM_PREC_FP64_FAST 1
#define CM_PREC_FP32_SPLIT 2

typedef struct {
    int32_t kind;              /* CM_BACKEND_* */
    int32_t precision;         /* CM_PREC_* */
    int32_t shell_theorem;     /* IGravityModelAwareBackend.EnableShellTheorem */
    int32_t morton_levels;     /* key bits / 3: 10 (30-bit keys) .. 21 */
    double theta;              /* BarnesHutBackend.Theta, clamped to >= 0.2 */
    int32_t leaf_capacity;     /* bodies per leaf bucket, 1..64 */
    int32_t group_size;        /* max bodies per target group (one wavefront), 1..64 */
    int32_t learned_correction;/* 1: add the MLP far-field correction (Barnes-Hut,
                                  FP32_SPLIT only; no reference counterpart) */
    double far_field_ratio;    /* kappa in [0, 0.5]: accepted cells farther than
                                  R/kappa from a target group's centre (R its
                                  half diagonal) are summed for the group as a
                                  second-order Taylor expansion of their
                                  monopole field (Barnes-Hut FP32_SPLIT without
                                  the learned correction; 0 = off; no reference
                                  counterpart) */
} cm_backend_config;

/* Pointers to the 18 BodySoA arrays (SoA/BodySoA.cs:42-69).  On upload a
 * NULL pointer means zeros, except old_acc_* which, when NULL, is seeded from
 * acc_* as BodySoA.CopyFrom does (SoA/BodySoA.cs:141-144).  On download a
 * NULL pointer skips that array. */
typedef struct {
    double *pos_x, *pos_y, *pos_z;
    double *vel_x, *vel_y, *vel_z;
    double *acc_x, *acc_y, *acc_z;
    double *old_acc_x, *old_acc_y, *old_acc_z;
    double *mass, *radius, *density;
    uint8_t *is_active, *is_collidable;
    int32_t *body_type;
} cm_soa_ptrs;

/* Timing and work counters of the last force evaluation. */
typedef struct {
    int64_t cell_interactions;  /* body-node monopole terms */
    int64_t body_interactions;  /* body-body terms (self excluded) */
    int32_t n_active;
    int32_t n_nodes;
    int32_t n_groups;
    int32_t flags;              /* bit 0: node pool exhausted */
    float ms_tree;              /* bbox + keys + sort + build + mass */
    float ms_sort;              /* radix sort only */
    float ms_force;             /* traversal / direct kernel */
    float ms_total;             /* whole force evaluation */
    /* accumulated since the last cm_reset_force_stats: */
    int64_t acc_interactions;   /* cell + body terms */
    double acc_ms_force;        /* summed force-kernel time of the timed evaluations, HIP events */
    double acc_ms_total;        /* summed force-evaluation time of the timed evaluations */
    int32_t acc_evaluations;    /* every force evaluation */
    int64_t far_cells;          /* (group, cell) far-field expansion terms */
    int64_t acc_far_cells;      /* the same, accumulated */
    int32_t acc_timed_evaluations; /* the timed ones: the first evaluation of each call and
                                    * every 4th after it (at most 1024 between two host
                                    * synchronisations) */
} cm_force_stats;

/* EnergyCalculator quantities (Validation/EnergyCalculator.cs:16-81) with
 * compensated (Kahan) FP64 reductions on the GPU. */
typedef struct {
    double kinetic;
    double potential;           /* -sum_{i<j} m_i m_j / sqrt(r^2 + eps^2) */
    double momentum[3];
    double angular_momentum[3];
    double angular_momentum_mag;
    double total_mass;
    int32_t active_count;
    /* sum m r / sum m over active bodies, 0 when the total mass is not
     * positive: the com fields of the engine's GPU energy snapshot
     * (GpuTreeSolver::compute_gpu_energy, engine/src/cuda/gpu_tree.cu:1296-1300,
     * gpu_reduce_com at engine/src/cuda/gpu_reduction.cu:481-517) */
    double center_of_mass[3];
} cm_energy_report;

/* CollisionEvent (Collisions/CollisionEvent.cs:3-11). */
typedef struct {
    int32_t body_a, body_b;
    double overlap_depth, distance;
    double normal_x, normal_y, normal_z;
    double relative_normal_speed;
} cm_collision_event;

/* ---- lifecycle ----------------------------------------------------------- */
/* Replaces celestial_init (engine/include/celestial/interop/native_api.h:33):
 * allocates device state for `capacity` bodies on HIP device `device`. */
int32_t cm_create(int32_t capacity, int32_t device, void** out_engine);
/* Replaces celestial_shutdown (native_api.h:36). */
int32_t cm_destroy(void* engine);
/* Replaces celestial_get_last_error_message (native_api.h:294), per handle. */
const char* cm_last_error(void* engine);
/* IGravityModelAwareBackend.EnableShellTheorem, BarnesHutBackend.Theta and
 * the NBodySolver.SelectBackend choice (Solvers/NBodySolver.cs:451-484). */
int32_t cm_configure(void* engine, const cm_backend_config* cfg);
int32_t cm_default_config(cm_backend_config* cfg);

/* ---- data transfer ------------------------------------------------------ */
/* BodySoA.CopyFrom-equivalent upload (SoA/BodySoA.cs:118-153), replacing
 * celestial_set_particles (native_api.h:41-47).  count <= capacity. */
int32_t cm_upload_soa(void* engine, const cm_soa_ptrs* src, int32_t count);
/* The part of BodySoA that IPhysicsComputeBackend.ComputeForces reads
 * (Solvers/IPhysicsComputeBackend.cs:21-33: positions, masses, radii, active
 * flags), for a host that keeps the bodies and calls the backend once per
 * step: pos_*, mass and is_active are required; radius is required with the
 * shell theorem (else uploaded when given); vel_* are uploaded only (and then
 * required) with the 1PN decorator on, body_type only with per-type
 * softening.  The device's other arrays keep their contents.  41 bytes per
 * body instead of the 126 of cm_upload_soa.  count <= capacity. */
int32_t cm_upload_force_inputs(void* engine, const cm_soa_ptrs* src, int32_t count);
/* BodySoA.CopyTo-equivalent download (SoA/BodySoA.cs:160-174), replacing
 * celestial_get_accelerations/positions/velocities (native_api.h:50-60). */
int32_t cm_download_soa(void* engine, const cm_soa_ptrs* dst, int32_t count);
int32_t cm_get_count(void* engine, int32_t* count);

/* ---- physics -------------------------------------------------------------- */
/* IPhysicsComputeBackend.ComputeForces(BodySoA, softening)
 * (Solvers/IPhysicsComputeBackend.cs:33); replaces celestial_compute_forces
 * (native_api.h:65).  Overwrites acc for every index < count; inactive
 * bodies get 0. */
int32_t cm_compute_forces(void* engine, double softening);
/* n_steps x SoAVerletIntegrator.Step(bodies, backend, softening, dt)
 * (Integrators/SoAVerletIntegrator.cs:48-115), device-resident; replaces
 * celestial_step (native_api.h:68). */
int32_t cm_verlet_step(void* engine, double softening, double dt,
                       int32_t n_steps);
/* EnergyCalculator.ComputeKE/ComputePE/ComputeMomentum/
 * ComputeAngularMomentum (Validation/EnergyCalculator.cs:16-81).
 * with_potential: 0 no PE; 1 the exact pair sum (EnergyCalculator.ComputePE,
 * NBodySolver.cs:313-337); 2 the Barnes-Hut potential energy
 * PE = 1/2 sum m_i phi_i of the engine's launch_gpu_potential_energy
 * (engine/src/cuda/potential_energy_kernel.cu:134-171), phi through this
 * library's FP64 group traversal with the configured theta (no shell
 * theorem), on a tree built for the current positions. */
int32_t cm_energy(void* engine, double softening, int32_t with_potential,
                  cm_energy_report* out);
/* cm_energy that allocates nothing (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same inputs,
 * outputs and results as cm_energy (bit for bit), but the per-body potential
 * rows and the reduction partials live in a static device array of the
 * library (32 MiB + 108 KiB per GPU, kept until the process ends) instead of
 * a per-call hipMalloc.  with_potential > 0 holds at most 4,194,304 bodies
 * (CM_ERR_INVALID_ARGUMENT beyond).  Every engine on one GPU shares that
 * array: calls that overlap on one GPU (from several threads or engines)
 * race and must be serialised by the caller. */
int32_t cm_energy_noalloc(void* engine, double softening, int32_t with_potential,
                          cm_energy_report* out);
/* Per-body shares of the potential energy, rows[i] = 1/2 m_i phi_i (0 for
 * inactive bodies; host array of count doubles): method 1 exact, method 2
 * Barnes-Hut as in cm_energy.  The per-body m_i phi_i the engine's
 * compute_potential_bh_kernel writes (potential_energy_kernel.cu:29-111),
 * halved. */
int32_t cm_potential_rows(void* engine, double softening, int32_t method, double* rows);
/* Engine::compute_max_acceleration (engine/src/sim/engine.cpp:468-500,
 * k_reduce_max_accel at engine/src/cuda/gpu_reduction.cu:269-283): the largest
 * |a| = sqrt(ax^2 + ay^2 + az^2) over active bodies of the accelerations the
 * engine holds (0 when none is active), the input of the adaptive time step. */
int32_t cm_max_acceleration(void* engine, double* out);
/* FP64 direct-sum accelerations on a subset of target bodies: the judge of
 * ForceErrorAnalyzer.CompareForces (Validation/ForceErrorAnalyzer.cs:67-92)
 * restricted to a sample.  Results are host arrays of length n_targets. */
int32_t cm_direct_sample(void* engine, double softening,
                         const int32_t* targets, int32_t n_targets,
                         double* acc_x, double* acc_y, double* acc_z);
/* ForceErrorAnalyzer.CompareForces(bodies, softening, theta)
 * (Validation/ForceErrorAnalyzer.cs:67-135): exact FP64 direct-sum forces,
 * then the Barnes-Hut backend's forces at `theta` (the engine's configured
 * precision, leaf capacity, group size and far field; global softening, no
 * shell theorem, no 1PN, as the analyzer's plain backends), and per active
 * body the relative error |a_bh - a_exact| / |a_exact|, skipping bodies with
 * |a_exact| < 1e-30; mean, max and RMS summed in body order, plus the
 * median (build-added).  n_targets = 0 compares every body (the exact sum is
 * O(N^2) on the GPU); otherwise only the listed targets, in list order
 * (exact forces as cm_direct_sample).  Like the reference, the bodies' acc
 * holds the Barnes-Hut forces afterwards. */
typedef struct {
    double mean_relative_error;    /* ComparisonResult.MeanRelativeError */
    double max_relative_error;     /* MaxRelativeError */
    double rms_relative_error;     /* RmsRelativeError */
    double median_relative_error;  /* no reference counterpart */
    int32_t body_count;            /* BodyCount: bodies compared */
    double theta;                  /* Theta */
} cm_force_comparison;
int32_t cm_compare_forces(void* engine, double softening, double theta,
                          const int32_t* targets, int32_t n_targets,
                          cm_force_comparison* out);
int32_t cm_get_force_stats(void* engine, cm_force_stats* out);
int32_t cm_reset_force_stats(void* engine);

/* ---- collisions ----------------------------------------------------------- */
/* CollisionDetector.Detect (Collisions/CollisionDetector.cs:35-68): events
 * sorted by overlap depth, deepest first.  Copies up to `cap` events to
 * `out` (may be NULL) and stores the total in *n_events.  The full list
 * stays on the device for cm_resolve_merges. */
int32_t cm_detect_collisions(void* engine, cm_collision_event* out,
                             int32_t cap, int32_t* n_events);
/* CollisionResolver.Resolve in CollisionMode.MergeOnly
 * (Collisions/CollisionResolver.cs:58-130, MergePolicy.cs:31-80) over the
 * events of the last cm_detect_collisions. */
int32_t cm_resolve_merges(void* engine, int32_t* n_merged);
/* CollisionResolver.Configure / Resolve with every outcome
 * (Collisions/CollisionResolver.cs:36-247): MergeOnly, BounceOnly and
 * Realistic modes (Types/CollisionMode.cs); Realistic picks Accretion
 * (compact body survives), Merge, Fragmentation or CatastrophicDisruption
 * from CollisionEnergyModel.Evaluate (Astrophysics/CollisionEnergyModel.cs:
 * 20-70).  Configure's clamps apply (restitution to [0, 1], mass-loss cap to
 * [0, 0.9]); a cap below 0.05 is rejected because the reference throws on the
 * first fragmentation-class event with it.  The two unused Configure
 * arguments are kept for the signature.  max_merges_per_frame /
 * max_merges_per_body are the engine's merge safeguards
 * (engine/include/celestial/sim/simulation_config.hpp:52-60,
 * engine/src/physics/collision_resolver.cpp:46-76), MergeOnly only, <= 0 = no
 * cap (the C# default).  Burst telemetry and accretion disks do not change the
 * bodies and are not produced. */
#define CM_COLLISION_MERGE_ONLY 0
#define CM_COLLISION_BOUNCE_ONLY 1
#define CM_COLLISION_REALISTIC 2
typedef struct {
    int32_t mode;                          /* CM_COLLISION_* */
    double default_restitution;            /* 0.15 */
    double fragmentation_specific_energy_threshold;  /* 0.6 (unused by Resolve) */
    double fragmentation_mass_loss_cap;    /* 0.3 */
    double capture_velocity_factor;        /* 0.9 (unused by Resolve) */
    int32_t promote_compact_remnants;      /* Resolve's promoteCompactRemnants */
    int32_t max_merges_per_frame;          /* <= 0: no cap (engine default 64) */
    int32_t max_merges_per_body;           /* <= 0: no cap (engine default 2) */
} cm_collision_config;
int32_t cm_default_collision_config(cm_collision_config* cfg);
int32_t cm_configure_collisions(void* engine, const cm_collision_config* cfg);
/* Resolves the events of the last cm_detect_collisions with the configured
 * mode; *n_resolved = LastCollisionCount. */
int32_t cm_resolve_collisions(void* engine, int32_t* n_resolved);
/* n_steps x NBodySolver.StepSoA on the device-resident state
 * (Solvers/NBodySolver.cs:383-429): one SoAVerletIntegrator.Step with the
 * configured backend (and 1PN decorator), then, when collisions = 1,
 * CollisionDetector.Detect and CollisionResolver.Resolve with the configured
 * collision mode (cm_configure_collisions), as NBodySolver does when
 * collisions are enabled.  Like the C# path it never compacts.
 * *n_resolved (may be NULL) = resolved events summed over the steps. */
int32_t cm_nbody_step(void* engine, double softening, double dt, int32_t n_steps,
                      int32_t collisions, int32_t* n_resolved);
/* SimulationState (Types/SimulationState.cs:9-25) as NBodySolver.Step returns
 * it after each step (Solvers/NBodySolver.cs:307-379).  Energies are
 * EnergyCalculator's with compensated sums; the potential is the exact pair
 * sum (with_potential = 1), the Barnes-Hut potential of cm_energy
 * (with_potential = 2) or 0 (with_potential = 0: kinetic-only totals). */
typedef struct {
    double time;                        /* Time: dt summed over the diagnosed steps
                                           since the last cm_upload_soa */
    int32_t body_count;                 /* BodyCount */
    int32_t active_body_count;          /* ActiveBodyCount */
    double kinetic_energy;              /* KineticEnergy */
    double potential_energy;            /* PotentialEnergy */
    double total_energy;                /* TotalEnergy */
    double total_momentum[3];           /* TotalMomentum */
    double energy_drift;                /* EnergyDrift: (E - E0) / |E0|, E0 the total
                                           energy after the first diagnosed step since
                                           the last upload; 0 when |E0| <= 1e-15 */
    int32_t collision_count;            /* CollisionCount: resolved this step */
    double current_dt;                  /* CurrentDt */
    double step_energy_delta;           /* StepEnergyDelta: E before - E after */
    double step_momentum_delta;         /* StepMomentumDelta: |P before - P after| */
    double step_angular_momentum_delta; /* StepAngularMomentumDelta: ||L| before - |L| after| */
} cm_simulation_state;
/* n_steps x NBodySolver.Step (Solvers/NBodySolver.cs:307-379): cm_nbody_step's
 * device-resident step (Verlet with the configured backend, then collisions
 * when collisions = 1) with the SimulationState of every step in
 * states[n_steps].  The energy sums of each step are reduced on the device
 * into one history buffer and read back once at the end (the reference
 * computes them on the CPU twice per step). */
int32_t cm_nbody_step_diag(void* engine, double softening, double dt, int32_t n_steps,
                           int32_t collisions, int32_t with_potential,
                           cm_simulation_state* states);
/* Order-preserving removal of inactive bodies, ParticleSystem::compact
 * (engine/src/physics/particle_system.cpp:156-184).  old_index (may be NULL)
 * receives, for each survivor, its index before compaction. */
int32_t cm_compact(void* engine, int32_t* new_count, int32_t* old_index);

/* ---- 1PN relativistic correction ---------------------------------------- */
/* PostNewtonianBackend (Solvers/PostNewtonianBackend.cs:33-43), the decorator
 * NBodySolver.SelectBackend puts around any backend when post-Newtonian
 * corrections are on (Solvers/NBodySolver.cs:480-481): after the configured
 * backend's Newtonian forces, every force evaluation of cm_compute_forces,
 * cm_verlet_step and cm_yoshida4_step adds
 * PostNewtonian1Correction.ApplyCorrections
 * (Extensions/PostNewtonian1Correction.cs:171-188): for each active body i,
 * the EIH term of ComputeCorrection (:66-165) summed over every j != i with
 * m_j > 0, softened with the call's softening, in FP64 bit for bit.
 * max_velocity_fraction_c is the engine port's field
 * (engine/include/celestial/physics/post_newtonian.hpp:13); the C# class
 * always uses 0.3, which is the default here. */
typedef struct {
    int32_t enabled;
    double max_velocity_fraction_c;       /* 0.3 */
    double schwarzschild_warning_factor;  /* 3.0: pairs closer than this many
                                             Schwarzschild radii are skipped */
} cm_pn_config;
int32_t cm_default_pn_config(cm_pn_config* cfg);
int32_t cm_set_post_newtonian(void* engine, const cm_pn_config* cfg);
/* OnSchwarzschildProximityWarning (PostNewtonian1Correction.cs:47,124-129):
 * the pairs the last force evaluation skipped as strong-field, in (i, j)
 * order.  Up to 4096 pairs per evaluation are kept: the first 4096 of the
 * reference's (i ascending, j ascending) loop, also when more occur (then a
 * second pass rewrites the kept slots in loop order); *n_total counts all
 * (saturated at INT32_MAX). */
typedef struct {
    int32_t body, other;
    double distance, schwarzschild_radius;
} cm_pn_warning;
int32_t cm_pn_warnings(void* engine, cm_pn_warning* out, int32_t cap,
                       int32_t* n_total);

/* ---- softening modes of the direct sum ----------------------------------- */
/* SofteningConfig (engine/include/celestial/sim/simulation_config.hpp:13-41)
 * as Engine::compute_cpu_brute_forces applies it
 * (engine/src/sim/engine.cpp:581-660): with a mode other than GLOBAL the
 * direct backends sweep every active j != i with the pairwise softening
 * (eps_i + eps_j) / 2, where eps_k = type_softening[body_type_k mod 16]
 * (PER_TYPE) or max(adaptive_scale * cbrt(m_k), adaptive_min) (ADAPTIVE).
 * GLOBAL keeps the call's softening and the C# backends' arithmetic.  As in
 * the engine, the Barnes-Hut backend always uses the global softening.
 * Replaces celestial_set_softening_mode / _set_type_softening /
 * _set_adaptive_softening_scale (engine/src/interop/native_api.cpp:339-358). */
#define CM_SOFTENING_GLOBAL 0
#define CM_SOFTENING_PER_TYPE 1
#define CM_SOFTENING_ADAPTIVE 2
typedef struct {
    int32_t mode;                 /* CM_SOFTENING_* */
    double type_softening[16];    /* per BodyType index; engine default 0 */
    double adaptive_scale;        /* 0.01 */
    double adaptive_min;          /* 1e-8 */
} cm_softening_config;
int32_t cm_default_softening_config(cm_softening_config* cfg);
int32_t cm_set_softening(void* engine, const cm_softening_config* cfg);

/* ---- Yoshida 4th-order integrator ------------------------------------------ */
/* n_steps x the engine's IntegratorType::Yoshida4 step
 * (engine/src/sim/engine.cpp:1057-1081): three SoAVerletIntegrator.Step
 * sub-steps of W1 dt, W0 dt, W1 dt with the coefficients of
 * engine/include/celestial/sim/yoshida.hpp:18-28, forces (and the 1PN term)
 * re-evaluated in every sub-step; device-resident. */
int32_t cm_yoshida4_step(void* engine, double softening, double dt,
                         int32_t n_steps);

/* ---- batched ensemble of independent systems ---------------------------- */
/* Many independent systems (each <= 2048 bodies, concatenated; system s owns
 * bodies offsets[s] .. offsets[s+1]-1) stepped with SoAVerletIntegrator.Step
 * (Integrators/SoAVerletIntegrator.cs:48-115) and CpuParallelBackend's
 * direct sweep (Solvers/CpuParallelBackend.cs:99-136) per system.  The
 * reference steps one system per NBodySolver; this batches thousands in one
 * persistent launch replayed from a hipGraph. */
typedef struct {
    int32_t threads;             /* workgroup size in use */
    int32_t tuned_threads;       /* autotuner choice (0: not tuned) */
    int32_t tuned_cost_order;    /* 1: systems ordered by N^2 descending */
    float tuned_ms_per_step;
    float last_ms;               /* last cm_ensemble_step call, HIP events */
    int32_t last_steps;
    int64_t interactions_total;  /* pair terms evaluated since creation */
} cm_ensemble_stats;

int32_t cm_ensemble_create(int32_t n_systems, const int32_t* offsets,
                           int32_t device, void** out_ensemble);
int32_t cm_ensemble_destroy(void* ensemble);
const char* cm_ensemble_last_error(void* ensemble);
/* precision: CM_PREC_FP64 or CM_PREC_FP32_SPLIT; threads: 0 (autotune on the
 * next step), 256, 512 or 1024; steps_per_launch: steps per captured launch */
int32_t cm_ensemble_configure(void* ensemble, int32_t precision, int32_t threads,
                              int32_t steps_per_launch);
int32_t cm_ensemble_upload(void* ensemble, const cm_soa_ptrs* src);
int32_t cm_ensemble_download(void* ensemble, const cm_soa_ptrs* dst);
int32_t cm_ensemble_step(void* ensemble, double softening, double dt,
                         int32_t n_steps);
/* per-system total energy KE + PE (EnergyCalculator quantities), FP64 with
 * compensated sums; energy has n_systems entries */
int32_t cm_ensemble_energy(void* ensemble, double softening, double* energy);
int32_t cm_ensemble_get_stats(void* ensemble, cm_ensemble_stats* out);

/* ---- learned far-field correction (no reference counterpart: the
 * reference's Barnes-Hut is monopole-only, BarnesHutBackend.cs:392-459) ---- */
/* A 16 -> 64 -> 64 -> 64 -> 16 ReLU MLP (bf16 weights and activations on
 * MFMA, fp32 accumulation) evaluated per (target group, accepted cell) pair of
 * the traversal; it predicts the cell's monopole truncation residual at the
 * group centre and its gradient, and the traversal adds
 * sum over cells (residual + gradient . (x - centre)) to each body when
 * cm_backend_config.learned_correction = 1 (FP32_SPLIT only).
 * Parameters: CM_MLP_PARAMS floats, per layer W[out][in] row-major then b[out]
 * (layers 64x16, 64x64, 64x64, 16x64).  Inputs per pair: r/|r| (3), h/|r|,
 * Q/(M h^2) (xx, yy, zz, xy, xz, yz), 6 zeros, with r = cell centre of mass
 * minus group centre, h the cell half size, Q the cell's second moments about
 * its centre of mass; outputs 0..2 residual / (M h^2 / |r|^4), 3..7 gradient
 * (xx, yy, xy, xz, yz) / (M h^2 / |r|^5), 8..15 unused. */
#define CM_MLP_IN 16
#define CM_MLP_OUT 16
#define CM_MLP_USED 8
#define CM_MLP_PARAMS 10448
/* He-uniform weights (SplitMix64 from seed), zero biases; resets Adam. */
int32_t cm_mlp_init(void* engine, uint64_t seed);
int32_t cm_mlp_set_params(void* engine, const float* params, int32_t n);
int32_t cm_mlp_get_params(void* engine, float* params, int32_t n);
/* outputs y[n][16] for inputs x[n][16] */
int32_t cm_mlp_forward(void* engine, const float* x, int32_t n, float* y);
/* gradient (CM_MLP_PARAMS floats) and value of the mean squared error over
 * the 8 used outputs of one batch x[n][16], targets t[n][8] */
int32_t cm_mlp_gradient(void* engine, const float* x, const float* t, int32_t n,
                        float* grad, float* loss);
/* Training pairs from the current state: builds the tree, samples up to
 * max_groups target groups (seeded), records each accepted cell of >= 2
 * bodies (at most max_pairs) with its inputs and its FP64 direct-sum targets
 * (softened by `softening`, like the traversal).  Kept on the device. */
int32_t cm_mlp_collect(void* engine, double softening, int32_t max_groups, uint64_t seed,
                       int32_t max_pairs, int32_t* n_pairs);
/* copies the first `cap` collected pairs (each array may be NULL): inputs
 * x[.][16], targets t[.][8], pair_info[.][4] = (sample rank, node, first body,
 * end body of the target group in sorted order) and the group centre [.][3];
 * pairs are ordered by sampled group (in first-body order), then node */
int32_t cm_mlp_get_pairs(void* engine, float* x, float* t, int32_t* pair_info,
                         double* centre, int32_t cap);
/* Adam (lr, beta 0.9 / 0.999) over the collected pairs, shuffled per epoch
 * (seeded); loss_per_epoch (may be NULL) receives the mean batch loss. */
int32_t cm_mlp_train(void* engine, int32_t epochs, int32_t batch, float lr, uint64_t seed,
                     float* loss_per_epoch);

/* ---- synthetic scenarios (host-side, deterministic from a seed) ----------- */
/* Planetary systems for the ensemble: system sizes uniform in [n_min, n_max];
 * each is a 1 Msun star plus planets with log-uniform masses in
 * [m_min, m_max], log-uniform semi-major axes in [a_min, a_max], Rayleigh
 * eccentricity (sigma_e), isotropic orientation, centre-of-mass frame.
 * Call with out == NULL to get offsets (n_systems + 1 entries) only; then
 * with arrays of offsets[n_systems] bodies. */
int32_t cm_generate_ensemble(int32_t n_systems, uint64_t seed, int32_t n_min,
                             int32_t n_max, double m_min, double m_max,
                             double a_min, double a_max, double sigma_e,
                             int32_t* offsets, const cm_soa_ptrs* out);
/* Cold thin disk around a 1 Msun star (no reference generator exists for it;
 * it stands in for the reference's body setup of collision scenarios):
 * body 0 is the star (radius star_radius, density consistent with it),
 * bodies 1..n-1 have log-uniform semi-major axes in [a_min, a_max], Rayleigh
 * eccentricity / inclination, uniform angles, log-uniform masses in
 * [m_min, m_max] and radii cbrt(3 (m / density) / (4 pi))
 * (DensityModel.ComputeRadius, Types/DensityModel.cs:60-66); all active and
 * collidable, centre-of-mass frame.  Writes pos, vel, mass, radius, density,
 * is_active, is_collidable and body_type. */
int32_t cm_generate_disk(int32_t n, uint64_t seed, double star_radius,
                         double a_min, double a_max, double sigma_e,
                         double sigma_i, double m_min, double m_max,
                         double density, const cm_soa_ptrs* out);
/* PlummerSphere.Generate (ScenarioGenerators/PlummerSphere.cs:41-116) with a
 * SplitMix64 stream; mass_fraction_cut < 1 truncates the sphere at the radius
 * enclosing that mass fraction; com_frame subtracts the centre-of-mass
 * position and velocity.  Writes pos/vel/mass/radius of n bodies. */
int32_t cm_generate_plummer(int32_t n, double total_mass, double scale_length,
                            uint64_t seed, double mass_fraction_cut,
                            int32_t com_frame, const cm_soa_ptrs* out);
/* PlummerSphere.Generate(count, totalMass, scaleLength, seed)
 * (ScenarioGenerators/PlummerSphere.cs:41-116) body for body: the same
 * NextDouble stream as .NET's System.Random(seed), no truncation, no
 * centre-of-mass shift.  seed >= 0 (the reference seeds a negative seed from
 * the clock, which no caller can reproduce).  Writes pos/vel/mass/radius. */
int32_t cm_generate_plummer_dotnet(int32_t n, double total_mass, double scale_length,
                                   int32_t seed, const cm_soa_ptrs* out);

#ifdef __cplusplus
}
#endif
#endif /* CMHIP_H */
