// k_diag.hip -- energy / momentum diagnostics with compensated FP64 sums.
//
// Quantities of EnergyCalculator (Validation/EnergyCalculator.cs:16-81):
// KE = sum 0.5 m |v|^2, P = sum v m, L = sum r x (v m) over active bodies and
// PE = -sum_{i<j} m_i m_j / sqrt(|r_j - r_i|^2 + eps^2)
// (Forces/NewtonianGravity.cs:68-77).  Every partial sum is carried as a
// (sum, compensation) pair and pairs are merged with Neumaier's two-sum rule
// (the device-side counterpart of the reference's Kahan finalisation,
// engine/src/cuda/gpu_reduction.cu:21-39), in a fixed launch shape, so the
// results are deterministic and accurate to a few ulps of the exact sums.
// PE is formed per body as a full row over j != i (tiled through LDS) and
// halved at the end, or (with_pe = 2) from each body's Barnes-Hut potential,
// 0.5 m_i phi_i, as the engine's launch_gpu_potential_energy does
// (engine/src/cuda/potential_energy_kernel.cu:134-171).  The same pass sums
// m x, m y, m z for the centre of mass of the engine's GPU energy snapshot
// (k_reduce_com, engine/src/cuda/gpu_reduction.cu:232-262, finalised as
// gpu_reduce_com does at :481-517).  k_max_accel / k_max_final are
// Engine::compute_max_acceleration (engine/src/sim/engine.cpp:468-500,
// k_reduce_max_accel at gpu_reduction.cu:269-283): the largest |a| over active
// bodies, 0 when there are none.  A maximum does not depend on the order in
// which it is taken, so that result is exact.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "cm_internal.h"

namespace cm {

constexpr int kRedBlocks = 512;
constexpr int kQ = 13;   // ke, px, py, pz, lx, ly, lz, mass, pe, count, mx, my, mz

struct Comp {
    double s, c;
};

__device__ __forceinline__ void comp_add(Comp& a, double x) {
    double t = a.s + x;
    double err = fabs(a.s) >= fabs(x) ? (a.s - t) + x : (x - t) + a.s;
    a.c += err;
    a.s = t;
}

__device__ __forceinline__ Comp comp_merge(Comp a, Comp b) {
    double t = a.s + b.s;
    double err = fabs(a.s) >= fabs(b.s) ? (a.s - t) + b.s : (b.s - t) + a.s;
    Comp r;
    r.s = t;
    r.c = a.c + b.c + err;
    return r;
}

__device__ void block_reduce_store(Comp (&v)[kQ], Comp* __restrict__ out) {
    __shared__ double ss[kQ][256];
    __shared__ double sc[kQ][256];
    int t = threadIdx.x;
    for (int q = 0; q < kQ; q++) { ss[q][t] = v[q].s; sc[q][t] = v[q].c; }
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) {
            for (int q = 0; q < kQ; q++) {
                Comp a{ss[q][t], sc[q][t]}, b{ss[q][t + o], sc[q][t + o]};
                Comp r = comp_merge(a, b);
                ss[q][t] = r.s; sc[q][t] = r.c;
            }
        }
        __syncthreads();
    }
    if (t == 0)
        for (int q = 0; q < kQ; q++) out[(size_t)q * kRedBlocks + blockIdx.x] = Comp{ss[q][0], sc[q][0]};
}

__global__ void __launch_bounds__(256)
k_energy_bodies(int n, const double* __restrict__ px, const double* __restrict__ py,
                const double* __restrict__ pz, const double* __restrict__ vx,
                const double* __restrict__ vy, const double* __restrict__ vz,
                const double* __restrict__ m, const uint8_t* __restrict__ act,
                const double* __restrict__ pe_rows, Comp* __restrict__ out) {
    Comp v[kQ];
    for (int q = 0; q < kQ; q++) v[q] = Comp{0.0, 0.0};
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        if (!act[i]) continue;
        double mi = m[i];
        double x = vx[i], y = vy[i], z = vz[i];
        comp_add(v[0], 0.5 * mi * (x * x + y * y + z * z));
        double mvx = x * mi, mvy = y * mi, mvz = z * mi;
        comp_add(v[1], mvx); comp_add(v[2], mvy); comp_add(v[3], mvz);
        double rx = px[i], ry = py[i], rz = pz[i];
        comp_add(v[4], ry * mvz - rz * mvy);
        comp_add(v[5], rz * mvx - rx * mvz);
        comp_add(v[6], rx * mvy - ry * mvx);
        comp_add(v[7], mi);
        if (pe_rows) comp_add(v[8], pe_rows[i]);
        v[9].s += 1.0;
        comp_add(v[10], mi * rx); comp_add(v[11], mi * ry); comp_add(v[12], mi * rz);
    }
    block_reduce_store(v, out);
}

__global__ void __launch_bounds__(256)
k_energy_final(int nblocks, const Comp* __restrict__ part, Comp* __restrict__ tmp,
               double* __restrict__ res) {
    Comp v[kQ];
    for (int q = 0; q < kQ; q++) {
        v[q] = Comp{0.0, 0.0};
        for (int b = threadIdx.x; b < nblocks; b += 256) v[q] = comp_merge(v[q], part[(size_t)q * kRedBlocks + b]);
    }
    block_reduce_store(v, tmp);
    __syncthreads();
    if (threadIdx.x == 0)
        for (int q = 0; q < kQ; q++) {
            Comp c = tmp[(size_t)q * kRedBlocks];
            res[q] = c.s + c.c;
        }
}

// Row sums of the pair potential: row_i = sum_{j != i, active} -m_i m_j / d_ij.
__global__ void __launch_bounds__(256)
k_pe_rows(int n, double eps2, const double* __restrict__ px,
          const double* __restrict__ py, const double* __restrict__ pz,
          const double* __restrict__ m, const uint8_t* __restrict__ act,
          double* __restrict__ rows) {
    __shared__ double tx[256], ty[256], tz[256], tm[256];
    int i = blockIdx.x * 256 + threadIdx.x;
    bool me = i < n && act[i];
    double xi = 0, yi = 0, zi = 0, mi = 0;
    if (i < n) { xi = px[i]; yi = py[i]; zi = pz[i]; mi = m[i]; }
    Comp r{0.0, 0.0};
    for (int base = 0; base < n; base += 256) {
        __syncthreads();
        int j = base + threadIdx.x;
        if (j < n) {
            tx[threadIdx.x] = px[j]; ty[threadIdx.x] = py[j]; tz[threadIdx.x] = pz[j];
            tm[threadIdx.x] = act[j] ? m[j] : 0.0;
        } else {
            tm[threadIdx.x] = 0.0;
            tx[threadIdx.x] = 0.0; ty[threadIdx.x] = 0.0; tz[threadIdx.x] = 0.0;
        }
        __syncthreads();
        if (!me) continue;
        int lim = min(256, n - base);
        double s = 0.0;
        for (int q = 0; q < lim; q++) {
            if (base + q == i || tm[q] == 0.0) continue;
            double dx = tx[q] - xi, dy = ty[q] - yi, dz = tz[q] - zi;
            double d = sqrt(dx * dx + dy * dy + dz * dz + eps2);
            s += -1.0 * mi * tm[q] / d;
        }
        comp_add(r, s);
    }
    if (i < n) rows[i] = me ? 0.5 * (r.s + r.c) : 0.0;
}

// Per-body shares of the potential energy (device array of count doubles):
// method 1 the exact row sums, method 2 the Barnes-Hut potential.
static int launch_rows(Engine& e, double softening, int method, double* rows) {
    const int n = e.count;
    if (method == 1) {
        hipLaunchKernelGGL(k_pe_rows, dim3((n + 255) / 256), dim3(256), 0, e.stream, n,
                           softening * softening, e.f[PX], e.f[PY], e.f[PZ],
                           e.f[MASS], e.act, rows);
        CM_HIP_TRY(e, hipGetLastError());
        return CM_OK;
    }
    return launch_bh_potential(e, softening, rows);
}

// A tree that outgrew the node pool yields no potential.
static int check_pool(Engine& e) {
    int32_t over = 0;
    CM_HIP_TRY(e, hipMemcpyAsync(&over, &e.hdr->overflow, 4, hipMemcpyDeviceToHost, e.stream));
    CM_HIP_TRY(e, hipStreamSynchronize(e.stream));
    return over ? set_error(e, CM_ERR_POOL_EXHAUSTED, "node pool exhausted") : CM_OK;
}

int potential_rows(Engine& e, double softening, int method, double* host_rows) {
    const int n = e.count;
    if (n == 0) return CM_OK;
    double* rows = nullptr;
    CM_HIP_TRY(e, hipMalloc(&rows, (size_t)n * 8));
    int rc = launch_rows(e, softening, method, rows);
    if (rc == CM_OK && method == 2) rc = check_pool(e);
    if (rc == CM_OK) {
        hipError_t err = hipMemcpyAsync(host_rows, rows, (size_t)n * 8, hipMemcpyDeviceToHost,
                                        e.stream);
        if (err == hipSuccess) err = hipStreamSynchronize(e.stream);
        if (err != hipSuccess) rc = set_error(e, CM_ERR_HIP, hipGetErrorString(err));
    }
    (void)hipFree(rows);
    return rc;
}

// KE, P, L, mass, COM (and PE from per-body rows when with_pe > 0) with the
// per-body rows in `rows` (count doubles) and the reduction partials in `red`
// (kEnergyRedDoubles doubles); both device buffers owned by the caller.
constexpr size_t kEnergyRedDoubles = 4 * (size_t)kQ * kRedBlocks + 2 * kQ + 64;
static_assert(kEnergyRedDoubles <= 64 * 1024, "energy partials fit the engine's diagnostics scratch");

// the kQ sums of one evaluation into dev_out (device, kQ doubles), no host sync
static int energy_sums_async(Engine& e, double softening, int with_pe, double* rows, double* red,
                             double* dev_out) {
    const int n = e.count;
    hipStream_t st = e.stream;
    if (with_pe && n > 0) {
        int rc = launch_rows(e, softening, with_pe, rows);
        if (rc == CM_OK && with_pe == 2) rc = check_pool(e);
        if (rc != CM_OK) return rc;
    }
    Comp* part = reinterpret_cast<Comp*>(red);
    Comp* tmp = part + (size_t)kQ * kRedBlocks;
    hipLaunchKernelGGL(k_energy_bodies, dim3(kRedBlocks), dim3(256), 0, st, n,
                       e.f[PX], e.f[PY], e.f[PZ], e.f[VX], e.f[VY], e.f[VZ],
                       e.f[MASS], e.act, with_pe ? rows : nullptr, part);
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, st, kRedBlocks, part, tmp, dev_out);
    CM_HIP_TRY(e, hipGetLastError());
    return CM_OK;
}

static void fill_report(const double* h, int with_pe, cm_energy_report* out) {
    out->kinetic = h[0];
    out->momentum[0] = h[1]; out->momentum[1] = h[2]; out->momentum[2] = h[3];
    out->angular_momentum[0] = h[4];
    out->angular_momentum[1] = h[5];
    out->angular_momentum[2] = h[6];
    out->angular_momentum_mag = sqrt(h[4] * h[4] + h[5] * h[5] + h[6] * h[6]);
    out->total_mass = h[7];
    out->potential = with_pe ? h[8] : 0.0;
    out->active_count = (int32_t)h[9];
    for (int k = 0; k < 3; k++) out->center_of_mass[k] = h[7] > 0.0 ? h[10 + k] / h[7] : 0.0;
}

static int energy_into(Engine& e, double softening, int with_pe, double* rows, double* red,
                       cm_energy_report* out) {
    double* res = red + 4 * (size_t)kQ * kRedBlocks;
    int rc = energy_sums_async(e, softening, with_pe, rows, red, res);
    if (rc != CM_OK) return rc;
    double h[kQ] = {0};
    hipError_t err = hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, e.stream);
    if (err == hipSuccess) err = hipStreamSynchronize(e.stream);
    if (err != hipSuccess) return set_error(e, CM_ERR_HIP, hipGetErrorString(err));
    fill_report(h, with_pe, out);
    return CM_OK;
}

int launch_energy(Engine& e, double softening, int with_pe, cm_energy_report* out) {
    const int n = e.count;
    double* rows = nullptr;
    if (with_pe && n > 0) CM_HIP_TRY(e, hipMalloc(&rows, (size_t)n * 8));
    int rc = energy_into(e, softening, with_pe, rows, e.red, out);
    if (rows) (void)hipFree(rows);
    return rc;
}

// ---- the allocation-free variant (cm_energy_noalloc) ------------------------
// Its per-body potential rows and reduction partials live in this file-scope
// device array (one copy per GPU, held until the process ends).  Host code
// never passes the array's name to a kernel: it reads the array's device
// address once per device with hipGetSymbolAddress (noalloc_scratch_init, at
// cm_create) and uses that address.  Every call on one GPU shares the array.
constexpr size_t kNoAllocRowDoubles = (size_t)kNoAllocMaxBodies;
static __device__ __align__(256) double scratch[kNoAllocRowDoubles + kEnergyRedDoubles];

static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

int noalloc_scratch_init(Engine& e) {
    if (e.device < 0 || e.device >= kMaxDevices)
        return set_error(e, CM_ERR_DEVICE_NOT_FOUND, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[e.device]) {
        void* p = nullptr;
        hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !p)
            return set_error(e, CM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                                 hipGetErrorString(err));
        g_scratch_ptr[e.device] = static_cast<double*>(p);
    }
    e.noalloc_scratch = g_scratch_ptr[e.device];
    return CM_OK;
}

int launch_energy_noalloc(Engine& e, double softening, int with_pe, cm_energy_report* out) {
    if (!e.noalloc_scratch)
        return set_error(e, CM_ERR_NOT_INITIALIZED, "scratch address not initialised");
    if (with_pe && e.count > kNoAllocMaxBodies)
        return set_error(e, CM_ERR_INVALID_ARGUMENT,
                         "cm_energy_noalloc with a potential holds at most " +
                             std::to_string(kNoAllocMaxBodies) + " bodies (count " +
                             std::to_string(e.count) + ")");
    double* rows = e.noalloc_scratch;
    double* red = e.noalloc_scratch + kNoAllocRowDoubles;
    return energy_into(e, softening, with_pe, rows, red, out);
}

__global__ void __launch_bounds__(256)
k_max_accel(int n, const double* __restrict__ ax, const double* __restrict__ ay,
            const double* __restrict__ az, const uint8_t* __restrict__ act,
            double* __restrict__ block_max) {
    __shared__ double sm[256];
    double best = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        if (!act[i]) continue;
        double x = ax[i], y = ay[i], z = az[i];
        best = fmax(best, sqrt(x * x + y * y + z * z));
    }
    sm[threadIdx.x] = best;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) block_max[blockIdx.x] = sm[0];
}

__global__ void __launch_bounds__(256)
k_max_final(int nblocks, double* __restrict__ block_max) {
    __shared__ double sm[256];
    double best = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += 256) best = fmax(best, block_max[b]);
    sm[threadIdx.x] = best;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) block_max[nblocks] = sm[0];
}

int max_acceleration(Engine& e, double* out) {
    const int n = e.count;
    if (n == 0) {
        *out = 0.0;
        return CM_OK;
    }
    hipLaunchKernelGGL(k_max_accel, dim3(kRedBlocks), dim3(256), 0, e.stream, n, e.f[AX],
                       e.f[AY], e.f[AZ], e.act, e.red);
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(256), 0, e.stream, kRedBlocks, e.red);
    CM_HIP_TRY(e, hipGetLastError());
    CM_HIP_TRY(e, hipMemcpyAsync(out, e.red + kRedBlocks, 8, hipMemcpyDeviceToHost, e.stream));
    CM_HIP_TRY(e, hipStreamSynchronize(e.stream));
    return CM_OK;
}

// ---- NBodySolver.Step with its SimulationState ---------------------------------
// n_steps x (StepSoA, then the diagnostics NBodySolver.Step computes around it,
// Solvers/NBodySolver.cs:307-379).  The energy sums of the state before the
// first step and after every step are reduced on the device into one history
// buffer (no host sync between steps unless collisions are on); the states
// are formed on the host at the end.
int nbody_step_diag(Engine& e, double softening, double dt, int n_steps, int collisions,
                    int with_pe, int (*verlet_one)(Engine&, double, double),
                    cm_simulation_state* states) {
    const int n = e.count;
    double* hist = nullptr;
    double* rows = nullptr;
    if (hipMalloc(&hist, (size_t)(n_steps + 1) * kQ * 8) != hipSuccess)
        return set_error(e, CM_ERR_OUT_OF_MEMORY, "diagnostics history");
    if (with_pe && n > 0 && hipMalloc(&rows, (size_t)n * 8) != hipSuccess) {
        (void)hipFree(hist);
        return set_error(e, CM_ERR_OUT_OF_MEMORY, "potential rows");
    }
    std::vector<int32_t> resolved(n_steps, 0);
    int rc = energy_sums_async(e, softening, with_pe, rows, e.red, hist);
    for (int k = 0; k < n_steps && rc == CM_OK; k++) {
        rc = verlet_one(e, softening, dt);
        if (rc == CM_OK && collisions) {
            int32_t ne = 0;
            rc = collide_detect(e, nullptr, 0, &ne);
            if (rc == CM_OK) rc = collide_resolve(e, e.coll_cfg, &resolved[k]);
        }
        if (rc == CM_OK)
            rc = energy_sums_async(e, softening, with_pe, rows, e.red, hist + (size_t)(k + 1) * kQ);
    }
    std::vector<double> h((size_t)(n_steps + 1) * kQ);
    if (rc == CM_OK) {
        hipError_t err = hipMemcpyAsync(h.data(), hist, h.size() * 8, hipMemcpyDeviceToHost, e.stream);
        if (err == hipSuccess) err = hipStreamSynchronize(e.stream);
        if (err != hipSuccess) rc = set_error(e, CM_ERR_HIP, hipGetErrorString(err));
    }
    (void)hipFree(hist);
    if (rows) (void)hipFree(rows);
    if (rc != CM_OK) return rc;
    for (int k = 0; k < n_steps; k++) {
        cm_energy_report b{}, a{};
        fill_report(&h[(size_t)k * kQ], with_pe, &b);
        fill_report(&h[(size_t)(k + 1) * kQ], with_pe, &a);
        const double eb = b.kinetic + b.potential, ea = a.kinetic + a.potential;
        e.sim_time += dt;
        if (!e.initial_energy_set) {   // NBodySolver.cs:345-349
            e.initial_energy = ea;
            e.initial_energy_set = true;
        }
        cm_simulation_state& s = states[k];
        s = cm_simulation_state{};
        s.time = e.sim_time;
        s.body_count = n;
        s.active_body_count = a.active_count;
        s.kinetic_energy = a.kinetic;
        s.potential_energy = a.potential;
        s.total_energy = ea;
        for (int q = 0; q < 3; q++) s.total_momentum[q] = a.momentum[q];
        s.energy_drift = fabs(e.initial_energy) > 1e-15
                             ? (ea - e.initial_energy) / fabs(e.initial_energy) : 0.0;
        s.collision_count = resolved[k];
        s.current_dt = dt;
        s.step_energy_delta = eb - ea;
        const double dpx = b.momentum[0] - a.momentum[0], dpy = b.momentum[1] - a.momentum[1],
                     dpz = b.momentum[2] - a.momentum[2];
        s.step_momentum_delta = sqrt(dpx * dpx + dpy * dpy + dpz * dpz);
        s.step_angular_momentum_delta = fabs(b.angular_momentum_mag - a.angular_momentum_mag);
    }
    return CM_OK;
}

// ---- ForceErrorAnalyzer.CompareForces ----------------------------------------
// Per-body relative error |a_test - a_exact| / |a_exact| (the reference's
// arithmetic, ForceErrorAnalyzer.cs:98-116); -1 marks a body the analyzer
// skips (inactive, or |a_exact| < 1e-30).
__global__ void __launch_bounds__(256)
k_rel_error(int n, const uint8_t* __restrict__ act, const double* __restrict__ ex,
            const double* __restrict__ ey, const double* __restrict__ ez,
            const double* __restrict__ tx, const double* __restrict__ ty,
            const double* __restrict__ tz, double* __restrict__ rel) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double r = -1.0;
    if (act[i]) {
        const double x = ex[i], y = ey[i], z = ez[i];
        const double mag = sqrt(x * x + y * y + z * z);
        if (!(mag < 1e-30)) {
            const double dx = tx[i] - x, dy = ty[i] - y, dz = tz[i] - z;
            r = sqrt(dx * dx + dy * dy + dz * dz) / mag;
        }
    }
    rel[i] = r;
}

// mean / max / RMS accumulated in body order as the reference's loop does
// (ForceErrorAnalyzer.cs:97-124), plus the median
static void error_statistics(std::vector<double>& errs, cm_force_comparison* out) {
    double sum = 0.0, sum2 = 0.0, mx = 0.0;
    for (double r : errs) {
        sum += r;
        sum2 += r * r;
        if (r > mx) mx = r;
    }
    const size_t c = errs.size();
    out->body_count = (int32_t)c;
    out->mean_relative_error = c ? sum / (double)c : 0.0;
    out->rms_relative_error = c ? sqrt(sum2 / (double)c) : 0.0;
    out->max_relative_error = mx;
    out->median_relative_error = 0.0;
    if (c) {
        std::sort(errs.begin(), errs.end());
        out->median_relative_error = (c & 1) ? errs[c / 2] : 0.5 * (errs[c / 2 - 1] + errs[c / 2]);
    }
}

int compare_forces(Engine& e, double softening, double theta, const int32_t* targets,
                   int32_t nt, cm_force_comparison* out) {
    const int n = e.count;
    *out = cm_force_comparison{};
    out->theta = theta;
    if (n == 0) return CM_OK;
    const cm_backend_config saved = e.cfg;
    const cm_softening_config saved_soft = e.soft;
    const cm_pn_config saved_pn = e.pn;
    // the analyzer's plain backends: global softening, no shell theorem, no 1PN
    e.soft.mode = CM_SOFTENING_GLOBAL;
    e.pn.enabled = 0;
    e.cfg.shell_theorem = 0;
    double* ex = nullptr;
    double* rel = nullptr;
    std::vector<double> hx, hy, hz;
    auto restore = [&](int rc) {
        e.cfg = saved;
        e.soft = saved_soft;
        e.pn = saved_pn;
        if (ex) (void)hipFree(ex);
        if (rel) (void)hipFree(rel);
        return rc;
    };
    int rc = CM_OK;
    if (nt == 0) {
        // exact forces of every body: the FP64 direct sum (CpuParallelBackend
        // order; CpuSingleThreadBackend's pair terms in a per-body order)
        if (hipMalloc(&ex, (size_t)n * 3 * 8) != hipSuccess || hipMalloc(&rel, (size_t)n * 8) != hipSuccess)
            return restore(set_error(e, CM_ERR_OUT_OF_MEMORY, "compare_forces buffers"));
        e.cfg.kind = CM_BACKEND_DIRECT;
        rc = launch_direct(e, softening);
        for (int k = 0; k < 3 && rc == CM_OK; k++)
            if (hipMemcpyAsync(ex + (size_t)k * n, e.f[AX + k], (size_t)n * 8,
                               hipMemcpyDeviceToDevice, e.stream) != hipSuccess)
                rc = set_error(e, CM_ERR_HIP, "copy of the exact forces");
    } else {
        hx.resize(nt); hy.resize(nt); hz.resize(nt);
        rc = launch_direct_sample(e, softening, targets, nt, hx.data(), hy.data(), hz.data());
    }
    if (rc != CM_OK) return restore(rc);
    // test forces: the configured Barnes-Hut backend at the analyzer's theta
    e.cfg.kind = CM_BACKEND_BARNES_HUT;
    e.cfg.theta = theta;
    rc = begin_run(e);
    if (rc == CM_OK) rc = launch_barnes_hut(e, softening);
    if (rc == CM_OK) rc = finalize_stats(e);
    if (rc == CM_OK && (e.stats.flags & 1))
        rc = set_error(e, CM_ERR_POOL_EXHAUSTED, "octree node pool exhausted");
    if (rc != CM_OK) return restore(rc);
    std::vector<double> errs;
    if (nt == 0) {
        hipLaunchKernelGGL(k_rel_error, dim3((n + 255) / 256), dim3(256), 0, e.stream, n, e.act,
                           ex, ex + n, ex + 2 * (size_t)n, e.f[AX], e.f[AY], e.f[AZ], rel);
        std::vector<double> h(n);
        hipError_t err = hipGetLastError();
        if (err == hipSuccess)
            err = hipMemcpyAsync(h.data(), rel, (size_t)n * 8, hipMemcpyDeviceToHost, e.stream);
        if (err == hipSuccess) err = hipStreamSynchronize(e.stream);
        if (err != hipSuccess) return restore(set_error(e, CM_ERR_HIP, hipGetErrorString(err)));
        errs.reserve(n);
        for (int i = 0; i < n; i++)
            if (h[i] >= 0.0) errs.push_back(h[i]);
    } else {
        std::vector<double> ax(n), ay(n), az(n);
        std::vector<uint8_t> act(n);
        hipError_t err = hipSuccess;
        double* dst[3] = {ax.data(), ay.data(), az.data()};
        for (int k = 0; k < 3 && err == hipSuccess; k++)
            err = hipMemcpyAsync(dst[k], e.f[AX + k], (size_t)n * 8, hipMemcpyDeviceToHost, e.stream);
        if (err == hipSuccess)
            err = hipMemcpyAsync(act.data(), e.act, (size_t)n, hipMemcpyDeviceToHost, e.stream);
        if (err == hipSuccess) err = hipStreamSynchronize(e.stream);
        if (err != hipSuccess) return restore(set_error(e, CM_ERR_HIP, hipGetErrorString(err)));
        errs.reserve(nt);
        for (int t = 0; t < nt; t++) {
            const int i = targets[t];
            if (!act[i]) continue;
            const double x = hx[t], y = hy[t], z = hz[t];
            const double mag = sqrt(x * x + y * y + z * z);
            if (mag < 1e-30) continue;
            const double dx = ax[i] - x, dy = ay[i] - y, dz = az[i] - z;
            errs.push_back(sqrt(dx * dx + dy * dy + dz * dz) / mag);
        }
    }
    error_statistics(errs, out);
    out->theta = theta;
    return restore(CM_OK);
}

}  // namespace cm
