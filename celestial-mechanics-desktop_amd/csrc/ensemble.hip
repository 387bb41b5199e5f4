// ensemble.hip -- many independent small systems of uneven size stepped
// together: direct softened gravity with the reference's Velocity Verlet
// (SoAVerletIntegrator.cs:48-115) and CpuParallelBackend's force sweep
// (Solvers/CpuParallelBackend.cs:99-136) per system.
//
// One persistent kernel launch advances every system by S steps: workgroups
// take systems from a device counter in descending cost order (N^2), keep the
// system's positions and masses in LDS for the whole launch and step it S
// times with workgroup barriers between the drift, force and kick phases --
// no grid-wide synchronisation is needed because systems never interact.
// The launch (counter reset + kernel) is captured once in a hipGraph and
// replayed for every chunk of S steps.  The workgroup size and the system
// order are chosen by a small autotuner that times the candidates on a copy
// of the state.
// FP64 reproduces the CPU loop bit for bit (-ffp-contract=off, IEEE sqrt and
// divide); FP32_SPLIT forms each pair offset in double and evaluates the rest
// of the pair term in float, folding per-target sums into double.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/cmhip.h"
#include "cm_internal.h"

namespace cm {

constexpr int kEnsMaxBodies = 2048;   // bodies per system held in LDS
constexpr int kEnsMaxParts = 16;      // source ranges per target of a small system

struct EnsArgs {
    const int32_t* offsets;
    const int32_t* order;
    uint32_t* counter;
    int n_systems;
    double *px, *py, *pz, *vx, *vy, *vz, *ax, *ay, *az, *oax, *oay, *oaz;
    const double* m;
    const uint8_t* act;
    double eps2, dt;
    int steps;
    unsigned long long* interactions;
};

template <int PREC>
__device__ __forceinline__ void pair_acc(double xi, double yi, double zi, double xj,
                                         double yj, double zj, double mj, double eps2,
                                         double& sx, double& sy, double& sz, float& fx,
                                         float& fy, float& fz) {
    double dx = xi - xj, dy = yi - yj, dz = zi - zj;
    if constexpr (PREC == CM_PREC_FP32_SPLIT) {
        float x = (float)dx, y = (float)dy, z = (float)dz;
        float r2 = __builtin_fmaf(x, x, __builtin_fmaf(y, y, __builtin_fmaf(z, z, (float)eps2)));
        float inv = __builtin_amdgcn_rsqf(r2);
        float c = (float)mj * (inv * inv * inv);
        fx = __builtin_fmaf(-c, x, fx);
        fy = __builtin_fmaf(-c, y, fy);
        fz = __builtin_fmaf(-c, z, fz);
    } else {
        double raw = dx * dx + dy * dy + dz * dz;
        double inv = 1.0 / sqrt(raw + eps2);
        double c = mj * inv * inv * inv;   // GravityKernel.cs:17-18
        sx -= c * dx; sy -= c * dy; sz -= c * dz;
    }
}

// FP32 split pair term: offset formed in double, the rest in float.
__device__ __forceinline__ void pair_f(const double4 p, const double4 q, float e2, float& fx,
                                       float& fy, float& fz) {
    const float x = (float)(p.x - q.x), y = (float)(p.y - q.y), z = (float)(p.z - q.z);
    const float r2 = __builtin_fmaf(x, x, __builtin_fmaf(y, y, __builtin_fmaf(z, z, e2)));
    const float inv = __builtin_amdgcn_rsqf(r2);
    const float c = (float)q.w * (inv * inv * inv);
    fx = __builtin_fmaf(-c, x, fx);
    fy = __builtin_fmaf(-c, y, fy);
    fz = __builtin_fmaf(-c, z, fz);
}

template <int PREC, int NT>
__global__ void __launch_bounds__(NT)
k_ens_run(EnsArgs a) {
    // the system's positions and masses (x, y, z, m; m = 0 for inactive bodies)
    __shared__ double4 s_p[kEnsMaxBodies];
    __shared__ uint8_t s_a[kEnsMaxBodies];
    __shared__ int s_sys;
    // FP32 partial sums of the source ranges of small systems: [part][slot][6]
    __shared__ float s_red[PREC == CM_PREC_FP32_SPLIT ? 6 * NT : 1];
    const int t = threadIdx.x;
    const double half = 0.5 * a.dt;
    unsigned long long work = 0ull;
    for (;;) {
        if (t == 0) s_sys = (int)atomicAdd(a.counter, 1u);
        __syncthreads();
        const int q = s_sys;
        __syncthreads();
        if (q >= a.n_systems) break;
        const int sys = a.order[q];
        const int b0 = a.offsets[sys];
        const int nb = a.offsets[sys + 1] - b0;
        for (int i = t; i < nb; i += NT) {
            const uint8_t act = a.act[b0 + i];
            s_p[i].w = act ? a.m[b0 + i] : 0.0;
            s_a[i] = act;
        }
        // final kick and rotate of target i from its acceleration
        auto finish = [&](int i, double sx, double sy, double sz) {
            const int g = b0 + i;
            a.ax[g] = sx; a.ay[g] = sy; a.az[g] = sz;
            if (s_a[i]) {
                a.vx[g] = a.vx[g] + half * sx;
                a.vy[g] = a.vy[g] + half * sy;
                a.vz[g] = a.vz[g] + half * sz;
            }
            a.oax[g] = sx; a.oay[g] = sy; a.oaz[g] = sz;
        };
        for (int step = 0; step < a.steps; step++) {
            // phases 1-2: half kick with the old acceleration, drift
            for (int i = t; i < nb; i += NT) {
                int g = b0 + i;
                double x = a.px[g], y = a.py[g], z = a.pz[g];
                if (s_a[i]) {
                    double vx = a.vx[g] + half * a.oax[g];
                    double vy = a.vy[g] + half * a.oay[g];
                    double vz = a.vz[g] + half * a.oaz[g];
                    a.vx[g] = vx; a.vy[g] = vy; a.vz[g] = vz;
                    x = x + a.dt * vx; y = y + a.dt * vy; z = z + a.dt * vz;
                    a.px[g] = x; a.py[g] = y; a.pz[g] = z;
                }
                s_p[i].x = x; s_p[i].y = y; s_p[i].z = z;
            }
            __syncthreads();
            // phase 3: forces, then phases 4-5 (second half kick, rotate)
            if constexpr (PREC == CM_PREC_FP32_SPLIT) {
                // two targets per thread share every source read; inactive
                // sources carry zero mass and, with eps2 >= kSelfTermMinEps2,
                // the self term is exactly zero, so neither needs a branch
                // (adding an exact zero leaves each float sum unchanged)
                const float e2 = (float)a.eps2;
                const bool skip_self = !(a.eps2 >= kSelfTermMinEps2);
                const int T = (nb + 1) / 2;   // target slots of two targets each
                if (T < NT) {
                    // fewer slots than threads: the idle threads take source
                    // ranges too; thread t sums targets (slot, slot + T) over
                    // range `part` of P contiguous ranges, and the P partial
                    // sums are added in range order
                    const int P = min(kEnsMaxParts, NT / T);
                    const int slot = t % T, part = t / T;
                    const int i0 = slot, i1 = slot + T;
                    const bool has1 = i1 < nb;
                    float ax0 = 0.0f, ay0 = 0.0f, az0 = 0.0f, ax1 = 0.0f, ay1 = 0.0f, az1 = 0.0f;
                    if (part < P) {
                        const double4 p0 = s_p[i0];
                        const double4 p1 = s_p[has1 ? i1 : i0];
                        const int j0 = part * nb / P, j1 = (part + 1) * nb / P;
                        if (!skip_self) {
                            for (int j = j0; j < j1; j++) {
                                const double4 qj = s_p[j];
                                pair_f(p0, qj, e2, ax0, ay0, az0);
                                pair_f(p1, qj, e2, ax1, ay1, az1);
                            }
                        } else {
                            for (int j = j0; j < j1; j++) {
                                const double4 qj = s_p[j];
                                if (j != i0) pair_f(p0, qj, e2, ax0, ay0, az0);
                                if (j != i1) pair_f(p1, qj, e2, ax1, ay1, az1);
                            }
                        }
                        float* r = &s_red[6 * (part * T + slot)];
                        r[0] = ax0; r[1] = ay0; r[2] = az0; r[3] = ax1; r[4] = ay1; r[5] = az1;
                    }
                    __syncthreads();
                    if (part == 0) {
                        for (int q = 1; q < P; q++) {
                            const float* r = &s_red[6 * (q * T + slot)];
                            ax0 += r[0]; ay0 += r[1]; az0 += r[2];
                            ax1 += r[3]; ay1 += r[4]; az1 += r[5];
                        }
                        const bool act0 = s_a[i0] != 0;
                        finish(i0, act0 ? (double)ax0 : 0.0, act0 ? (double)ay0 : 0.0,
                               act0 ? (double)az0 : 0.0);
                        if (has1) {
                            const bool act1 = s_a[i1] != 0;
                            finish(i1, act1 ? (double)ax1 : 0.0, act1 ? (double)ay1 : 0.0,
                                   act1 ? (double)az1 : 0.0);
                        }
                    }
                } else
                for (int i0 = t; i0 < nb; i0 += 2 * NT) {
                    const int i1 = i0 + NT;
                    const bool has1 = i1 < nb;
                    const double4 p0 = s_p[i0];
                    const double4 p1 = s_p[has1 ? i1 : i0];
                    float ax0 = 0.0f, ay0 = 0.0f, az0 = 0.0f, ax1 = 0.0f, ay1 = 0.0f, az1 = 0.0f;
                    if (!skip_self) {
                        for (int j = 0; j < nb; j++) {
                            const double4 qj = s_p[j];
                            pair_f(p0, qj, e2, ax0, ay0, az0);
                            pair_f(p1, qj, e2, ax1, ay1, az1);
                        }
                    } else {
                        for (int j = 0; j < nb; j++) {
                            const double4 qj = s_p[j];
                            if (j != i0) pair_f(p0, qj, e2, ax0, ay0, az0);
                            if (j != i1) pair_f(p1, qj, e2, ax1, ay1, az1);
                        }
                    }
                    const bool act0 = s_a[i0] != 0;
                    finish(i0, act0 ? (double)ax0 : 0.0, act0 ? (double)ay0 : 0.0,
                           act0 ? (double)az0 : 0.0);
                    if (has1) {
                        const bool act1 = s_a[i1] != 0;
                        finish(i1, act1 ? (double)ax1 : 0.0, act1 ? (double)ay1 : 0.0,
                               act1 ? (double)az1 : 0.0);
                    }
                }
            } else {
                // FP64: CpuParallelBackend's sweep (j ascending, self and
                // inactive skipped), bit for bit
                for (int i = t; i < nb; i += NT) {
                    double sx = 0.0, sy = 0.0, sz = 0.0;
                    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
                    if (s_a[i]) {
                        const double4 pi = s_p[i];
                        for (int j = 0; j < nb; j++) {
                            if (j == i || !s_a[j]) continue;
                            const double4 qj = s_p[j];
                            pair_acc<PREC>(pi.x, pi.y, pi.z, qj.x, qj.y, qj.z, qj.w, a.eps2,
                                           sx, sy, sz, fx, fy, fz);
                        }
                    }
                    finish(i, sx, sy, sz);
                }
            }
            __syncthreads();
        }
        if (t == 0) work += (unsigned long long)nb * (unsigned long long)(nb - 1) * a.steps;
    }
    if (t == 0 && work) atomicAdd(a.interactions, work);
}

__global__ void k_ens_reset(uint32_t* counter) { *counter = 0u; }

// Per-system KE + PE (EnergyCalculator.cs:16-81 quantities) with compensated
// sums: one workgroup per system.
__global__ void __launch_bounds__(256)
k_ens_energy(const int32_t* __restrict__ offsets, const double* __restrict__ px,
             const double* __restrict__ py, const double* __restrict__ pz,
             const double* __restrict__ vx, const double* __restrict__ vy,
             const double* __restrict__ vz, const double* __restrict__ m,
             const uint8_t* __restrict__ act, double eps2, double* __restrict__ out) {
    __shared__ double s_s[256], s_c[256];
    const int sys = blockIdx.x;
    const int b0 = offsets[sys], b1 = offsets[sys + 1];
    double s = 0.0, c = 0.0;
    auto add = [&](double x) {
        double tt = s + x;
        c += fabs(s) >= fabs(x) ? (s - tt) + x : (x - tt) + s;
        s = tt;
    };
    for (int i = b0 + threadIdx.x; i < b1; i += 256) {
        if (!act[i]) continue;
        add(0.5 * m[i] * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]));
        for (int j = i + 1; j < b1; j++) {
            if (!act[j]) continue;
            double rx = px[j] - px[i], ry = py[j] - py[i], rz = pz[j] - pz[i];
            add(-1.0 * m[i] * m[j] / sqrt(rx * rx + ry * ry + rz * rz + eps2));
        }
    }
    s_s[threadIdx.x] = s;
    s_c[threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        double S = 0.0, Cc = 0.0;
        for (int k = 0; k < 256; k++) {
            double x = s_s[k];
            double tt = S + x;
            Cc += fabs(S) >= fabs(x) ? (S - tt) + x : (x - tt) + S;
            S = tt;
            Cc += s_c[k];
        }
        out[sys] = S + Cc;
    }
}

struct Ensemble {
    int device = 0;
    hipStream_t stream = nullptr;
    int n_systems = 0;
    int n_bodies = 0;
    std::vector<int32_t> h_offsets;
    int32_t* offsets = nullptr;
    int32_t* order = nullptr;
    uint32_t* counter = nullptr;
    double* f[13] = {};   // px py pz vx vy vz ax ay az oax oay oaz m
    uint8_t* act = nullptr;
    double* energy = nullptr;
    unsigned long long* inter = nullptr;
    int precision = CM_PREC_FP32_SPLIT;
    int threads = 0;               // 0: autotune on the next step call
    int steps_per_launch = 100;
    int cost_order = 1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double graph_key[3] = {0, 0, 0};   // eps2, dt, steps of the captured graph
    std::string err;
    cm_ensemble_stats stats{};
};

static int ens_err(Ensemble& e, int code, const std::string& m) {
    e.err = m;
    return code;
}

#define ENS_TRY(e, call)                                                       \
    do {                                                                       \
        hipError_t _s = (call);                                                \
        if (_s != hipSuccess)                                                  \
            return ens_err((e), CM_ERR_HIP, std::string(#call) + ": " +        \
                                                hipGetErrorString(_s));        \
    } while (0)

static EnsArgs ens_args(Ensemble& e, double eps2, double dt, int steps, double* const* f) {
    EnsArgs a;
    a.offsets = e.offsets; a.order = e.order; a.counter = e.counter;
    a.n_systems = e.n_systems;
    a.px = f[0]; a.py = f[1]; a.pz = f[2]; a.vx = f[3]; a.vy = f[4]; a.vz = f[5];
    a.ax = f[6]; a.ay = f[7]; a.az = f[8]; a.oax = f[9]; a.oay = f[10]; a.oaz = f[11];
    a.m = f[12]; a.act = e.act; a.eps2 = eps2; a.dt = dt; a.steps = steps;
    a.interactions = e.inter;
    return a;
}

static void launch_run(Ensemble& e, const EnsArgs& a, int threads, int precision) {
    int dev_blocks = 256 * (threads <= 256 ? 4 : (threads <= 512 ? 2 : 1));
    int nb = std::max(1, std::min(e.n_systems, dev_blocks));
    hipLaunchKernelGGL(k_ens_reset, dim3(1), dim3(1), 0, e.stream, e.counter);
#define ENS_CASE(P, T)                                                                   \
    if (precision == P && threads == T)                                                  \
        hipLaunchKernelGGL((k_ens_run<P, T>), dim3(nb), dim3(T), 0, e.stream, a);
    ENS_CASE(CM_PREC_FP64, 256) ENS_CASE(CM_PREC_FP64, 512) ENS_CASE(CM_PREC_FP64, 1024)
    ENS_CASE(CM_PREC_FP32_SPLIT, 256) ENS_CASE(CM_PREC_FP32_SPLIT, 512)
    ENS_CASE(CM_PREC_FP32_SPLIT, 1024)
#undef ENS_CASE
}

static int set_order(Ensemble& e, bool by_cost) {
    std::vector<int32_t> ord(e.n_systems);
    std::iota(ord.begin(), ord.end(), 0);
    if (by_cost) {
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
            long long na = e.h_offsets[a + 1] - e.h_offsets[a];
            long long nb = e.h_offsets[b + 1] - e.h_offsets[b];
            return na > nb;
        });
    }
    ENS_TRY(e, hipMemcpy(e.order, ord.data(), ord.size() * 4, hipMemcpyHostToDevice));
    return CM_OK;
}

// Autotuner: times one launch of 2 steps per (workgroup size, order)
// candidate on a scratch copy of the state and keeps the fastest.
static int autotune(Ensemble& e, double eps2, double dt) {
    double* scratch[13];
    size_t b = (size_t)e.n_bodies * 8;
    for (int k = 0; k < 13; k++) {
        ENS_TRY(e, hipMalloc(&scratch[k], std::max<size_t>(b, 8)));
        ENS_TRY(e, hipMemcpy(scratch[k], e.f[k], b, hipMemcpyDeviceToDevice));
    }
    hipEvent_t t0, t1;
    ENS_TRY(e, hipEventCreate(&t0));
    ENS_TRY(e, hipEventCreate(&t1));
    float best = 1e30f;
    int best_t = 256, best_o = 1;
    for (int order = 1; order >= 0; order--) {
        int rc = set_order(e, order == 1);
        if (rc != CM_OK) return rc;
        for (int threads : {256, 512, 1024}) {
            EnsArgs a = ens_args(e, eps2, dt, 2, scratch);
            launch_run(e, a, threads, e.precision);   // warm
            ENS_TRY(e, hipEventRecord(t0, e.stream));
            launch_run(e, a, threads, e.precision);
            ENS_TRY(e, hipEventRecord(t1, e.stream));
            ENS_TRY(e, hipEventSynchronize(t1));
            float ms = 0.0f;
            ENS_TRY(e, hipEventElapsedTime(&ms, t0, t1));
            if (ms < best) { best = ms; best_t = threads; best_o = order; }
        }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    for (int k = 0; k < 13; k++) (void)hipFree(scratch[k]);
    e.threads = best_t;
    e.cost_order = best_o;
    e.stats.tuned_threads = best_t;
    e.stats.tuned_cost_order = best_o;
    e.stats.tuned_ms_per_step = best / 2.0f;
    return set_order(e, best_o == 1);
}

static void drop_graph(Ensemble& e) {
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
    if (e.graph) (void)hipGraphDestroy(e.graph);
    e.exec = nullptr;
    e.graph = nullptr;
}

}  // namespace cm

using cm::Ensemble;

extern "C" {

int32_t cm_ensemble_create(int32_t n_systems, const int32_t* offsets, int32_t device,
                           void** out) {
    if (!out || n_systems <= 0 || !offsets) return CM_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (offsets[0] != 0) return CM_ERR_INVALID_ARGUMENT;
    for (int s = 0; s < n_systems; s++) {
        int nb = offsets[s + 1] - offsets[s];
        if (nb < 1 || nb > cm::kEnsMaxBodies) return CM_ERR_INVALID_ARGUMENT;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return CM_ERR_DEVICE_NOT_FOUND;
    Ensemble* e = new (std::nothrow) Ensemble();
    if (!e) return CM_ERR_OUT_OF_MEMORY;
    e->device = device;
    e->n_systems = n_systems;
    e->n_bodies = offsets[n_systems];
    e->h_offsets.assign(offsets, offsets + n_systems + 1);
    bool ok = hipSetDevice(device) == hipSuccess &&
              hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) == hipSuccess;
    size_t nbod = (size_t)e->n_bodies;
    auto alloc = [&](void** p, size_t bytes) {
        if (ok && hipMalloc(p, std::max<size_t>(bytes, 256)) != hipSuccess) ok = false;
    };
    alloc((void**)&e->offsets, (size_t)(n_systems + 1) * 4);
    alloc((void**)&e->order, (size_t)n_systems * 4);
    alloc((void**)&e->counter, 4);
    for (int k = 0; k < 13; k++) alloc((void**)&e->f[k], nbod * 8);
    alloc((void**)&e->act, nbod);
    alloc((void**)&e->energy, (size_t)n_systems * 8);
    alloc((void**)&e->inter, 8);
    if (ok) ok = hipMemcpy(e->offsets, offsets, (size_t)(n_systems + 1) * 4,
                           hipMemcpyHostToDevice) == hipSuccess;
    if (ok) ok = hipMemset(e->inter, 0, 8) == hipSuccess;
    if (ok) ok = cm::set_order(*e, true) == CM_OK;
    if (!ok) {
        cm_ensemble_destroy(e);
        return CM_ERR_OUT_OF_MEMORY;
    }
    *out = e;
    return CM_OK;
}

int32_t cm_ensemble_destroy(void* h) {
    Ensemble* e = static_cast<Ensemble*>(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    cm::drop_graph(*e);
    void* ptrs[] = {e->offsets, e->order, e->counter, e->act, e->energy, e->inter};
    for (void* p : ptrs) if (p) (void)hipFree(p);
    for (int k = 0; k < 13; k++) if (e->f[k]) (void)hipFree(e->f[k]);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return CM_OK;
}

const char* cm_ensemble_last_error(void* h) {
    Ensemble* e = static_cast<Ensemble*>(h);
    return e ? e->err.c_str() : "null ensemble handle";
}

int32_t cm_ensemble_configure(void* h, int32_t precision, int32_t threads,
                              int32_t steps_per_launch) {
    Ensemble* e = static_cast<Ensemble*>(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (precision != CM_PREC_FP64 && precision != CM_PREC_FP32_SPLIT)
        return cm::ens_err(*e, CM_ERR_INVALID_ARGUMENT, "precision must be FP64 or FP32_SPLIT");
    if (threads != 0 && threads != 256 && threads != 512 && threads != 1024)
        return cm::ens_err(*e, CM_ERR_INVALID_ARGUMENT, "threads must be 0, 256, 512 or 1024");
    if (steps_per_launch < 1)
        return cm::ens_err(*e, CM_ERR_INVALID_ARGUMENT, "steps_per_launch must be >= 1");
    e->precision = precision;
    e->threads = threads;
    e->steps_per_launch = steps_per_launch;
    cm::drop_graph(*e);
    return CM_OK;
}

int32_t cm_ensemble_upload(void* h, const cm_soa_ptrs* src) {
    Ensemble* e = static_cast<Ensemble*>(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!src) return cm::ens_err(*e, CM_ERR_INVALID_ARGUMENT, "null source");
    ENS_TRY(*e, hipSetDevice(e->device));
    double* const p[13] = {src->pos_x, src->pos_y, src->pos_z, src->vel_x, src->vel_y,
                           src->vel_z, src->acc_x, src->acc_y, src->acc_z, src->old_acc_x,
                           src->old_acc_y, src->old_acc_z, src->mass};
    size_t b = (size_t)e->n_bodies * 8;
    for (int k = 0; k < 13; k++) {
        const double* from = p[k];
        if (!from && k >= 9 && k <= 11) from = p[k - 3];   // BodySoA.CopyFrom seeds OldAcc
        if (from) ENS_TRY(*e, hipMemcpy(e->f[k], from, b, hipMemcpyHostToDevice));
        else ENS_TRY(*e, hipMemset(e->f[k], 0, b));
    }
    if (src->is_active)
        ENS_TRY(*e, hipMemcpy(e->act, src->is_active, e->n_bodies, hipMemcpyHostToDevice));
    else
        ENS_TRY(*e, hipMemset(e->act, 1, e->n_bodies));
    return CM_OK;
}

int32_t cm_ensemble_download(void* h, const cm_soa_ptrs* dst) {
    Ensemble* e = static_cast<Ensemble*>(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!dst) return cm::ens_err(*e, CM_ERR_INVALID_ARGUMENT, "null destination");
    ENS_TRY(*e, hipSetDevice(e->device));
    ENS_TRY(*e, hipStreamSynchronize(e->stream));
    double* const p[13] = {dst->pos_x, dst->pos_y, dst->pos_z, dst->vel_x, dst->vel_y,
                           dst->vel_z, dst->acc_x, dst->acc_y, dst->acc_z, dst->old_acc_x,
                           dst->old_acc_y, dst->old_acc_z, dst->mass};
    size_t b = (size_t)e->n_bodies * 8;
    for (int k = 0; k < 13; k++)
        if (p[k]) ENS_TRY(*e, hipMemcpy(p[k], e->f[k], b, hipMemcpyDeviceToHost));
    if (dst->is_active)
        ENS_TRY(*e, hipMemcpy(dst->is_active, e->act, e->n_bodies, hipMemcpyDeviceToHost));
    return CM_OK;
}

int32_t cm_ensemble_step(void* h, double softening, double dt, int32_t n_steps) {
    Ensemble* e = static_cast<Ensemble*>(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (n_steps < 0 || !(softening >= 0.0))
        return cm::ens_err(*e, CM_ERR_INVALID_ARGUMENT, "bad step arguments");
    ENS_TRY(*e, hipSetDevice(e->device));
    double eps2 = softening * softening;
    if (e->threads == 0) {
        int rc = cm::autotune(*e, eps2, dt);
        if (rc != CM_OK) return rc;
        cm::drop_graph(*e);
    }
    hipEvent_t t0, t1;
    ENS_TRY(*e, hipEventCreate(&t0));
    ENS_TRY(*e, hipEventCreate(&t1));
    ENS_TRY(*e, hipEventRecord(t0, e->stream));
    int done = 0;
    const int chunk = e->steps_per_launch;
    if (n_steps >= chunk) {
        if (!e->exec || e->graph_key[0] != eps2 || e->graph_key[1] != dt ||
            e->graph_key[2] != (double)chunk) {
            cm::drop_graph(*e);
            ENS_TRY(*e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            cm::EnsArgs a = cm::ens_args(*e, eps2, dt, chunk, e->f);
            cm::launch_run(*e, a, e->threads, e->precision);
            ENS_TRY(*e, hipStreamEndCapture(e->stream, &e->graph));
            ENS_TRY(*e, hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
            e->graph_key[0] = eps2; e->graph_key[1] = dt; e->graph_key[2] = chunk;
        }
        for (; done + chunk <= n_steps; done += chunk)
            ENS_TRY(*e, hipGraphLaunch(e->exec, e->stream));
    }
    if (done < n_steps) {
        cm::EnsArgs a = cm::ens_args(*e, eps2, dt, n_steps - done, e->f);
        cm::launch_run(*e, a, e->threads, e->precision);
        ENS_TRY(*e, hipGetLastError());
    }
    ENS_TRY(*e, hipEventRecord(t1, e->stream));
    ENS_TRY(*e, hipEventSynchronize(t1));
    float ms = 0.0f;
    ENS_TRY(*e, hipEventElapsedTime(&ms, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    unsigned long long inter = 0;
    ENS_TRY(*e, hipMemcpy(&inter, e->inter, 8, hipMemcpyDeviceToHost));
    e->stats.last_ms = ms;
    e->stats.last_steps = n_steps;
    e->stats.interactions_total = (int64_t)inter;
    e->stats.threads = e->threads;
    return CM_OK;
}

int32_t cm_ensemble_energy(void* h, double softening, double* energy) {
    Ensemble* e = static_cast<Ensemble*>(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!energy) return cm::ens_err(*e, CM_ERR_INVALID_ARGUMENT, "null output");
    ENS_TRY(*e, hipSetDevice(e->device));
    hipLaunchKernelGGL(cm::k_ens_energy, dim3(e->n_systems), dim3(256), 0, e->stream,
                       e->offsets, e->f[0], e->f[1], e->f[2], e->f[3], e->f[4], e->f[5],
                       e->f[12], e->act, softening * softening, e->energy);
    ENS_TRY(*e, hipGetLastError());
    ENS_TRY(*e, hipStreamSynchronize(e->stream));
    ENS_TRY(*e, hipMemcpy(energy, e->energy, (size_t)e->n_systems * 8, hipMemcpyDeviceToHost));
    return CM_OK;
}

int32_t cm_ensemble_get_stats(void* h, cm_ensemble_stats* out) {
    Ensemble* e = static_cast<Ensemble*>(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!out) return cm::ens_err(*e, CM_ERR_INVALID_ARGUMENT, "null output");
    *out = e->stats;
    return CM_OK;
}

}  // extern "C"
