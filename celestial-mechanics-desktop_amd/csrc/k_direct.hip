// k_direct.hip -- FP64 direct-sum force kernels.
//
// Each kernel gives every target body its own thread and streams the sources
// through LDS in tiles of 256 in ascending index order, so every body's sum
// is accumulated in exactly the order of the reference's CPU loops:
//   k_direct_pairwise: CpuSingleThreadBackend, Newton-3 branch
//     (Solvers/CpuSingleThreadBackend.cs:79-118).  Body k first receives
//     +(m_i inv^3) (x_i - x_k) from every active i < k in ascending i (the
//     ax[j] += factor_ji * dx updates), then adds its own row sum over j > k.
//   k_direct_sweep: CpuParallelBackend (Solvers/CpuParallelBackend.cs:99-136)
//     and the shell-theorem loop of CpuSingleThreadBackend (:124-158) with
//     GravityKernel.AccelerationCoeffFromSource (Solvers/GravityKernel.cs:9-33).
//   k_direct_softened: the engine's brute force with per-type or adaptive
//     softening (engine/src/sim/engine.cpp:587-650).
// Built with -ffp-contract=off and IEEE division / square root: results are
// bit-identical to the reference's FP64 arithmetic.
#include <hip/hip_runtime.h>

#include <vector>

#include "cm_internal.h"
#include "cm_math.h"

namespace cm {

constexpr int kTile = 256;

__device__ __forceinline__ double gravity_coeff(double dist_sq, double m,
                                                double r, double eps2,
                                                bool shell) {
    if (!shell || r <= 0.0) {
        double inv = 1.0 / sqrt(dist_sq + eps2);
        return m * inv * inv * inv;
    }
    double dist = sqrt(dist_sq);
    if (dist < r) {
        double denom = r * r + eps2;
        double inv_denom_sqrt = 1.0 / sqrt(denom);
        return m * inv_denom_sqrt / denom;
    }
    double inv = 1.0 / sqrt(dist_sq + eps2);
    return m * inv * inv * inv;
}

struct SrcTile {
    double x[kTile], y[kTile], z[kTile], m[kTile], r[kTile];
    uint8_t a[kTile];
};

__device__ __forceinline__ void load_tile(SrcTile& t, int base, int n,
                                          const double* px, const double* py,
                                          const double* pz, const double* m,
                                          const double* rad, const uint8_t* act) {
    int j = base + threadIdx.x;
    if (j < n) {
        t.x[threadIdx.x] = px[j];
        t.y[threadIdx.x] = py[j];
        t.z[threadIdx.x] = pz[j];
        t.m[threadIdx.x] = m[j];
        t.r[threadIdx.x] = rad ? rad[j] : 0.0;
        t.a[threadIdx.x] = act[j];
    } else {
        t.a[threadIdx.x] = 0;
    }
}

__global__ void __launch_bounds__(kTile)
k_direct_pairwise(int n, double eps2, const double* __restrict__ px,
                  const double* __restrict__ py, const double* __restrict__ pz,
                  const double* __restrict__ m, const uint8_t* __restrict__ act,
                  double* __restrict__ ax, double* __restrict__ ay,
                  double* __restrict__ az) {
    __shared__ SrcTile t;
    int k = blockIdx.x * kTile + threadIdx.x;
    bool me = k < n && act[k];
    double xk = 0, yk = 0, zk = 0, mk = 0;
    if (k < n) { xk = px[k]; yk = py[k]; zk = pz[k]; mk = m[k]; }
    double rx = 0.0, ry = 0.0, rz = 0.0;   // received from i < k
    double sx = 0.0, sy = 0.0, sz = 0.0;   // own row over j > k
    for (int base = 0; base < n; base += kTile) {
        __syncthreads();
        load_tile(t, base, n, px, py, pz, m, nullptr, act);
        __syncthreads();
        if (!me) continue;
        int lim = min(kTile, n - base);
        for (int q = 0; q < lim; q++) {
            int j = base + q;
            if (j == k || !t.a[q]) continue;
            if (j < k) {
                double dx = t.x[q] - xk;
                double dy = t.y[q] - yk;
                double dz = t.z[q] - zk;
                double d2 = dx * dx + dy * dy + dz * dz + eps2;
                double inv = 1.0 / sqrt(d2);
                double inv3 = inv * inv * inv;
                double f = t.m[q] * inv3;
                rx += f * dx; ry += f * dy; rz += f * dz;
            } else {
                double dx = xk - t.x[q];
                double dy = yk - t.y[q];
                double dz = zk - t.z[q];
                double d2 = dx * dx + dy * dy + dz * dz + eps2;
                double inv = 1.0 / sqrt(d2);
                double inv3 = inv * inv * inv;
                double f = t.m[q] * inv3;
                sx -= f * dx; sy -= f * dy; sz -= f * dz;
            }
        }
    }
    (void)mk;
    if (k < n) {
        if (me) { ax[k] = rx + sx; ay[k] = ry + sy; az[k] = rz + sz; }
        else { ax[k] = 0.0; ay[k] = 0.0; az[k] = 0.0; }
    }
}

__global__ void __launch_bounds__(kTile)
k_direct_sweep(int n, double eps2, int shell, const double* __restrict__ px,
               const double* __restrict__ py, const double* __restrict__ pz,
               const double* __restrict__ m, const double* __restrict__ rad,
               const uint8_t* __restrict__ act, double* __restrict__ ax,
               double* __restrict__ ay, double* __restrict__ az) {
    __shared__ SrcTile t;
    int i = blockIdx.x * kTile + threadIdx.x;
    bool me = i < n && act[i];
    double xi = 0, yi = 0, zi = 0;
    if (i < n) { xi = px[i]; yi = py[i]; zi = pz[i]; }
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int base = 0; base < n; base += kTile) {
        __syncthreads();
        load_tile(t, base, n, px, py, pz, m, rad, act);
        __syncthreads();
        if (!me) continue;
        int lim = min(kTile, n - base);
        for (int q = 0; q < lim; q++) {
            int j = base + q;
            if (j == i || !t.a[q]) continue;
            double dx = xi - t.x[q];
            double dy = yi - t.y[q];
            double dz = zi - t.z[q];
            double raw = dx * dx + dy * dy + dz * dz;
            double c = gravity_coeff(raw, t.m[q], t.r[q], eps2, shell != 0);
            sx -= c * dx; sy -= c * dy; sz -= c * dz;
        }
    }
    if (i < n) { ax[i] = sx; ay[i] = sy; az[i] = sz; }
}

// Engine::compute_cpu_brute_forces with a SofteningMode other than Global
// (engine/src/sim/engine.cpp:587-650): active j != i in ascending j, pairwise
// eps = (eps_i + eps_j) * 0.5 and the engine's operation order
// (inv3 = inv*inv*inv, factor = m_j * inv3).  eps_k is the per-type table
// entry (type mod 16, as gravity_kernel.cu:131 masks it) or
// max(scale * cbrt(m_k), min) with a correctly rounded cube root; each tile
// computes its sources' eps once.
struct SoftTable { double eps[16]; };

__device__ __forceinline__ double soft_eps(int mode, const SoftTable& tb, double scale,
                                           double mn, double mass, int32_t type) {
    if (mode == CM_SOFTENING_PER_TYPE) return tb.eps[type & 15];
    double e = scale * cbrt_rn(mass);
    return e < mn ? mn : e;   // std::max(e, mn)
}

__global__ void __launch_bounds__(kTile)
k_direct_softened(int n, int mode, SoftTable tb, double scale, double mn,
                  const double* __restrict__ px, const double* __restrict__ py,
                  const double* __restrict__ pz, const double* __restrict__ m,
                  const uint8_t* __restrict__ act, const int32_t* __restrict__ type,
                  double* __restrict__ ax, double* __restrict__ ay,
                  double* __restrict__ az) {
    __shared__ SrcTile t;   // t.r holds the sources' eps
    int i = blockIdx.x * kTile + threadIdx.x;
    bool me = i < n && act[i];
    double xi = 0, yi = 0, zi = 0, ei = 0;
    if (me) {
        xi = px[i]; yi = py[i]; zi = pz[i];
        ei = soft_eps(mode, tb, scale, mn, m[i], type[i]);
    }
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int base = 0; base < n; base += kTile) {
        __syncthreads();
        int j = base + threadIdx.x;
        if (j < n) {
            t.x[threadIdx.x] = px[j];
            t.y[threadIdx.x] = py[j];
            t.z[threadIdx.x] = pz[j];
            t.m[threadIdx.x] = m[j];
            t.a[threadIdx.x] = act[j];
            t.r[threadIdx.x] = act[j] ? soft_eps(mode, tb, scale, mn, m[j], type[j]) : 0.0;
        } else {
            t.a[threadIdx.x] = 0;
        }
        __syncthreads();
        if (!me) continue;
        int lim = min(kTile, n - base);
        for (int q = 0; q < lim; q++) {
            if (base + q == i || !t.a[q]) continue;
            double e = (ei + t.r[q]) * 0.5;
            double eps2 = e * e;
            double dx = xi - t.x[q];
            double dy = yi - t.y[q];
            double dz = zi - t.z[q];
            double d2 = dx * dx + dy * dy + dz * dz + eps2;
            double inv = 1.0 / sqrt(d2);
            double inv3 = inv * inv * inv;
            double f = t.m[q] * inv3;
            sx -= f * dx; sy -= f * dy; sz -= f * dz;
        }
    }
    if (i < n) { ax[i] = sx; ay[i] = sy; az[i] = sz; }
}

// Judge sums for sampled targets: target t, source chunk c -> partial.
// Partials over chunks of kSampleChunk sources are folded in chunk order, so
// the result is deterministic (it differs from a single sequential sweep
// only by the rounding of the chunk partition).
constexpr int kSampleChunk = 16384;

__global__ void __launch_bounds__(256)
k_direct_sample(int n, int nt, double eps2, const int32_t* __restrict__ tg,
                const double* __restrict__ px, const double* __restrict__ py,
                const double* __restrict__ pz, const double* __restrict__ m,
                const uint8_t* __restrict__ act, double* __restrict__ part) {
    __shared__ SrcTile t;
    int ti = blockIdx.x * 256 + threadIdx.x;
    int chunk = blockIdx.y;
    int i = ti < nt ? tg[ti] : -1;
    bool me = i >= 0 && act[i];
    double xi = 0, yi = 0, zi = 0;
    if (i >= 0) { xi = px[i]; yi = py[i]; zi = pz[i]; }
    double sx = 0.0, sy = 0.0, sz = 0.0;
    int c0 = chunk * kSampleChunk, c1 = min(n, c0 + kSampleChunk);
    for (int base = c0; base < c1; base += kTile) {
        __syncthreads();
        load_tile(t, base, c1, px, py, pz, m, nullptr, act);
        __syncthreads();
        if (!me) continue;
        int lim = min(kTile, c1 - base);
        for (int q = 0; q < lim; q++) {
            int j = base + q;
            if (j == i || !t.a[q]) continue;
            double dx = xi - t.x[q];
            double dy = yi - t.y[q];
            double dz = zi - t.z[q];
            double raw = dx * dx + dy * dy + dz * dz;
            double inv = 1.0 / sqrt(raw + eps2);
            double c = t.m[q] * inv * inv * inv;
            sx -= c * dx; sy -= c * dy; sz -= c * dz;
        }
    }
    if (ti < nt) {
        size_t o = ((size_t)chunk * nt + ti) * 3;
        part[o] = sx; part[o + 1] = sy; part[o + 2] = sz;
    }
}

__global__ void k_sample_fold(int nt, int nchunks, const double* __restrict__ part,
                              double* __restrict__ out) {
    int ti = blockIdx.x * blockDim.x + threadIdx.x;
    if (ti >= nt) return;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int c = 0; c < nchunks; c++) {
        size_t o = ((size_t)c * nt + ti) * 3;
        sx += part[o]; sy += part[o + 1]; sz += part[o + 2];
    }
    out[3 * ti] = sx; out[3 * ti + 1] = sy; out[3 * ti + 2] = sz;
}

int launch_direct(Engine& e, double softening) {
    int n = e.count;
    if (n == 0) return CM_OK;
    double eps2 = softening * softening;
    int nb = (n + kTile - 1) / kTile;
    if (e.detail_timing()) CM_HIP_TRY(e, hipEventRecord(e.ev[0], e.stream));
    e.mark(0);
    e.mark(1);
    bool pairwise = e.cfg.kind == CM_BACKEND_DIRECT_PAIRWISE && !e.cfg.shell_theorem;
    if (e.soft.mode != CM_SOFTENING_GLOBAL) {
        SoftTable tb;
        for (int k = 0; k < 16; k++) tb.eps[k] = e.soft.type_softening[k];
        hipLaunchKernelGGL(k_direct_softened, dim3(nb), dim3(kTile), 0, e.stream, n,
                           e.soft.mode, tb, e.soft.adaptive_scale, e.soft.adaptive_min,
                           e.f[PX], e.f[PY], e.f[PZ], e.f[MASS], e.act, e.type,
                           e.f[AX], e.f[AY], e.f[AZ]);
    } else if (pairwise) {
        hipLaunchKernelGGL(k_direct_pairwise, dim3(nb), dim3(kTile), 0, e.stream,
                           n, eps2, e.f[PX], e.f[PY], e.f[PZ], e.f[MASS], e.act,
                           e.f[AX], e.f[AY], e.f[AZ]);
    } else {
        hipLaunchKernelGGL(k_direct_sweep, dim3(nb), dim3(kTile), 0, e.stream, n,
                           eps2, e.cfg.shell_theorem, e.f[PX], e.f[PY], e.f[PZ],
                           e.f[MASS], e.f[RAD], e.act, e.f[AX], e.f[AY], e.f[AZ]);
    }
    CM_HIP_TRY(e, hipGetLastError());
    if (e.detail_timing()) CM_HIP_TRY(e, hipEventRecord(e.ev[5], e.stream));
    e.mark(2);
    e.stats = cm_force_stats{};
    e.stats.n_active = -1;
    e.stats.body_interactions = -1;
    e.stats.cell_interactions = 0;
    return CM_OK;
}

int launch_direct_sample(Engine& e, double softening, const int32_t* targets,
                         int32_t nt, double* ax, double* ay, double* az) {
    if (nt == 0) return CM_OK;
    int n = e.count;
    int nchunks = (n + kSampleChunk - 1) / kSampleChunk;
    if (nchunks == 0) nchunks = 1;
    size_t part_bytes = (size_t)nchunks * nt * 3 * 8;
    double* part = nullptr;
    double* out = nullptr;
    int32_t* tg = nullptr;
    CM_HIP_TRY(e, hipMalloc(&part, part_bytes));
    CM_HIP_TRY(e, hipMalloc(&out, (size_t)nt * 3 * 8));
    CM_HIP_TRY(e, hipMalloc(&tg, (size_t)nt * 4));
    int rc = CM_OK;
    do {
        if (hipMemcpyAsync(tg, targets, (size_t)nt * 4, hipMemcpyHostToDevice, e.stream) != hipSuccess) { rc = CM_ERR_HIP; break; }
        dim3 grid((nt + 255) / 256, nchunks);
        hipLaunchKernelGGL(k_direct_sample, grid, dim3(256), 0, e.stream, n, nt,
                           softening * softening, tg, e.f[PX], e.f[PY], e.f[PZ],
                           e.f[MASS], e.act, part);
        hipLaunchKernelGGL(k_sample_fold, dim3((nt + 255) / 256), dim3(256), 0,
                           e.stream, nt, nchunks, part, out);
        if (hipGetLastError() != hipSuccess) { rc = CM_ERR_HIP; break; }
        std::vector<double> h((size_t)nt * 3);
        if (hipMemcpyAsync(h.data(), out, h.size() * 8, hipMemcpyDeviceToHost, e.stream) != hipSuccess) { rc = CM_ERR_HIP; break; }
        if (hipStreamSynchronize(e.stream) != hipSuccess) { rc = CM_ERR_HIP; break; }
        for (int t = 0; t < nt; t++) {
            ax[t] = h[3 * t]; ay[t] = h[3 * t + 1]; az[t] = h[3 * t + 2];
        }
    } while (0);
    (void)hipFree(part);
    (void)hipFree(out);
    (void)hipFree(tg);
    if (rc != CM_OK) return set_error(e, rc, "direct sample kernel failed");
    return CM_OK;
}

}  // namespace cm
