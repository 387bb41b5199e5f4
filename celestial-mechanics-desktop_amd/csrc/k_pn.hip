// k_pn.hip -- the 1PN (Einstein-Infeld-Hoffmann) correction of
// PostNewtonian1Correction (src/CelestialMechanics.Physics/Extensions/
// PostNewtonian1Correction.cs:66-188), applied after the Newtonian forces as
// PostNewtonianBackend does (Solvers/PostNewtonianBackend.cs:33-43).
//
// One thread per target body, sources streamed through LDS in tiles of 256 in
// ascending index order, so each body's correction is summed in the
// reference's j order; FP64 with IEEE sqrt and division and no contraction,
// so the result is the reference's bit for bit.  The pair term needs ~80
// double operations plus a square root and a division: the kernel is bound
// by the FP64 VALU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "cm_internal.h"

namespace cm {

constexpr int kPnTile = 256;
constexpr int kPnWarnCap = 4096;   // strong-field pairs kept per evaluation

struct PnBuf {
    cm_pn_warning* warn = nullptr;
    unsigned long long* count = nullptr;   // all strong-field pairs (64-bit: n^2 may pass 2^31)
    int32_t* rowcnt = nullptr;   // strong-field pairs of each row i (capacity entries)
    int32_t* rowoff = nullptr;   // their exclusive prefix sums, after an overflow
};

// PhysicalConstants (src/CelestialMechanics.Math/PhysicalConstants.cs:12-50):
// C_Sim = C * TimeUnit / AU, folded in double as C# folds the constant.
constexpr double kCSi = 2.998e8;
constexpr double kTimeUnit = 5.0226e6;
constexpr double kAu = 1.496e11;

struct PnArgs {
    int n;
    double eps2, inv_c2, max_vel2, rs_factor, warn_factor;
    const double *px, *py, *pz, *vx, *vy, *vz, *m;
    const uint8_t* act;
    double *ax, *ay, *az;
    cm_pn_warning* warn;
    unsigned long long* count;
    int32_t* rowcnt;
};

struct PnTile {
    double x[kPnTile], y[kPnTile], z[kPnTile];
    double vx[kPnTile], vy[kPnTile], vz[kPnTile];
    double m[kPnTile], v2[kPnTile];
};

__global__ void __launch_bounds__(kPnTile) k_pn(PnArgs a) {
    __shared__ PnTile t;
    const int i = blockIdx.x * kPnTile + threadIdx.x;
    const bool me = i < a.n && a.act[i];
    double xi = 0.0, yi = 0.0, zi = 0.0, vxi = 0.0, vyi = 0.0, vzi = 0.0, mi = 0.0;
    if (me) {
        xi = a.px[i]; yi = a.py[i]; zi = a.pz[i];
        vxi = a.vx[i]; vyi = a.vy[i]; vzi = a.vz[i];
        mi = a.m[i];
    }
    const double vi2 = vxi * vxi + vyi * vyi + vzi * vzi;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    int nwarn = 0;   // this row's strong-field pairs
    for (int base = 0; base < a.n; base += kPnTile) {
        __syncthreads();
        const int j = base + threadIdx.x;
        if (j < a.n) {
            const double vx = a.vx[j], vy = a.vy[j], vz = a.vz[j];
            t.x[threadIdx.x] = a.px[j];
            t.y[threadIdx.x] = a.py[j];
            t.z[threadIdx.x] = a.pz[j];
            t.vx[threadIdx.x] = vx;
            t.vy[threadIdx.x] = vy;
            t.vz[threadIdx.x] = vz;
            t.m[threadIdx.x] = a.m[j];
            t.v2[threadIdx.x] = vx * vx + vy * vy + vz * vz;   // vj2 of :118
        }
        __syncthreads();
        if (!me) continue;
        const int lim = min(kPnTile, a.n - base);
        // Every pair term is computed; the reference's four skips (j == i,
        // m_j <= 0, relative speed above the cut, strong field) only select
        // whether it is added, so the loop has no divergent branches and the
        // sums are the reference's bit for bit.
#pragma unroll 2
        for (int q = 0; q < lim; q++) {
            const int jj = base + q;
            const double mj = t.m[q];
            const double dx = t.x[q] - xi, dy = t.y[q] - yi, dz = t.z[q] - zi;
            const double dist2 = dx * dx + dy * dy + dz * dz + a.eps2;
            const double dist = sqrt(dist2);
            const double inv = 1.0 / dist;
            const double vxj = t.vx[q], vyj = t.vy[q], vzj = t.vz[q];
            const double dvx = vxi - vxj, dvy = vyi - vyj, dvz = vzi - vzj;
            const double vrel2 = dvx * dvx + dvy * dvy + dvz * dvz;
            const double rs = a.rs_factor * (mi + mj);
            const bool live = jj != i && !(mj <= 0.0) && !(vrel2 > a.max_vel2);
            const bool strong = dist < a.warn_factor * rs;
            if (live && strong) {
                nwarn++;
                const unsigned long long k = atomicAdd(a.count, 1ull);
                if (k < (unsigned long long)kPnWarnCap) {
                    cm_pn_warning w;
                    w.body = i; w.other = jj; w.distance = dist; w.schwarzschild_radius = rs;
                    a.warn[k] = w;
                }
            }
            const double nx = dx * inv, ny = dy * inv, nz = dz * inv;
            const double vj2 = t.v2[q];
            const double vivj = vxi * vxj + vyi * vyj + vzi * vzj;
            const double nvi = nx * vxi + ny * vyi + nz * vzi;
            const double nvj = nx * vxj + ny * vyj + nz * vzj;
            const double gmi_r = 1.0 * mi * inv;   // G_Sim = 1
            const double gmj_r = 1.0 * mj * inv;
            const double pre = gmj_r * inv * a.inv_c2;
            const double t1 = -vi2 - 2.0 * vj2 + 4.0 * vivj + 1.5 * nvj * nvj
                              + 5.0 * gmi_r + 4.0 * gmj_r;
            const double t2 = 4.0 * nvi - 3.0 * nvj;
            const bool add = live && !strong;
            const double cx = sx + pre * (t1 * nx + t2 * dvx);
            const double cy = sy + pre * (t1 * ny + t2 * dvy);
            const double cz = sz + pre * (t1 * nz + t2 * dvz);
            sx = add ? cx : sx;
            sy = add ? cy : sy;
            sz = add ? cz : sz;
        }
    }
    if (me) {
        a.ax[i] += sx;
        a.ay[i] += sy;
        a.az[i] += sz;
    }
    if (i < a.n) a.rowcnt[i] = nwarn;
}

// More than kPnWarnCap strong-field pairs: the slots hold whichever pairs won
// the atomic counter.  The two kernels below (each a no-op otherwise) replace
// them by the first kPnWarnCap pairs of the reference's (i ascending,
// j ascending) loop: the exclusive prefix sums of the per-row counts give
// each row its first slot, and every row that starts below the cap recomputes
// its strong-field pairs in j order with k_pn's arithmetic (the same doubles)
// and writes them from there.
constexpr int kPnScanThreads = 1024;

__global__ void __launch_bounds__(kPnScanThreads)
k_pn_row_offsets(int n, const unsigned long long* __restrict__ count,
                 const int32_t* __restrict__ rowcnt, int32_t* __restrict__ rowoff) {
    if (*count <= (unsigned long long)kPnWarnCap) return;
    __shared__ long long s_w[kPnScanThreads / 64];
    __shared__ long long s_carry;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) s_carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += kPnScanThreads) {
        const int i = base + t;
        const long long v = i < n ? rowcnt[i] : 0;
        long long x = v;   // inclusive scan within the wave
        for (int o = 1; o < 64; o <<= 1) {
            const long long y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) s_w[w] = x;
        __syncthreads();
        long long before = s_carry;
        for (int k = 0; k < w; k++) before += s_w[k];
        // rows starting at or past the cap need no offset: clamp to it
        if (i < n) rowoff[i] = (int32_t)min(before + x - v, (long long)kPnWarnCap);
        __syncthreads();
        if (t == 0) {
            long long tot = 0;
            for (int k = 0; k < kPnScanThreads / 64; k++) tot += s_w[k];
            s_carry += tot;
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kPnTile)
k_pn_row_warnings(PnArgs a, const int32_t* __restrict__ rowoff) {
    if (*a.count <= (unsigned long long)kPnWarnCap) return;
    const int i = blockIdx.x * kPnTile + threadIdx.x;
    if (i >= a.n || a.rowcnt[i] <= 0 || rowoff[i] < 0 || rowoff[i] >= kPnWarnCap) return;
    const int first = rowoff[i];
    const int room = min(a.rowcnt[i], kPnWarnCap - first);
    const double xi = a.px[i], yi = a.py[i], zi = a.pz[i];
    const double vxi = a.vx[i], vyi = a.vy[i], vzi = a.vz[i];
    const double mi = a.m[i];
    int k = 0;
    for (int j = 0; j < a.n && k < room; j++) {
        const double mj = a.m[j];
        const double dx = a.px[j] - xi, dy = a.py[j] - yi, dz = a.pz[j] - zi;
        const double dist2 = dx * dx + dy * dy + dz * dz + a.eps2;
        const double dist = sqrt(dist2);
        const double dvx = vxi - a.vx[j], dvy = vyi - a.vy[j], dvz = vzi - a.vz[j];
        const double vrel2 = dvx * dvx + dvy * dvy + dvz * dvz;
        const double rs = a.rs_factor * (mi + mj);
        const bool live = j != i && !(mj <= 0.0) && !(vrel2 > a.max_vel2);
        if (live && dist < a.warn_factor * rs) {
            cm_pn_warning w;
            w.body = i; w.other = j; w.distance = dist; w.schwarzschild_radius = rs;
            a.warn[first + k] = w;
            k++;
        }
    }
}

static PnBuf* pn_buf(Engine& e) { return static_cast<PnBuf*>(e.pn_buf); }

void pn_free(Engine& e) {
    PnBuf* b = pn_buf(e);
    if (!b) return;
    if (b->warn) (void)hipFree(b->warn);
    if (b->count) (void)hipFree(b->count);
    if (b->rowcnt) (void)hipFree(b->rowcnt);
    if (b->rowoff) (void)hipFree(b->rowoff);
    delete b;
    e.pn_buf = nullptr;
}

static int pn_ensure(Engine& e) {
    if (e.pn_buf) return CM_OK;
    PnBuf* b = new (std::nothrow) PnBuf();
    if (!b) return set_error(e, CM_ERR_OUT_OF_MEMORY, "1PN buffers");
    e.pn_buf = b;
    const size_t rows = (size_t)std::max(e.capacity, 1) * 4;
    if (hipMalloc(&b->warn, sizeof(cm_pn_warning) * kPnWarnCap) != hipSuccess ||
        hipMalloc(&b->count, 8) != hipSuccess || hipMalloc(&b->rowcnt, rows) != hipSuccess ||
        hipMalloc(&b->rowoff, rows) != hipSuccess) {
        pn_free(e);
        return set_error(e, CM_ERR_OUT_OF_MEMORY, "1PN buffers");
    }
    CM_HIP_TRY(e, hipMemsetAsync(b->count, 0, 8, e.stream));
    return CM_OK;
}

int launch_pn(Engine& e, double softening) {
    if (!e.pn.enabled) return CM_OK;
    int rc = pn_ensure(e);
    if (rc != CM_OK) return rc;
    PnBuf* b = pn_buf(e);
    CM_HIP_TRY(e, hipMemsetAsync(b->count, 0, 8, e.stream));
    if (e.count == 0) return CM_OK;
    const double c_sim = kCSi * kTimeUnit / kAu;
    const double c_sim2 = c_sim * c_sim;
    const double max_vel = e.pn.max_velocity_fraction_c * c_sim;
    PnArgs a;
    a.n = e.count;
    a.eps2 = softening * softening;
    a.inv_c2 = 1.0 / c_sim2;
    a.max_vel2 = max_vel * max_vel;
    a.rs_factor = 2.0 * 1.0 / c_sim2;   // SchwarzschildFactorSim = 2 G / c^2
    a.warn_factor = e.pn.schwarzschild_warning_factor;
    a.px = e.f[PX]; a.py = e.f[PY]; a.pz = e.f[PZ];
    a.vx = e.f[VX]; a.vy = e.f[VY]; a.vz = e.f[VZ];
    a.m = e.f[MASS];
    a.act = e.act;
    a.ax = e.f[AX]; a.ay = e.f[AY]; a.az = e.f[AZ];
    a.warn = b->warn;
    a.count = b->count;
    a.rowcnt = b->rowcnt;
    const int nb = (e.count + kPnTile - 1) / kPnTile;
    hipLaunchKernelGGL(k_pn, dim3(nb), dim3(kPnTile), 0, e.stream, a);
    hipLaunchKernelGGL(k_pn_row_offsets, dim3(1), dim3(kPnScanThreads), 0, e.stream, a.n,
                       b->count, b->rowcnt, b->rowoff);
    hipLaunchKernelGGL(k_pn_row_warnings, dim3(nb), dim3(kPnTile), 0, e.stream, a, b->rowoff);
    CM_HIP_TRY(e, hipGetLastError());
    return CM_OK;
}

}  // namespace cm

using cm::Engine;

extern "C" {

int32_t cm_default_pn_config(cm_pn_config* cfg) {
    if (!cfg) return CM_ERR_INVALID_ARGUMENT;
    cfg->enabled = 0;
    cfg->max_velocity_fraction_c = 0.3;       // PostNewtonian1Correction.cs:29
    cfg->schwarzschild_warning_factor = 3.0;  // :36
    return CM_OK;
}

int32_t cm_set_post_newtonian(void* h, const cm_pn_config* cfg) {
    Engine* e = static_cast<Engine*>(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!cfg) return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "null 1PN config");
    if (cfg->enabled != 0 && cfg->enabled != 1)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "enabled must be 0 or 1");
    if (!(cfg->max_velocity_fraction_c >= 0.0) || !(cfg->schwarzschild_warning_factor >= 0.0))
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT,
                             "1PN velocity fraction and warning factor must be >= 0");
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    e->pn = *cfg;
    if (cfg->enabled) return cm::pn_ensure(*e);
    // no evaluation adds the correction any more: forget the last warnings
    if (cm::PnBuf* b = cm::pn_buf(*e))
        CM_HIP_TRY(*e, hipMemsetAsync(b->count, 0, 8, e->stream));
    return CM_OK;
}

int32_t cm_pn_warnings(void* h, cm_pn_warning* out, int32_t cap, int32_t* n_total) {
    Engine* e = static_cast<Engine*>(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!n_total || cap < 0 || (cap > 0 && !out))
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "bad warning buffer");
    *n_total = 0;
    cm::PnBuf* b = cm::pn_buf(*e);
    if (!b) return CM_OK;
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    unsigned long long total = 0;
    CM_HIP_TRY(*e, hipMemcpyAsync(&total, b->count, 8, hipMemcpyDeviceToHost, e->stream));
    CM_HIP_TRY(*e, hipStreamSynchronize(e->stream));
    const int kept = (int)std::min(total, (unsigned long long)cm::kPnWarnCap);
    std::vector<cm_pn_warning> w((size_t)kept);
    if (kept > 0) {
        CM_HIP_TRY(*e, hipMemcpyAsync(w.data(), b->warn, sizeof(cm_pn_warning) * kept,
                                      hipMemcpyDeviceToHost, e->stream));
        CM_HIP_TRY(*e, hipStreamSynchronize(e->stream));
    }
    // the reference reports them from its (i ascending, j ascending) loops
    std::sort(w.begin(), w.end(), [](const cm_pn_warning& x, const cm_pn_warning& y) {
        return x.body != y.body ? x.body < y.body : x.other < y.other;
    });
    for (int k = 0; k < std::min(kept, (int)cap); k++) out[k] = w[k];
    *n_total = (int32_t)std::min(total, (unsigned long long)INT32_MAX);
    return CM_OK;
}

}  // extern "C"
