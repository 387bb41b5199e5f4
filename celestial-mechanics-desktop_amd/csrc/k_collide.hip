// k_collide.hip -- collision detection, MergeOnly resolution and compaction.
//
// Detection (CollisionDetector.Detect, Collisions/CollisionDetector.cs:35-135):
//   * cell size max(2 * max radius of active bodies, 1e-6) as in
//     SpatialHashBroadPhase (SpatialHashBroadPhase.cs:25-33); a body's cell key
//     packs floor(x / cell) per axis into 21-bit fields (wrapping: aliased
//     cells only add candidates, never lose one);
//   * active bodies sorted by (cell key, index) with the radix sort of the
//     tree build; every active, collidable body scans the 27 neighbouring
//     cells (9 key ranges) and runs TryAddEvent's narrow phase against each
//     candidate of larger index, so the event set is exactly the i < j pairs
//     with d^2 < (r_i + r_j)^2 -- what the spatial hash yields;
//   * events are ordered like the reference's: (i, j) ascending, then a
//     stable sort by overlap depth, deepest first (two radix sorts).
// Resolution (CollisionResolver.Resolve in MergeOnly mode,
// Collisions/CollisionResolver.cs:58-130, with MergePolicy.Merge,
// MergePolicy.cs:31-80): events are applied in list order semantics by
// rounds: an event runs once every earlier pending event that shares one of
// its bodies has run, so events applied together touch disjoint bodies and
// the result equals the sequential loop's, bit for bit.
// Compaction (ParticleSystem::compact, engine/src/physics/particle_system.cpp:
// 156-184): order-preserving scatter of all 18 arrays after a scan of the
// active flags.
// Built with -ffp-contract=off; the radius uses a correctly rounded cube root
// computed by the same fixed IEEE sequence as the CPU restatement.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <vector>

#include "cm_internal.h"

namespace cm {

struct CollEvent {   // layout of cm_collision_event
    int32_t a, b;
    double overlap, dist, nx, ny, nz, rns;
};
static_assert(sizeof(CollEvent) == sizeof(cm_collision_event), "event layout");

struct CollHeader {
    double max_radius;
    double inv_cell;
    int32_t n_events;
    int32_t n_merged;
    int32_t n_active;
    int32_t pad;
};

struct Coll {
    int cap_bodies = 0;
    int cap_events = 0;
    uint64_t *ka = nullptr, *kb = nullptr;
    int32_t *va = nullptr, *vb = nullptr;
    uint32_t* hist = nullptr;
    CollEvent* raw = nullptr;      // detected, unordered
    CollEvent* ev = nullptr;       // ordered
    uint64_t *eka = nullptr, *ekb = nullptr;
    int32_t *eva = nullptr, *evb = nullptr;
    int32_t* first = nullptr;      // per body: earliest pending event
    uint8_t* done = nullptr;       // per event
    CollHeader* hdr = nullptr;
    int32_t* scan = nullptr;       // compaction: per-block sums
    int32_t* new_index = nullptr;
    int n_events = 0;              // of the last detection
};

static Coll* coll(Engine& e) { return static_cast<Coll*>(e.coll); }

void collide_free(Engine& e) {
    Coll* c = coll(e);
    if (!c) return;
    void* ptrs[] = {c->ka, c->kb, c->va, c->vb, c->hist, c->raw, c->ev, c->eka,
                    c->ekb, c->eva, c->evb, c->first, c->done, c->hdr, c->scan,
                    c->new_index};
    for (void* p : ptrs) if (p) (void)hipFree(p);
    delete c;
    e.coll = nullptr;
}

static int ensure_coll(Engine& e) {
    if (e.coll) return CM_OK;
    Coll* c = new Coll();
    e.coll = c;
    size_t n = (size_t)e.capacity;
    size_t ne = std::max<size_t>(4096, 4 * n);
    c->cap_bodies = (int)n;
    c->cap_events = (int)std::min<size_t>(ne, (size_t)INT32_MAX / 2);
    size_t big = std::max(n, (size_t)c->cap_events);
    size_t tiles = (big + kSortTile - 1) / kSortTile;
    bool ok = true;
    auto alloc = [&](void** p, size_t bytes) {
        if (ok && hipMalloc(p, std::max<size_t>(bytes, 256)) != hipSuccess) ok = false;
    };
    alloc((void**)&c->ka, n * 8);
    alloc((void**)&c->kb, n * 8);
    alloc((void**)&c->va, n * 4);
    alloc((void**)&c->vb, n * 4);
    alloc((void**)&c->hist, 256 * (tiles + 1) * 4);
    alloc((void**)&c->raw, (size_t)c->cap_events * sizeof(CollEvent));
    alloc((void**)&c->ev, (size_t)c->cap_events * sizeof(CollEvent));
    alloc((void**)&c->eka, (size_t)c->cap_events * 8);
    alloc((void**)&c->ekb, (size_t)c->cap_events * 8);
    alloc((void**)&c->eva, (size_t)c->cap_events * 4);
    alloc((void**)&c->evb, (size_t)c->cap_events * 4);
    alloc((void**)&c->first, n * 4);
    alloc((void**)&c->done, (size_t)c->cap_events);
    alloc((void**)&c->hdr, sizeof(CollHeader));
    alloc((void**)&c->scan, ((n + 255) / 256 + 1) * 4);
    alloc((void**)&c->new_index, n * 4);
    if (!ok) {
        collide_free(e);
        return set_error(e, CM_ERR_OUT_OF_MEMORY, "collision scratch allocation failed");
    }
    return CM_OK;
}

// ------------------------------------------------------------ detection ----
constexpr uint64_t kCellMask = (1ull << 21) - 1ull;

__device__ __forceinline__ int64_t cell_coord(double x, double inv) {
    double u = floor(x * inv);
    if (!(u > -4.0e18)) u = -4.0e18;   // also maps NaN
    if (u > 4.0e18) u = 4.0e18;
    return (int64_t)u;
}

__device__ __forceinline__ uint64_t cell_key(int64_t cx, int64_t cy, int64_t cz) {
    return (((uint64_t)cx & kCellMask) << 42) | (((uint64_t)cy & kCellMask) << 21) |
           ((uint64_t)cz & kCellMask);
}

__global__ void __launch_bounds__(256)
k_max_radius(int n, const double* __restrict__ r, const uint8_t* __restrict__ act,
             CollHeader* __restrict__ h) {
    __shared__ double s[256];
    __shared__ int sc[256];
    double m = 0.0;
    int c = 0;
    for (int i = threadIdx.x; i < n; i += 256)
        if (act[i]) { m = r[i] > m ? r[i] : m; c++; }
    s[threadIdx.x] = m;
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s[threadIdx.x] = s[threadIdx.x + o] > s[threadIdx.x] ? s[threadIdx.x + o] : s[threadIdx.x];
            sc[threadIdx.x] += sc[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double cell = s[0] * 2.0;
        cell = cell > 1e-6 ? cell : 1e-6;   // Math.Max(maxRadius * 2.0, 1e-6)
        h->max_radius = s[0];
        h->inv_cell = 1.0 / cell;
        h->n_events = 0;
        h->n_merged = 0;
        h->n_active = sc[0];
    }
}

__global__ void __launch_bounds__(256)
k_cell_keys(int n, const double* __restrict__ px, const double* __restrict__ py,
            const double* __restrict__ pz, const uint8_t* __restrict__ act,
            const CollHeader* __restrict__ h, uint64_t* __restrict__ keys,
            int32_t* __restrict__ vals) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint64_t k = ~0ull;
    if (act[i]) {
        double inv = h->inv_cell;
        k = cell_key(cell_coord(px[i], inv), cell_coord(py[i], inv), cell_coord(pz[i], inv));
    }
    keys[i] = k;
    vals[i] = i;
}

__device__ __forceinline__ int lower_key(const uint64_t* __restrict__ k, int n, uint64_t x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (k[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// CollisionDetector.TryAddEvent (CollisionDetector.cs:70-135) for i < j.
__device__ __forceinline__ void try_event(int i, int j, const double* px, const double* py,
                                          const double* pz, const double* vx,
                                          const double* vy, const double* vz,
                                          const double* m, const double* r,
                                          CollHeader* h, CollEvent* out, int cap) {
    double dx = px[i] - px[j];
    double dy = py[i] - py[j];
    double dz = pz[i] - pz[j];
    double dist2 = dx * dx + dy * dy + dz * dz;
    double sumr = r[i] + r[j];
    double sumr2 = sumr * sumr;
    if (dist2 >= sumr2) return;
    double dist = sqrt(dist2);
    double overlap = sumr - dist;
    double nx = 0.0, ny = 0.0, nz = 0.0;
    if (dist > 1e-12) { nx = dx / dist; ny = dy / dist; nz = dz / dist; }
    double rvx = vx[j] - vx[i];
    double rvy = vy[j] - vy[i];
    double rvz = vz[j] - vz[i];
    double rns = rvx * nx + rvy * ny + rvz * nz;
    int a = i, b = j;
    if (m[j] > m[i]) {
        a = j; b = i;
        nx = -nx; ny = -ny; nz = -nz;
        rns = -rns;
    }
    int slot = atomicAdd(&h->n_events, 1);
    if (slot < cap) {
        CollEvent ev;
        ev.a = a; ev.b = b; ev.overlap = overlap; ev.dist = dist;
        ev.nx = nx; ev.ny = ny; ev.nz = nz; ev.rns = rns;
        out[slot] = ev;
    }
}

__global__ void __launch_bounds__(256)
k_detect(int n, const uint64_t* __restrict__ keys, const int32_t* __restrict__ vals,
         const double* __restrict__ px, const double* __restrict__ py,
         const double* __restrict__ pz, const double* __restrict__ vx,
         const double* __restrict__ vy, const double* __restrict__ vz,
         const double* __restrict__ m, const double* __restrict__ r,
         const uint8_t* __restrict__ col, CollHeader* __restrict__ h,
         CollEvent* __restrict__ out, int cap) {
    int p = blockIdx.x * 256 + threadIdx.x;
    const int na = h->n_active;
    if (p >= na) return;
    int i = vals[p];
    if (!col[i]) return;
    double inv = h->inv_cell;
    int64_t cx = cell_coord(px[i], inv), cy = cell_coord(py[i], inv), cz = cell_coord(pz[i], inv);
    for (int ddx = -1; ddx <= 1; ddx++) {
        for (int ddy = -1; ddy <= 1; ddy++) {
            uint64_t k0 = cell_key(cx + ddx, cy + ddy, cz - 1);
            uint64_t k2 = cell_key(cx + ddx, cy + ddy, cz + 1);
            int nr = 1;
            uint64_t rl[3], rh[3];
            if (k2 == k0 + 2ull) { rl[0] = k0; rh[0] = k2; }
            else {
                nr = 3;
                for (int dz = -1; dz <= 1; dz++) {
                    uint64_t kk = cell_key(cx + ddx, cy + ddy, cz + dz);
                    rl[dz + 1] = kk; rh[dz + 1] = kk;
                }
            }
            for (int q = 0; q < nr; q++) {
                for (int s = lower_key(keys, na, rl[q]); s < na && keys[s] <= rh[q]; s++) {
                    int j = vals[s];
                    if (j <= i || !col[j]) continue;
                    try_event(i, j, px, py, pz, vx, vy, vz, m, r, h, out, cap);
                }
            }
        }
    }
}

__global__ void __launch_bounds__(256)
k_event_pair_keys(const CollHeader* __restrict__ h, int cap,
                  const CollEvent* __restrict__ raw, uint64_t* __restrict__ k,
                  int32_t* __restrict__ v) {
    int e = blockIdx.x * 256 + threadIdx.x;
    int ne = min(h->n_events, cap);
    if (e >= ne) return;
    uint32_t i = (uint32_t)min(raw[e].a, raw[e].b), j = (uint32_t)max(raw[e].a, raw[e].b);
    k[e] = ((uint64_t)i << 32) | j;
    v[e] = e;
}

// overlap > 0 for every event: larger doubles have larger bit patterns, so
// the complement sorts the deepest first
__global__ void __launch_bounds__(256)
k_event_depth_keys(const CollHeader* __restrict__ h, int cap,
                   const CollEvent* __restrict__ raw, const int32_t* __restrict__ v,
                   uint64_t* __restrict__ k) {
    int e = blockIdx.x * 256 + threadIdx.x;
    int ne = min(h->n_events, cap);
    if (e >= ne) return;
    k[e] = ~(uint64_t)__double_as_longlong(raw[v[e]].overlap);
}

__global__ void __launch_bounds__(256)
k_event_gather(const CollHeader* __restrict__ h, int cap,
               const CollEvent* __restrict__ raw, const int32_t* __restrict__ v,
               CollEvent* __restrict__ ev, uint8_t* __restrict__ done) {
    int e = blockIdx.x * 256 + threadIdx.x;
    int ne = min(h->n_events, cap);
    if (e >= ne) return;
    ev[e] = raw[v[e]];
    done[e] = 0;
}

int collide_detect(Engine& e, cm_collision_event* out, int32_t cap, int32_t* ne_out) {
    int rc = ensure_coll(e);
    if (rc != CM_OK) return rc;
    Coll* c = coll(e);
    const int n = e.count;
    hipStream_t st = e.stream;
    *ne_out = 0;
    c->n_events = 0;
    if (n == 0) return CM_OK;
    hipLaunchKernelGGL(k_max_radius, dim3(1), dim3(256), 0, st, n, e.f[RAD], e.act, c->hdr);
    int gb = (n + 255) / 256;
    hipLaunchKernelGGL(k_cell_keys, dim3(gb), dim3(256), 0, st, n, e.f[PX], e.f[PY],
                       e.f[PZ], e.act, c->hdr, c->ka, c->va);
    uint64_t *k = c->ka, *kt = c->kb;
    int32_t *v = c->va, *vt = c->vb;
    radix_sort_pairs(st, n, 0, 8, &k, &v, &kt, &vt, c->hist);
    hipLaunchKernelGGL(k_detect, dim3(gb), dim3(256), 0, st, n, k, v, e.f[PX], e.f[PY],
                       e.f[PZ], e.f[VX], e.f[VY], e.f[VZ], e.f[MASS], e.f[RAD], e.col,
                       c->hdr, c->raw, c->cap_events);
    CM_HIP_TRY(e, hipGetLastError());
    CollHeader hh;
    CM_HIP_TRY(e, hipMemcpyAsync(&hh, c->hdr, sizeof(hh), hipMemcpyDeviceToHost, st));
    CM_HIP_TRY(e, hipStreamSynchronize(st));
    if (hh.n_events > c->cap_events)
        return set_error(e, CM_ERR_POOL_EXHAUSTED, "collision event capacity exceeded");
    int ne = hh.n_events;
    if (ne > 0) {
        int eb = (ne + 255) / 256;
        hipLaunchKernelGGL(k_event_pair_keys, dim3(eb), dim3(256), 0, st, c->hdr,
                           c->cap_events, c->raw, c->eka, c->eva);
        uint64_t *ek = c->eka, *ekt = c->ekb;
        int32_t *ev = c->eva, *evt = c->evb;
        radix_sort_pairs(st, ne, 0, 8, &ek, &ev, &ekt, &evt, c->hist);
        hipLaunchKernelGGL(k_event_depth_keys, dim3(eb), dim3(256), 0, st, c->hdr,
                           c->cap_events, c->raw, ev, ek);
        radix_sort_pairs(st, ne, 0, 8, &ek, &ev, &ekt, &evt, c->hist);
        hipLaunchKernelGGL(k_event_gather, dim3(eb), dim3(256), 0, st, c->hdr,
                           c->cap_events, c->raw, ev, c->ev, c->done);
        CM_HIP_TRY(e, hipGetLastError());
        if (out && cap > 0)
            CM_HIP_TRY(e, hipMemcpyAsync(out, c->ev, (size_t)std::min(cap, ne) * sizeof(CollEvent),
                                         hipMemcpyDeviceToHost, st));
        CM_HIP_TRY(e, hipStreamSynchronize(st));
    }
    c->n_events = ne;
    *ne_out = ne;
    return CM_OK;
}

// ----------------------------------------------------------- resolution ----
// Correctly rounded cube root by a fixed IEEE sequence (see DensityModel use).
__device__ double cube_residual(double c, double x) {
    double c2 = c * c;
    double e2 = fma(c, c, -c2);
    double p = c2 * c;
    double ep = fma(c2, c, -p);
    double qv = e2 * c;
    double eq = fma(e2, c, -qv);
    return ((p - x) + ep) + (qv + eq);
}

__device__ double cbrt_rn(double x) {
    if (x == 0.0 || !isfinite(x)) return x;
    bool neg = x < 0.0;
    double a = neg ? -x : x;
    int ex;
    double f = frexp(a, &ex);
    int q = ex >= 0 ? ex / 3 : -((-ex + 2) / 3);
    int r = ex - 3 * q;
    double t = ldexp(f, r);
    double y = 0.5874 + t * (0.4820 + t * (-0.0901 + t * 0.0078));
    for (int it = 0; it < 4; it++) {
        double y2 = y * y;
        y = y - (y2 * y - t) / (3.0 * y2);
    }
    y = ldexp(y, q);
    double lo = nextafter(y, 0.0), hi = nextafter(y, INFINITY);
    double rb = fabs(cube_residual(y, a));
    double rl = fabs(cube_residual(lo, a));
    double rh = fabs(cube_residual(hi, a));
    double best = y;
    if (rl < rb) { best = lo; rb = rl; }
    if (rh < rb) best = hi;
    return neg ? -best : best;
}

// DensityModel.ComputeBodyRadius (Types/DensityModel.cs:54-100).
__device__ double body_radius(double mass, double density, int32_t type) {
    if (type == 7) {   // BodyType.BlackHole: Schwarzschild radius in AU
        if (mass <= 0.0) return 0.0;
        double kg = mass * 1.989e30;
        double rm = 2.0 * 6.67430e-11 * kg / (2.998e8 * 2.998e8);
        return rm / 1.496e11;
    }
    if (mass <= 0.0 || density <= 0.0) return 0.0;
    double volume = mass / density;
    return cbrt_rn(3.0 * volume / (4.0 * 3.14159265358979323846));
}

struct SoaPtrs {
    double *px, *py, *pz, *vx, *vy, *vz, *ax, *ay, *az, *oax, *oay, *oaz, *m, *rad, *dens;
    uint8_t *act, *col;
    int32_t* type;
};

__global__ void __launch_bounds__(1024)
k_resolve(int ne, const CollEvent* __restrict__ ev, uint8_t* __restrict__ done,
          int32_t* __restrict__ first, SoaPtrs s, CollHeader* __restrict__ h) {
    __shared__ int s_pending;
    __shared__ int s_merged;
    if (threadIdx.x == 0) s_merged = 0;
    for (;;) {
        if (threadIdx.x == 0) s_pending = 0;
        __syncthreads();
        for (int k = threadIdx.x; k < ne; k += blockDim.x) {
            if (done[k]) continue;
            first[ev[k].a] = INT_MAX;
            first[ev[k].b] = INT_MAX;
        }
        __syncthreads();
        for (int k = threadIdx.x; k < ne; k += blockDim.x) {
            if (done[k]) continue;
            atomicMin(&first[ev[k].a], k);
            atomicMin(&first[ev[k].b], k);
        }
        __syncthreads();
        for (int k = threadIdx.x; k < ne; k += blockDim.x) {
            if (done[k]) continue;
            int a = ev[k].a, b = ev[k].b;
            if (first[a] != k || first[b] != k) { atomicAdd(&s_pending, 1); continue; }
            done[k] = 1;
            if (!s.act[a] || !s.act[b]) continue;          // CollisionResolver.cs:76-77
            double ma = s.m[a], mb = s.m[b];
            if (ma <= 0.0 || mb <= 0.0) continue;           // :79-81
            atomicAdd(&s_merged, 1);
            double mt = ma + mb;
            if (mt <= 0.0) continue;                        // MergePolicy.cs:37
            double inv = 1.0 / mt;
            s.vx[a] = (ma * s.vx[a] + mb * s.vx[b]) * inv;
            s.vy[a] = (ma * s.vy[a] + mb * s.vy[b]) * inv;
            s.vz[a] = (ma * s.vz[a] + mb * s.vz[b]) * inv;
            s.px[a] = (ma * s.px[a] + mb * s.px[b]) * inv;
            s.py[a] = (ma * s.py[a] + mb * s.py[b]) * inv;
            s.pz[a] = (ma * s.pz[a] + mb * s.pz[b]) * inv;
            s.m[a] = mt;
            s.dens[a] = (ma * s.dens[a] + mb * s.dens[b]) * inv;
            if (s.type[b] == 7) s.type[a] = 7;
            s.rad[a] = body_radius(mt, s.dens[a], s.type[a]);
            s.act[b] = 0;
            s.col[b] = 0;
            s.ax[b] = 0.0; s.ay[b] = 0.0; s.az[b] = 0.0;
            s.oax[b] = 0.0; s.oay[b] = 0.0; s.oaz[b] = 0.0;
        }
        __syncthreads();
        int pending = s_pending;
        __syncthreads();
        if (pending == 0) break;
    }
    if (threadIdx.x == 0) h->n_merged = s_merged;
}

static SoaPtrs soa_ptrs(Engine& e) {
    SoaPtrs s;
    s.px = e.f[PX]; s.py = e.f[PY]; s.pz = e.f[PZ];
    s.vx = e.f[VX]; s.vy = e.f[VY]; s.vz = e.f[VZ];
    s.ax = e.f[AX]; s.ay = e.f[AY]; s.az = e.f[AZ];
    s.oax = e.f[OAX]; s.oay = e.f[OAY]; s.oaz = e.f[OAZ];
    s.m = e.f[MASS]; s.rad = e.f[RAD]; s.dens = e.f[DENS];
    s.act = e.act; s.col = e.col; s.type = e.type;
    return s;
}

int collide_resolve(Engine& e, int32_t* merged) {
    Coll* c = coll(e);
    if (merged) *merged = 0;
    if (!c || c->n_events == 0) return CM_OK;
    hipStream_t st = e.stream;
    hipLaunchKernelGGL(k_resolve, dim3(1), dim3(1024), 0, st, c->n_events, c->ev, c->done,
                       c->first, soa_ptrs(e), c->hdr);
    CM_HIP_TRY(e, hipGetLastError());
    CollHeader hh;
    CM_HIP_TRY(e, hipMemcpyAsync(&hh, c->hdr, sizeof(hh), hipMemcpyDeviceToHost, st));
    CM_HIP_TRY(e, hipStreamSynchronize(st));
    c->n_events = 0;   // events are consumed
    if (merged) *merged = hh.n_merged;
    return CM_OK;
}

// ----------------------------------------------------------- compaction ----
__global__ void __launch_bounds__(256)
k_flag_sums(int n, const uint8_t* __restrict__ act, int32_t* __restrict__ bsum) {
    __shared__ int s[256];
    int i = blockIdx.x * 256 + threadIdx.x;
    s[threadIdx.x] = (i < n && act[i]) ? 1 : 0;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(1024)
k_scan_blocks(int nb, int32_t* __restrict__ bsum) {
    __shared__ int s[1024];
    int per = (nb + 1023) / 1024;
    int b = threadIdx.x * per, e = min(nb, b + per);
    int acc = 0;
    for (int i = b; i < e; i++) acc += bsum[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int v = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    int run = threadIdx.x ? s[threadIdx.x - 1] : 0;
    for (int i = b; i < e; i++) {
        int v = bsum[i];
        bsum[i] = run;
        run += v;
    }
    if (threadIdx.x == 1023) bsum[nb] = s[1023];
}

__global__ void __launch_bounds__(256)
k_new_index(int n, const uint8_t* __restrict__ act, const int32_t* __restrict__ bsum,
            int32_t* __restrict__ idx) {
    __shared__ int s[256];
    int i = blockIdx.x * 256 + threadIdx.x;
    int f = (i < n && act[i]) ? 1 : 0;
    s[threadIdx.x] = f;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        int v = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    if (i < n) idx[i] = f ? bsum[blockIdx.x] + s[threadIdx.x] - 1 : -1;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_scatter(int n, const int32_t* __restrict__ idx, const T* __restrict__ src,
          T* __restrict__ dst) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n && idx[i] >= 0) dst[idx[i]] = src[i];
}

__global__ void __launch_bounds__(256)
k_old_index(int n, const int32_t* __restrict__ idx, int32_t* __restrict__ old) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n && idx[i] >= 0) old[idx[i]] = i;
}

int collide_compact(Engine& e, int32_t* new_count, int32_t* old_index) {
    int rc = ensure_coll(e);
    if (rc != CM_OK) return rc;
    Coll* c = coll(e);
    const int n = e.count;
    hipStream_t st = e.stream;
    if (n == 0) {
        if (new_count) *new_count = 0;
        return CM_OK;
    }
    int nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_flag_sums, dim3(nb), dim3(256), 0, st, n, e.act, c->scan);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st, nb, c->scan);
    hipLaunchKernelGGL(k_new_index, dim3(nb), dim3(256), 0, st, n, e.act, c->scan,
                       c->new_index);
    // every array goes through the sort buffer ka (n doubles / n x 8 bytes)
    double* tmp = reinterpret_cast<double*>(c->ka);
    for (int k = 0; k < NF64; k++) {
        hipLaunchKernelGGL(k_scatter<double>, dim3(nb), dim3(256), 0, st, n, c->new_index,
                           e.f[k], tmp);
        CM_HIP_TRY(e, hipMemcpyAsync(e.f[k], tmp, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
    }
    uint8_t* t8 = reinterpret_cast<uint8_t*>(c->ka);
    hipLaunchKernelGGL(k_scatter<uint8_t>, dim3(nb), dim3(256), 0, st, n, c->new_index, e.act, t8);
    CM_HIP_TRY(e, hipMemcpyAsync(e.act, t8, n, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_scatter<uint8_t>, dim3(nb), dim3(256), 0, st, n, c->new_index, e.col, t8);
    CM_HIP_TRY(e, hipMemcpyAsync(e.col, t8, n, hipMemcpyDeviceToDevice, st));
    int32_t* t32 = reinterpret_cast<int32_t*>(c->ka);
    hipLaunchKernelGGL(k_scatter<int32_t>, dim3(nb), dim3(256), 0, st, n, c->new_index, e.type, t32);
    CM_HIP_TRY(e, hipMemcpyAsync(e.type, t32, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    int32_t* old = c->va;
    hipLaunchKernelGGL(k_old_index, dim3(nb), dim3(256), 0, st, n, c->new_index, old);
    CM_HIP_TRY(e, hipGetLastError());
    int32_t total = 0;
    CM_HIP_TRY(e, hipMemcpyAsync(&total, c->scan + nb, 4, hipMemcpyDeviceToHost, st));
    CM_HIP_TRY(e, hipStreamSynchronize(st));
    if (old_index && total > 0)
        CM_HIP_TRY(e, hipMemcpy(old_index, old, (size_t)total * 4, hipMemcpyDeviceToHost));
    e.count = total;
    c->n_events = 0;
    if (new_count) *new_count = total;
    return CM_OK;
}

}  // namespace cm
