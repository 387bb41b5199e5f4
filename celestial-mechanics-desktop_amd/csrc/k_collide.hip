// k_collide.hip -- collision detection, resolution in every CollisionMode and
// compaction.
//
// Detection (CollisionDetector.Detect, Collisions/CollisionDetector.cs:35-135):
//   * cell size max(2 * max radius of active bodies, 1e-6) as in
//     SpatialHashBroadPhase (SpatialHashBroadPhase.cs:25-33); a body's cell key
//     packs floor(x / cell) per axis into 21-bit fields (wrapping: aliased
//     cells only add candidates, never lose one);
//   * each cell goes to one of 2^B buckets (B = log2(capacity) + 2, so most
//     buckets hold one cell; the cells of a 4x4x4 block get consecutive
//     buckets, the blocks hashed); active bodies are sorted by (bucket,
//     index) with the 32-bit radix sort of the tree build and each bucket's
//     first position is recorded; every active, collidable body looks up the
//     buckets of its 27 neighbouring cells (each distinct bucket once) and
//     runs TryAddEvent's narrow phase against each candidate of larger index.
//     Buckets shared by several cells only add candidates, which the narrow
//     phase rejects, so the event set is exactly the i < j pairs with
//     d^2 < (r_i + r_j)^2 -- what the spatial hash yields;
//   * events are ordered like the reference's: (i, j) ascending, then a
//     stable sort by overlap depth, deepest first (two radix sorts, or one
//     bitonic sort in LDS for up to 2048 events).
// Resolution (CollisionResolver.Resolve, Collisions/CollisionResolver.cs:
// 58-247, every mode and outcome, with MergePolicy.Merge, MergePolicy.cs:
// 31-80): events are applied in list order semantics by rounds: an event runs
// once every earlier pending event that shares one of its bodies has run, so
// events applied together touch disjoint bodies and the result equals the
// sequential loop's, bit for bit.  The engine's merge caps depend on the
// merges of all earlier events, so a one-thread pass decides first which
// events may merge.
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
#include "cm_math.h"

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
    uint8_t* caps_act = nullptr;   // merge caps: active flags as the decision pass sees them
    int32_t* caps_count = nullptr; // merge caps: merges per body this call
    uint8_t* caps_allow = nullptr; // merge caps: per event
    int32_t* bstart = nullptr;     // per cell bucket: first sorted position (stale
                                   // entries fail the key check of the lookup)
    double4* sp = nullptr;         // x, y, z, radius of the active bodies in bucket order
    uint8_t* scol = nullptr;       // their collidable flags
    int bucket_bits = 0;
    int n_events = 0;              // of the last detection
};

static Coll* coll(Engine& e) { return static_cast<Coll*>(e.coll); }

void collide_free(Engine& e) {
    Coll* c = coll(e);
    if (!c) return;
    void* ptrs[] = {c->ka, c->kb, c->va, c->vb, c->hist, c->raw, c->ev, c->eka,
                    c->ekb, c->eva, c->evb, c->first, c->done, c->hdr, c->scan,
                    c->new_index, c->caps_act, c->caps_count, c->caps_allow, c->bstart,
                    c->sp, c->scol};
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
    int bits = 10;
    while (bits < 26 && ((size_t)1 << bits) < 4 * n) bits++;
    c->bucket_bits = bits;
    alloc((void**)&c->bstart, ((size_t)1 << bits) * 4);
    alloc((void**)&c->sp, n * sizeof(double4));
    alloc((void**)&c->scol, n);
    if (ok && hipMemset(c->bstart, 0xff, ((size_t)1 << bits) * 4) != hipSuccess) ok = false;
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

// Largest radius and count of the active bodies: per-block partials, then
// one block folds them and sets the hash cell (SpatialHashBroadPhase.cs:25-33).
constexpr int kRadBlocks = 256;

__global__ void __launch_bounds__(256)
k_max_radius_part(int n, const double* __restrict__ r, const uint8_t* __restrict__ act,
                  double* __restrict__ part_r, int32_t* __restrict__ part_c) {
    __shared__ double s[256];
    __shared__ int sc[256];
    double m = 0.0;
    int c = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
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
    if (threadIdx.x == 0) { part_r[blockIdx.x] = s[0]; part_c[blockIdx.x] = sc[0]; }
}

__global__ void __launch_bounds__(256)
k_max_radius(int nparts, const double* __restrict__ part_r, const int32_t* __restrict__ part_c,
             CollHeader* __restrict__ h) {
    __shared__ double s[256];
    __shared__ int sc[256];
    s[threadIdx.x] = threadIdx.x < nparts ? part_r[threadIdx.x] : 0.0;
    sc[threadIdx.x] = threadIdx.x < nparts ? part_c[threadIdx.x] : 0;
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

// Bucket of cell (cx, cy, cz): cells are grouped in blocks of 4 x 4 x 4; the
// block's packed key is hashed (top bits - 6 bits of the SplitMix64
// finaliser) and the cell's place in its block fills the low 6 bits, so the
// cells of one block have consecutive buckets and the bodies of a block are
// contiguous after the sort (neighbouring threads then share most of their
// neighbour cells).
__device__ __forceinline__ uint32_t cell_bucket(int64_t cx, int64_t cy, int64_t cz, int bits) {
    uint64_t z = cell_key(cx >> 2, cy >> 2, cz >> 2) + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    const uint32_t local = (uint32_t)(((cx & 3) << 4) | ((cy & 3) << 2) | (cz & 3));
    return ((uint32_t)(z >> (64 - (bits - 6))) << 6) | local;
}

// Bucket of each active body's cell (inactive bodies: 2^bits, after every
// bucket) and its index, the pairs the radix sort orders.
__global__ void __launch_bounds__(256)
k_cell_keys(int n, int bits, const double* __restrict__ px, const double* __restrict__ py,
            const double* __restrict__ pz, const uint8_t* __restrict__ act,
            const CollHeader* __restrict__ h, uint32_t* __restrict__ keys,
            int32_t* __restrict__ vals) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint32_t k = 1u << bits;
    if (act[i]) {
        double inv = h->inv_cell;
        k = cell_bucket(cell_coord(px[i], inv), cell_coord(py[i], inv), cell_coord(pz[i], inv),
                        bits);
    }
    keys[i] = k;
    vals[i] = i;
}

// First sorted position of every bucket present among the active bodies, and
// the bodies' positions, radii and collidable flags in sorted order (so that
// the candidates of one bucket are read from consecutive addresses).
__global__ void __launch_bounds__(256)
k_bucket_starts(const CollHeader* __restrict__ h, const uint32_t* __restrict__ keys,
                const int32_t* __restrict__ vals, const double* __restrict__ px,
                const double* __restrict__ py, const double* __restrict__ pz,
                const double* __restrict__ r, const uint8_t* __restrict__ col,
                int32_t* __restrict__ bstart, double4* __restrict__ sp,
                uint8_t* __restrict__ scol) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    const int na = h->n_active;
    if (p >= na) return;
    const uint32_t b = keys[p];
    if (p == 0 || keys[p - 1] != b) bstart[b] = p;
    const int i = vals[p];
    sp[p] = make_double4(px[i], py[i], pz[i], r[i]);
    scol[p] = col[i];
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

// One active, collidable body per thread (sorted position p).  The distance
// test runs on the sorted copies; a pair that overlaps goes through
// try_event on the body arrays (the same values, so the same event).
__global__ void __launch_bounds__(256)
k_detect(int n, int bits, const uint32_t* __restrict__ keys, const int32_t* __restrict__ vals,
         const int32_t* __restrict__ bstart, const double4* __restrict__ sp,
         const uint8_t* __restrict__ scol,
         const double* __restrict__ px, const double* __restrict__ py,
         const double* __restrict__ pz, const double* __restrict__ vx,
         const double* __restrict__ vy, const double* __restrict__ vz,
         const double* __restrict__ m, const double* __restrict__ r,
         CollHeader* __restrict__ h, CollEvent* __restrict__ out, int cap) {
    int p = blockIdx.x * 256 + threadIdx.x;
    const int na = h->n_active;
    if (p >= na) return;
    if (!scol[p]) return;
    const int i = vals[p];
    const double4 me = sp[p];
    double inv = h->inv_cell;
    int64_t cx = cell_coord(me.x, inv), cy = cell_coord(me.y, inv), cz = cell_coord(me.z, inv);
    // the 27 neighbouring cells' buckets, each distinct bucket scanned once
    uint32_t seen[27];
    int ns = 0;
    for (int ddx = -1; ddx <= 1; ddx++) {
        for (int ddy = -1; ddy <= 1; ddy++) {
            for (int ddz = -1; ddz <= 1; ddz++) {
                const uint32_t bk = cell_bucket(cx + ddx, cy + ddy, cz + ddz, bits);
                bool dup = false;
                for (int q = 0; q < ns; q++) dup = dup || seen[q] == bk;
                if (dup) continue;
                seen[ns++] = bk;
                int s = bstart[bk];
                if (s < 0 || s >= na || keys[s] != bk) continue;   // bucket empty
                for (; s < na && keys[s] == bk; s++) {
                    const int j = vals[s];
                    if (j <= i || !scol[s]) continue;
                    const double4 o = sp[s];
                    const double dx = me.x - o.x, dy = me.y - o.y, dz = me.z - o.z;
                    const double dist2 = dx * dx + dy * dy + dz * dz;
                    const double sumr = me.w + o.w;
                    if (dist2 >= sumr * sumr) continue;
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

// Up to kSmallEvents events are ordered in one workgroup: a bitonic sort in
// LDS on the unique composite key (depth descending, then (i, j) ascending),
// the order the two stable radix sorts below produce.
constexpr int kSmallEvents = 2048;

__global__ void __launch_bounds__(1024)
k_event_sort_small(const CollHeader* __restrict__ h, int cap, const CollEvent* __restrict__ raw,
                   CollEvent* __restrict__ ev, uint8_t* __restrict__ done) {
    __shared__ uint64_t s_hi[kSmallEvents], s_lo[kSmallEvents];
    __shared__ int32_t s_ix[kSmallEvents];
    const int ne = min(h->n_events, cap);
    int P = 1;
    while (P < ne) P <<= 1;
    for (int e = threadIdx.x; e < P; e += blockDim.x) {
        if (e < ne) {
            const CollEvent x = raw[e];
            const uint32_t i = (uint32_t)min(x.a, x.b), j = (uint32_t)max(x.a, x.b);
            s_hi[e] = ~(uint64_t)__double_as_longlong(x.overlap);   // overlap > 0
            s_lo[e] = ((uint64_t)i << 32) | j;
        } else {
            s_hi[e] = ~0ull;
            s_lo[e] = ~0ull;
        }
        s_ix[e] = e;
    }
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int jj = k >> 1; jj > 0; jj >>= 1) {
            for (int e = threadIdx.x; e < P; e += blockDim.x) {
                const int o = e ^ jj;
                if (o > e) {
                    const bool up = (e & k) == 0;
                    const bool gt = s_hi[e] > s_hi[o] || (s_hi[e] == s_hi[o] && s_lo[e] > s_lo[o]);
                    if (gt == up) {
                        uint64_t t = s_hi[e]; s_hi[e] = s_hi[o]; s_hi[o] = t;
                        t = s_lo[e]; s_lo[e] = s_lo[o]; s_lo[o] = t;
                        const int32_t q = s_ix[e]; s_ix[e] = s_ix[o]; s_ix[o] = q;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
        ev[e] = raw[s_ix[e]];
        done[e] = 0;
    }
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
    {
        const int nparts = std::min(kRadBlocks, (n + 255) / 256);
        double* part_r = e.red;                                   // diagnostics scratch
        int32_t* part_c = reinterpret_cast<int32_t*>(e.red + kRadBlocks);
        hipLaunchKernelGGL(k_max_radius_part, dim3(nparts), dim3(256), 0, st, n, e.f[RAD],
                           e.act, part_r, part_c);
        hipLaunchKernelGGL(k_max_radius, dim3(1), dim3(256), 0, st, nparts, part_r, part_c,
                           c->hdr);
    }
    int gb = (n + 255) / 256;
    const int bits = c->bucket_bits;
    uint32_t* k = reinterpret_cast<uint32_t*>(c->ka);
    uint32_t* kt = reinterpret_cast<uint32_t*>(c->kb);
    int32_t *v = c->va, *vt = c->vb;
    hipLaunchKernelGGL(k_cell_keys, dim3(gb), dim3(256), 0, st, n, bits, e.f[PX], e.f[PY],
                       e.f[PZ], e.act, c->hdr, k, v);
    radix_sort_pairs32(st, n, 0, (bits + 1 + 7) / 8, &k, &v, &kt, &vt, c->hist);
    hipLaunchKernelGGL(k_bucket_starts, dim3(gb), dim3(256), 0, st, c->hdr, k, v, e.f[PX],
                       e.f[PY], e.f[PZ], e.f[RAD], e.col, c->bstart, c->sp, c->scol);
    hipLaunchKernelGGL(k_detect, dim3(gb), dim3(256), 0, st, n, bits, k, v, c->bstart, c->sp,
                       c->scol, e.f[PX], e.f[PY], e.f[PZ], e.f[VX], e.f[VY], e.f[VZ],
                       e.f[MASS], e.f[RAD], c->hdr, c->raw, c->cap_events);
    CM_HIP_TRY(e, hipGetLastError());
    CollHeader hh;
    CM_HIP_TRY(e, hipMemcpyAsync(&hh, c->hdr, sizeof(hh), hipMemcpyDeviceToHost, st));
    CM_HIP_TRY(e, hipStreamSynchronize(st));
    if (hh.n_events > c->cap_events)
        return set_error(e, CM_ERR_POOL_EXHAUSTED, "collision event capacity exceeded");
    int ne = hh.n_events;
    if (ne > 0 && ne <= kSmallEvents) {
        hipLaunchKernelGGL(k_event_sort_small, dim3(1), dim3(1024), 0, st, c->hdr,
                           c->cap_events, c->raw, c->ev, c->done);
        CM_HIP_TRY(e, hipGetLastError());
        if (out && cap > 0)
            CM_HIP_TRY(e, hipMemcpyAsync(out, c->ev, (size_t)std::min(cap, ne) * sizeof(CollEvent),
                                         hipMemcpyDeviceToHost, st));
        CM_HIP_TRY(e, hipStreamSynchronize(st));
    } else if (ne > 0) {
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
struct SoaPtrs {
    double *px, *py, *pz, *vx, *vy, *vz, *ax, *ay, *az, *oax, *oay, *oaz, *m, *rad, *dens;
    uint8_t *act, *col;
    int32_t* type;
};

struct ResolveCfg {
    int mode;              // CM_COLLISION_*
    double restitution;    // default restitution of types without their own
    double mass_loss_cap;  // fragmentation mass-loss cap
    int promote;           // PromoteCompactRemnant after merges
};

// .NET Math.Max / Math.Clamp for doubles.
__device__ __forceinline__ double cs_max(double a, double b) {
    return (a > b || isnan(a)) ? a : b;
}
__device__ __forceinline__ double cs_clamp(double v, double lo, double hi) {
    if (v < lo) return lo;
    if (v > hi) return hi;
    return v;
}

constexpr double kSolarMassKg = 1.989e30;
constexpr double kAuM = 1.496e11;
constexpr double kTimeUnitS = 5.0226e6;
constexpr double kGSi = 6.67430e-11;

// CollisionEnergyModel.Evaluate (Astrophysics/CollisionEnergyModel.cs:20-70):
// the classification and the ejecta mass, the fields that change the state.
__device__ void collision_energy(double m1, double m2, double eff_radius_au,
                                 double rel_speed, double max_loss, bool& is_merge,
                                 bool& is_frag, double& ejecta_kg) {
    const double m1kg = cs_max(m1, 1e-15) * kSolarMassKg;
    const double m2kg = cs_max(m2, 1e-15) * kSolarMassKg;
    const double mt = m1kg + m2kg;
    const double rm = cs_max(eff_radius_au, 1e-12) * kAuM;
    const double v = (cs_max(rel_speed, 0.0) * kAuM) / kTimeUnitS;
    const double mu = (m1kg * m2kg) / cs_max(mt, 1e-15);
    const double ec = 0.5 * mu * v * v;
    const double eb = (3.0 / 5.0) * kGSi * mt * mt / cs_max(rm, 1.0);
    const double ratio = eb > 0.0 ? ec / eb : 0.0;
    const bool merge = ec < 0.5 * eb;
    const bool frag = !merge && ec < eb;
    double f = 0.0;
    if (frag) {
        const double nrm = cs_clamp((ratio - 0.5) / 0.5, 0.0, 1.0);
        f = cs_clamp(0.08 + nrm * 0.32, 0.05, max_loss);
    } else if (!merge) {
        const double nrm = cs_clamp(ratio - 1.0, 0.0, 12.0) / 12.0;
        const double ccap = cs_max(max_loss, 0.55);
        f = cs_clamp(0.5 + nrm * 0.4, 0.45, ccap);
    }
    is_merge = merge;
    is_frag = frag;
    ejecta_kg = mt * f;
}

// CollisionResolver.GetRestitutionForType (Collisions/CollisionResolver.cs:456-470);
// BodyType: Star 0, Planet 1, GasGiant 2, RockyPlanet 3, Moon 4, Asteroid 5,
// NeutronStar 6, BlackHole 7, Comet 8, Custom 9 (Types/BodyType.cs).
__device__ __forceinline__ double restitution_for(int32_t t, double dflt) {
    switch (t) {
        case 7: return 0.0;
        case 6: return 0.02;
        case 0: return 0.05;
        case 2: return 0.08;
        case 1: case 3: return 0.2;
        case 5: return 0.32;
        case 8: return 0.38;
        case 4: return 0.18;
        default: return dflt;
    }
}

__device__ __forceinline__ bool is_compact(int32_t t) { return t == 6 || t == 7; }

// DensityModel.GetDefaultDensity (Types/DensityModel.cs:76-87).
__device__ __forceinline__ double default_density(int32_t t) {
    switch (t) {
        case 0: return 1000.0;
        case 1: case 3: return 5000.0;
        case 2: return 1300.0;
        case 4: return 3300.0;
        case 5: return 3000.0;
        case 6: return 1e12;
        case 8: return 500.0;
        case 7: return 1.0;
        default: return 1000.0;
    }
}

// ResolveImpulse (CollisionResolver.cs:378-405).
__device__ void impulse(const SoaPtrs& s, int a, int b, double nx, double ny, double nz,
                        double rns, double e) {
    if (rns >= 0.0) return;
    const double ma = s.m[a], mb = s.m[b];
    if (ma <= 0.0 || mb <= 0.0) return;
    const double ia = 1.0 / ma, ib = 1.0 / mb;
    const double j = -(1.0 + e) * rns / (ia + ib);
    s.vx[a] -= j * ia * nx; s.vy[a] -= j * ia * ny; s.vz[a] -= j * ia * nz;
    s.vx[b] += j * ib * nx; s.vy[b] += j * ib * ny; s.vz[b] += j * ib * nz;
}

// PositionalCorrection (CollisionResolver.cs:407-437).
__device__ void positional(const SoaPtrs& s, int a, int b, double depth, double nx,
                           double ny, double nz) {
    if (depth <= 0.0) return;
    const double ma = s.m[a], mb = s.m[b];
    if (ma <= 0.0 || mb <= 0.0) return;
    const double ia = 1.0 / ma, ib = 1.0 / mb;
    const double c = 0.6 * cs_max(depth - 1e-4, 0.0) / (ia + ib);
    s.px[a] -= c * ia * nx; s.py[a] -= c * ia * ny; s.pz[a] -= c * ia * nz;
    s.px[b] += c * ib * nx; s.py[b] += c * ib * ny; s.pz[b] += c * ib * nz;
}

// ConservePairMomentum (CollisionResolver.cs:527-546).
__device__ void conserve_pair(const SoaPtrs& s, int a, int b, double dx, double dy,
                              double dz) {
    const double ma = cs_max(s.m[a], 1e-15), mb = cs_max(s.m[b], 1e-15);
    const double cx = ma * s.vx[a] + mb * s.vx[b];
    const double cy = ma * s.vy[a] + mb * s.vy[b];
    const double cz = ma * s.vz[a] + mb * s.vz[b];
    const double ex = dx - cx, ey = dy - cy, ez = dz - cz;
    const double it = 1.0 / (ma + mb);
    s.vx[a] += ex * it; s.vy[a] += ey * it; s.vz[a] += ez * it;
    s.vx[b] += ex * it; s.vy[b] += ey * it; s.vz[b] += ez * it;
}

__device__ void deactivate(const SoaPtrs& s, int b) {
    s.act[b] = 0;
    s.col[b] = 0;
    s.ax[b] = 0.0; s.ay[b] = 0.0; s.az[b] = 0.0;
    s.oax[b] = 0.0; s.oay[b] = 0.0; s.oaz[b] = 0.0;
}

// MergePolicy.Merge (Collisions/MergePolicy.cs:31-80).
__device__ void merge_policy(const SoaPtrs& s, int a, int b) {
    const double ma = s.m[a], mb = s.m[b];
    const double mt = ma + mb;
    if (mt <= 0.0) return;                          // MergePolicy.cs:37
    const double inv = 1.0 / mt;
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
    deactivate(s, b);
}

// PromoteCompactRemnant (CollisionResolver.cs:577-607).
__device__ void promote(const SoaPtrs& s, int a) {
    const int32_t cur = s.type[a];
    if (!(cur == 0 || cur == 6 || cur == 7)) return;
    const double mass = s.m[a];
    int32_t tgt = cur;
    if (mass >= 3.0) tgt = 7;
    else if (mass >= 1.6) tgt = 6;
    if (tgt == cur) return;
    const double d = default_density(tgt);
    s.type[a] = tgt;
    s.dens[a] = d;
    s.rad[a] = body_radius(mass, d, tgt);
}

// CollisionResolver.Resolve's body for one event (CollisionResolver.cs:72-244):
// returns whether the event was resolved (counted in LastCollisionCount).
__device__ bool resolve_event(const SoaPtrs& s, const CollEvent& e, const ResolveCfg& c,
                              bool allowed) {
    int a = e.a, b = e.b;
    if (!s.act[a] || !s.act[b]) return false;      // :76-77
    double ma = s.m[a], mb = s.m[b];
    if (ma <= 0.0 || mb <= 0.0) return false;       // :79-81
    double nx = s.px[b] - s.px[a], ny = s.py[b] - s.py[a], nz = s.pz[b] - s.pz[a];
    const double dist = sqrt(nx * nx + ny * ny + nz * nz);
    if (dist > 1e-12) {
        const double inv = 1.0 / dist;
        nx *= inv; ny *= inv; nz *= inv;
    } else {
        nx = 1.0; ny = 0.0; nz = 0.0;
    }
    const double rvx = s.vx[b] - s.vx[a], rvy = s.vy[b] - s.vy[a], rvz = s.vz[b] - s.vz[a];
    const double rns = rvx * nx + rvy * ny + rvz * nz;
    const double rsp = sqrt(rvx * rvx + rvy * rvy + rvz * rvz);
    bool is_merge, is_frag;
    double ej_kg;
    collision_energy(ma, mb, 0.5 * cs_max(s.rad[a] + s.rad[b], 1e-12), rsp, c.mass_loss_cap,
                     is_merge, is_frag, ej_kg);
    // SelectOutcome (:249-267): 0 Bounce, 1 Merge, 2 Fragmentation, 3 Accretion,
    // 4 CatastrophicDisruption; the mode overrides it (:123-126)
    int outcome;
    if (is_compact(s.type[a]) || is_compact(s.type[b])) outcome = 3;
    else if (is_merge) outcome = 1;
    else if (is_frag) outcome = 2;
    else outcome = 4;
    if (c.mode == CM_COLLISION_BOUNCE_ONLY) outcome = 0;
    else if (c.mode == CM_COLLISION_MERGE_ONLY) outcome = 1;
    if (outcome == 3 && !is_compact(s.type[a]) && is_compact(s.type[b])) {
        const int t = a; a = b; b = t;              // EnsureAccretorSurvivor (:565-575)
    }
    ma = s.m[a];
    mb = s.m[b];
    if (outcome == 1 || outcome == 3) {
        if (!allowed) return false;                 // the engine's merge caps
        merge_policy(s, a, b);
        if (c.promote) promote(s, a);
    } else if (outcome == 2) {
        const double r = cs_clamp(0.5 * (restitution_for(s.type[a], c.restitution) +
                                         restitution_for(s.type[b], c.restitution)),
                                  0.0, 1.0);
        impulse(s, a, b, nx, ny, nz, rns, r * 0.35);
        positional(s, a, b, e.overlap, nx, ny, nz);
        const double ej = ej_kg / kSolarMassKg;
        const double ef = cs_clamp(ej / cs_max(ma + mb, 1e-15), 0.02, c.mass_loss_cap);
        const double keep = 1.0 - ef;
        const double pbx = ma * s.vx[a] + mb * s.vx[b];
        const double pby = ma * s.vy[a] + mb * s.vy[b];
        const double pbz = ma * s.vz[a] + mb * s.vz[b];
        s.m[a] = cs_max(1e-12, ma * keep);
        s.m[b] = cs_max(1e-12, mb * keep);
        s.rad[a] = body_radius(s.m[a], s.dens[a], s.type[a]);
        s.rad[b] = body_radius(s.m[b], s.dens[b], s.type[b]);
        conserve_pair(s, a, b, pbx, pby, pbz);
    } else if (outcome == 4) {
        // ResolveCatastrophicDisruption (:285-333)
        const double mt = ma + mb;
        if (mt > 0.0) {
            const double sm = cs_max(1e-12, mt - ej_kg / kSolarMassKg);
            const double qx = ma * s.vx[a] + mb * s.vx[b];
            const double qy = ma * s.vy[a] + mb * s.vy[b];
            const double qz = ma * s.vz[a] + mb * s.vz[b];
            s.px[a] = (ma * s.px[a] + mb * s.px[b]) / mt;
            s.py[a] = (ma * s.py[a] + mb * s.py[b]) / mt;
            s.pz[a] = (ma * s.pz[a] + mb * s.pz[b]) / mt;
            s.vx[a] = qx / sm; s.vy[a] = qy / sm; s.vz[a] = qz / sm;
            s.m[a] = sm;
            const int32_t ta = s.type[a], tb = s.type[b];
            const int32_t st = (is_compact(ta) || is_compact(tb)) ? (is_compact(ta) ? ta : tb)
                                                                  : (ma >= mb ? ta : tb);
            s.type[a] = st;
            s.rad[a] = body_radius(s.m[a], s.dens[a], st);
            deactivate(s, b);
        }
    } else {
        const double r = cs_clamp(0.5 * (restitution_for(s.type[a], c.restitution) +
                                         restitution_for(s.type[b], c.restitution)),
                                  0.0, 1.0);
        impulse(s, a, b, nx, ny, nz, rns, r);
        positional(s, a, b, e.overlap, nx, ny, nz);
    }
    return true;
}

// The event list in the reference's order semantics, by rounds: an event runs
// once every earlier pending event that shares one of its bodies has run.
// Each event reads and writes only its own two bodies, so events applied
// together commute and the result is the sequential loop's, bit for bit.
__global__ void __launch_bounds__(1024)
k_resolve(int ne, const CollEvent* __restrict__ ev, uint8_t* __restrict__ done,
          int32_t* __restrict__ first, SoaPtrs s, ResolveCfg cfg,
          const uint8_t* __restrict__ allow, CollHeader* __restrict__ h) {
    __shared__ int s_pending;
    __shared__ int s_resolved;
    if (threadIdx.x == 0) s_resolved = 0;
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
            const CollEvent e = ev[k];
            if (first[e.a] != k || first[e.b] != k) { atomicAdd(&s_pending, 1); continue; }
            done[k] = 1;
            if (resolve_event(s, e, cfg, allow ? allow[k] != 0 : true))
                atomicAdd(&s_resolved, 1);
        }
        __syncthreads();
        int pending = s_pending;
        __syncthreads();
        if (pending == 0) break;
    }
    if (threadIdx.x == 0) h->n_merged = s_resolved;
}

// The engine's merge safeguards (engine/src/physics/collision_resolver.cpp:
// 46-76) for MergeOnly: which events may merge, decided in list order from
// the active flags and per-body merge counts alone (a merge keeps the
// survivor's mass positive and deactivates the absorbed body, so no
// arithmetic is needed).  One thread; it stops at the frame cap.
__global__ void k_merge_caps(int ne, const CollEvent* __restrict__ ev,
                             const double* __restrict__ m, uint8_t* __restrict__ act_copy,
                             int32_t* __restrict__ count, uint8_t* __restrict__ allow,
                             int max_frame, int max_body) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int total = 0;
    for (int k = 0; k < ne; k++) {
        const int a = ev[k].a, b = ev[k].b;
        if (!act_copy[a] || !act_copy[b]) continue;
        if (m[a] <= 0.0 || m[b] <= 0.0) continue;
        if (max_frame > 0 && total >= max_frame) break;
        if (max_body > 0 && (count[a] >= max_body || count[b] >= max_body)) continue;
        allow[k] = 1;
        act_copy[b] = 0;
        count[a]++;
        total++;
    }
}

__global__ void __launch_bounds__(256)
k_caps_init(int n, int ne, const uint8_t* __restrict__ act, uint8_t* __restrict__ act_copy,
            int32_t* __restrict__ count, uint8_t* __restrict__ allow) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { act_copy[i] = act[i]; count[i] = 0; }
    if (i < ne) allow[i] = 0;
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

int collide_resolve(Engine& e, const cm_collision_config& cc, int32_t* merged) {
    Coll* c = coll(e);
    if (merged) *merged = 0;
    if (!c || c->n_events == 0) return CM_OK;
    hipStream_t st = e.stream;
    const int ne = c->n_events;
    ResolveCfg cfg;
    cfg.mode = cc.mode;
    cfg.restitution = cc.default_restitution;
    cfg.mass_loss_cap = cc.fragmentation_mass_loss_cap;
    cfg.promote = cc.promote_compact_remnants;
    const uint8_t* allow = nullptr;
    if (cc.mode == CM_COLLISION_MERGE_ONLY &&
        (cc.max_merges_per_frame > 0 || cc.max_merges_per_body > 0)) {
        if (!c->caps_act) {
            size_t nb = (size_t)c->cap_bodies;
            if (hipMalloc(&c->caps_act, std::max<size_t>(nb, 256)) != hipSuccess ||
                hipMalloc(&c->caps_count, std::max<size_t>(nb * 4, 256)) != hipSuccess ||
                hipMalloc(&c->caps_allow, std::max<size_t>((size_t)c->cap_events, 256)) != hipSuccess)
                return set_error(e, CM_ERR_OUT_OF_MEMORY, "merge-cap scratch allocation failed");
        }
        const int nmax = std::max(e.count, ne);
        hipLaunchKernelGGL(k_caps_init, dim3((nmax + 255) / 256), dim3(256), 0, st, e.count, ne,
                           e.act, c->caps_act, c->caps_count, c->caps_allow);
        hipLaunchKernelGGL(k_merge_caps, dim3(1), dim3(64), 0, st, ne, c->ev, e.f[MASS],
                           c->caps_act, c->caps_count, c->caps_allow, cc.max_merges_per_frame,
                           cc.max_merges_per_body);
        allow = c->caps_allow;
    }
    hipLaunchKernelGGL(k_resolve, dim3(1), dim3(1024), 0, st, ne, c->ev, c->done,
                       c->first, soa_ptrs(e), cfg, allow, c->hdr);
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
    // each double array is scattered into the sort buffer ka (capacity x 8
    // bytes, the size of a body array), which then takes the array's place:
    // the old array becomes the next scratch, so every array moves once
    for (int k = 0; k < NF64; k++) {
        double* tmp = reinterpret_cast<double*>(c->ka);
        hipLaunchKernelGGL(k_scatter<double>, dim3(nb), dim3(256), 0, st, n, c->new_index,
                           e.f[k], tmp);
        c->ka = reinterpret_cast<uint64_t*>(e.f[k]);
        e.f[k] = tmp;
    }
    uint8_t* t8 = reinterpret_cast<uint8_t*>(c->ka);
    hipLaunchKernelGGL(k_scatter<uint8_t>, dim3(nb), dim3(256), 0, st, n, c->new_index, e.act, t8);
    CM_HIP_TRY(e, hipMemcpyAsync(e.act, t8, n, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_scatter<uint8_t>, dim3(nb), dim3(256), 0, st, n, c->new_index, e.col, t8);
    CM_HIP_TRY(e, hipMemcpyAsync(e.col, t8, n, hipMemcpyDeviceToDevice, st));
    int32_t* t32 = reinterpret_cast<int32_t*>(c->ka);
    hipLaunchKernelGGL(k_scatter<int32_t>, dim3(nb), dim3(256), 0, st, n, c->new_index, e.type, t32);
    CM_HIP_TRY(e, hipMemcpyAsync(e.type, t32, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    // the traversal's per-body work counts follow their bodies, so the next
    // evaluation's group order (heavy groups first, fat groups in parts)
    // reads each group's own work
    hipLaunchKernelGGL(k_scatter<int32_t>, dim3(nb), dim3(256), 0, st, n, c->new_index, e.bcost,
                       t32);
    CM_HIP_TRY(e, hipMemcpyAsync(e.bcost, t32, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
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
