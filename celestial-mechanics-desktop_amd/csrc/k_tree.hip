// k_tree.hip -- octree construction for the Barnes-Hut backend.
//
// Pipeline (all on one stream, no host round trip):
//   1. k_bbox / k_bbox_final: min/max over active bodies and the reference's
//      root cube (BarnesHutBackend.cs:150-170): centre (min+max)/2, half size
//      max extent / 2, clamped to >= 1e-10, times 1.001.
//   2. k_morton: levels x 3-bit Morton key per active body over that cube,
//      x bit most significant in each group, so the top group is the C#
//      octant number (OctreeNode.cs:185-192); inactive bodies get a sentinel
//      that sorts after every real key.
//   3. Stable LSD radix sort of (key, body index), 8-bit digits: per-tile
//      LDS digit histograms, one scan, and a scatter that ranks equal digits
//      with 64-lane ballots in original order (stable).
//   4. k_gather: sorted copies of position / mass / radius (coalesced reads
//      for the tree and the traversal).
//   5. k_build_level (one launch per level): level-order octree.  A node with
//      count > leaf_capacity below the deepest level gets its non-empty
//      children (binary searches on the sorted keys), allocated after the
//      whole level by a single-pass scan with decoupled look-back, so children
//      of a node are contiguous and node numbering is deterministic.
//   6. k_mass_level (deepest level first): leaf sums over its bodies in sorted
//      order, internal sums over its children in order; com = sum(m x) / M.
// Built with -ffp-contract=off: the box, keys and centres of mass are
// bit-identical to the CPU restatement of the same definitions.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cm_internal.h"

namespace cm {

constexpr int kBBoxBlocks = 1024;

// ---------------------------------------------------------------- bbox ----
__global__ void __launch_bounds__(256)
k_bbox(int n, const double* __restrict__ px, const double* __restrict__ py,
       const double* __restrict__ pz, const uint8_t* __restrict__ act,
       double* __restrict__ part, int32_t* __restrict__ cnt) {
    __shared__ double s[6][256];
    __shared__ int sc[256];
    double mn0 = 1.7976931348623157e308, mn1 = mn0, mn2 = mn0;
    double mx0 = -mn0, mx1 = -mn0, mx2 = -mn0;
    int c = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        if (!act[i]) continue;
        double x = px[i], y = py[i], z = pz[i];
        mn0 = x < mn0 ? x : mn0; mn1 = y < mn1 ? y : mn1; mn2 = z < mn2 ? z : mn2;
        mx0 = x > mx0 ? x : mx0; mx1 = y > mx1 ? y : mx1; mx2 = z > mx2 ? z : mx2;
        c++;
    }
    int t = threadIdx.x;
    s[0][t] = mn0; s[1][t] = mn1; s[2][t] = mn2;
    s[3][t] = mx0; s[4][t] = mx1; s[5][t] = mx2;
    sc[t] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) {
            for (int k = 0; k < 3; k++) {
                double a = s[k][t + o];
                if (a < s[k][t]) s[k][t] = a;
                double b = s[k + 3][t + o];
                if (b > s[k + 3][t]) s[k + 3][t] = b;
            }
            sc[t] += sc[t + o];
        }
        __syncthreads();
    }
    if (t == 0) {
        for (int k = 0; k < 6; k++) part[k * kBBoxBlocks + blockIdx.x] = s[k][0];
        cnt[blockIdx.x] = sc[0];
    }
}

__global__ void __launch_bounds__(1024)
k_bbox_final(int nblocks, const double* __restrict__ part,
             const int32_t* __restrict__ cnt, TreeHeader* __restrict__ hdr,
             int4* __restrict__ nd_info) {
    __shared__ double s[6][1024];
    __shared__ int sc[1024];
    int t = threadIdx.x;
    double big = 1.7976931348623157e308;
    for (int k = 0; k < 3; k++) {
        s[k][t] = t < nblocks ? part[k * kBBoxBlocks + t] : big;
        s[k + 3][t] = t < nblocks ? part[(k + 3) * kBBoxBlocks + t] : -big;
    }
    sc[t] = t < nblocks ? cnt[t] : 0;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (t < o) {
            for (int k = 0; k < 3; k++) {
                double a = s[k][t + o];
                if (a < s[k][t]) s[k][t] = a;
                double b = s[k + 3][t + o];
                if (b > s[k + 3][t]) s[k + 3][t] = b;
            }
            sc[t] += sc[t + o];
        }
        __syncthreads();
    }
    if (t < kMaxLevels + 2) hdr->tile_counter[t] = 0u;
    if (t == 0) {
        double minx = s[0][0], miny = s[1][0], minz = s[2][0];
        double maxx = s[3][0], maxy = s[4][0], maxz = s[5][0];
        double cx = (minx + maxx) * 0.5;
        double cy = (miny + maxy) * 0.5;
        double cz = (minz + maxz) * 0.5;
        double ex = maxx - minx, ey = maxy - miny, ez = maxz - minz;
        double e = ey > ez ? ey : ez;
        e = ex > e ? ex : e;
        double h = e * 0.5;
        h = h > 1e-10 ? h : 1e-10;
        h *= 1.001;
        hdr->box[0] = cx; hdr->box[1] = cy; hdr->box[2] = cz; hdr->box[3] = h;
        int na = sc[0];
        hdr->n_active = na;
        hdr->n_nodes = 1;
        hdr->overflow = 0;
        hdr->n_levels_used = 0;
        hdr->level_start[0] = 0;
        hdr->level_start[1] = 1;
        hdr->n_cell_int = 0ull;
        hdr->n_body_int = 0ull;
        hdr->n_groups = 0;
        hdr->work_next = 0u;
        nd_info[0] = make_int4(-1, 0, 0, na);
    }
}

// -------------------------------------------------------------- morton ----
__device__ __forceinline__ uint64_t spread3(uint64_t x) {
    x &= 0x1fffffull;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__device__ __forceinline__ uint32_t quantise(double x, double c, double h,
                                             int levels) {
    double cells = ldexp(1.0, levels);
    double scale = cells / (2.0 * h);
    double u = (x - (c - h)) * scale;
    if (!(u >= 0.0)) u = 0.0;
    if (u > cells - 1.0) u = cells - 1.0;
    return (uint32_t)u;
}

__global__ void __launch_bounds__(256)
k_morton(int n, int levels, const TreeHeader* __restrict__ hdr,
         const double* __restrict__ px, const double* __restrict__ py,
         const double* __restrict__ pz, const uint8_t* __restrict__ act,
         uint64_t* __restrict__ keys, int32_t* __restrict__ vals) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint64_t sentinel = (levels >= 21) ? ~0ull : ((2ull << (3 * levels)) - 1ull);
    uint64_t k = sentinel;
    if (act[i]) {
        double h = hdr->box[3];
        uint32_t qx = quantise(px[i], hdr->box[0], h, levels);
        uint32_t qy = quantise(py[i], hdr->box[1], h, levels);
        uint32_t qz = quantise(pz[i], hdr->box[2], h, levels);
        k = (spread3(qx) << 2) | (spread3(qy) << 1) | spread3(qz);
    }
    keys[i] = k;
    vals[i] = i;
}

// ---------------------------------------------------------- radix sort ----
__global__ void __launch_bounds__(kSortThreads)
k_sort_hist(int n, int shift, int ntiles, const uint64_t* __restrict__ keys,
            uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0u;
    __syncthreads();
    int base = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortTile / kSortThreads; r++) {
        int i = base + r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Block-wide exclusive scan of one value per thread (kSortThreads threads).
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_tmp,
                                                    uint32_t* total) {
    int t = threadIdx.x;
    s_tmp[t] = v;
    __syncthreads();
    for (int o = 1; o < kSortThreads; o <<= 1) {
        uint32_t x = t >= o ? s_tmp[t - o] : 0u;
        __syncthreads();
        s_tmp[t] += x;
        __syncthreads();
    }
    uint32_t incl = s_tmp[t];
    *total = s_tmp[kSortThreads - 1];
    __syncthreads();
    return incl - v;
}

// One block per digit: exclusive scan of that digit's per-tile counts along
// the tiles; the digit's total goes to totals[digit].
__global__ void __launch_bounds__(kSortThreads)
k_sort_scan(int ntiles, uint32_t* __restrict__ hist, uint32_t* __restrict__ totals) {
    __shared__ uint32_t s_tmp[kSortThreads];
    uint32_t* row = hist + (size_t)blockIdx.x * ntiles;
    int per = (ntiles + kSortThreads - 1) / kSortThreads;
    int b = threadIdx.x * per, e = min(ntiles, b + per);
    uint32_t s = 0;
    for (int i = b; i < e; i++) s += row[i];
    uint32_t tot;
    uint32_t run = block_excl_scan(s, s_tmp, &tot);
    for (int i = b; i < e; i++) {
        uint32_t v = row[i];
        row[i] = run;
        run += v;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kSortThreads)
k_sort_scatter(int n, int shift, int ntiles, const uint64_t* __restrict__ kin,
               const int32_t* __restrict__ vin, const uint32_t* __restrict__ hist,
               const uint32_t* __restrict__ totals,
               uint64_t* __restrict__ kout, int32_t* __restrict__ vout) {
    constexpr int kWaves = kSortThreads / kWave;
    __shared__ uint32_t base[256];
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[kWaves][256];
    int t = threadIdx.x;
    int lane = t & (kWave - 1);
    int w = t / kWave;
    {   // digit base = all smaller digits' totals + this digit in earlier tiles
        uint32_t tot;
        uint32_t dbase = block_excl_scan(totals[t], run, &tot);
        base[t] = dbase + hist[(size_t)t * ntiles + blockIdx.x];
    }
    run[t] = 0u;
    uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int tile0 = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortTile / kSortThreads; r++) {
        for (int q = 0; q < kWaves; q++) wcnt[q][t] = 0u;
        __syncthreads();
        int i = tile0 + r * kSortThreads + t;
        bool valid = i < n;
        uint64_t key = valid ? kin[i] : 0ull;
        int32_t val = valid ? vin[i] : 0;
        uint32_t d = (uint32_t)(key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < 8; b++) {
            bool bit = (d >> b) & 1u;
            uint64_t bal = __ballot(valid && bit);
            peers &= bit ? bal : ~bal;
        }
        uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {   // thread t owns digit t: offsets of each wave within the round
            uint32_t acc = run[t];
            for (int q = 0; q < kWaves; q++) {
                uint32_t c = wcnt[q][t];
                wcnt[q][t] = acc;
                acc += c;
            }
            run[t] = acc;
        }
        __syncthreads();
        if (valid) {
            uint32_t pos = base[d] + wcnt[w][d] + rank;
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256)
k_gather(int n, const int32_t* __restrict__ order, const double* __restrict__ px,
         const double* __restrict__ py, const double* __restrict__ pz,
         const double* __restrict__ m, const double* __restrict__ r,
         double* __restrict__ sorted, int cap) {
    int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    int o = order[k];
    sorted[k] = px[o];
    sorted[(size_t)cap + k] = py[o];
    sorted[2 * (size_t)cap + k] = pz[o];
    sorted[3 * (size_t)cap + k] = m[o];
    sorted[4 * (size_t)cap + k] = r[o];
}

// ---------------------------------------------------------- tree build ----
__device__ __forceinline__ int lower_digit(const uint64_t* __restrict__ keys,
                                           int lo, int hi, int shift, uint32_t d) {
    // first index in [lo, hi) whose digit at `shift` is >= d
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (((uint32_t)(keys[mid] >> shift) & 7u) < d) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ int wave_incl_scan(int v) {
    int lane = threadIdx.x & (kWave - 1);
    for (int o = 1; o < kWave; o <<= 1) {
        int t = __shfl_up(v, o, kWave);
        if (lane >= o) v += t;
    }
    return v;
}

__global__ void __launch_bounds__(kBuildTile)
k_build_level(int L, int levels, int leaf_cap, int node_cap, int btiles,
              TreeHeader* __restrict__ hdr, int4* __restrict__ info,
              const uint64_t* __restrict__ keys,
              unsigned long long* __restrict__ lookback) {
    constexpr int kWaves = kBuildTile / kWave;
    __shared__ int s_tile;
    __shared__ int s_wsum[kWaves];
    __shared__ int s_prefix;
    const int ls = hdr->level_start[L];
    const int le = hdr->level_start[L + 1];
    const int size = le - ls;
    const int ntiles = (size + kBuildTile - 1) / kBuildTile;
    int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
    if (ntiles == 0) {
        if (blockIdx.x == 0 && t == 0) {
            hdr->level_start[L + 2] = le;
        }
        return;
    }
    unsigned long long* status = lookback + (size_t)L * btiles;
    for (;;) {
        if (t == 0) s_tile = (int)atomicAdd(&hdr->tile_counter[L], 1u);
        __syncthreads();
        int tile = s_tile;
        __syncthreads();
        if (tile >= ntiles) break;
        int k = ls + tile * kBuildTile + t;
        bool valid = k < le;
        int cs[8], cc[8];
        int nch = 0;
        int4 me = valid ? info[k] : make_int4(0, 0, 0, 0);
        if (valid && me.w > leaf_cap && L < levels) {
            int shift = 3 * (levels - L - 1);
            int lo = me.z, hi = me.z + me.w;
            // the 7 boundaries (first index with digit >= d) are searched
            // together, one step of each per round, so their loads overlap
            int b[9], blo[8], bhi[8];
            b[0] = lo; b[8] = hi;
            for (int d = 1; d < 8; d++) { blo[d] = lo; bhi[d] = hi; }
            for (int round = 0; round < 32; round++) {
                bool any = false;
                for (int d = 1; d < 8; d++) {
                    if (blo[d] < bhi[d]) {
                        int mid = (blo[d] + bhi[d]) >> 1;
                        if (((uint32_t)(keys[mid] >> shift) & 7u) < (uint32_t)d) blo[d] = mid + 1;
                        else bhi[d] = mid;
                        any = true;
                    }
                }
                if (!any) break;
            }
            for (int d = 1; d < 8; d++) b[d] = blo[d];
            for (int d = 0; d < 8; d++) {
                int c = b[d + 1] - b[d];
                if (c > 0) { cs[nch] = b[d]; cc[nch] = c; nch++; }
            }
        }
        // block exclusive scan of nch
        int incl = wave_incl_scan(nch);
        if (lane == kWave - 1) s_wsum[w] = incl;
        __syncthreads();
        int woff = 0, agg = 0;
        for (int q = 0; q < kWaves; q++) {
            if (q < w) woff += s_wsum[q];
            agg += s_wsum[q];
        }
        int excl = woff + incl - nch;
        // decoupled look-back over tiles of this level, 64 predecessors per
        // round (one per lane of the first wavefront)
        if (w == 0) {
            unsigned long long prefix = 0ull;
            if (tile == 0) {
                if (lane == 0)
                    __hip_atomic_store(&status[0], (2ull << 32) | (unsigned)agg,
                                       __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                if (lane == 0)
                    __hip_atomic_store(&status[tile], (1ull << 32) | (unsigned)agg,
                                       __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                int p = tile - 1;
                for (;;) {
                    int idx = p - lane;
                    unsigned long long v = idx >= 0
                        ? __hip_atomic_load(&status[idx], __ATOMIC_ACQUIRE,
                                            __HIP_MEMORY_SCOPE_AGENT)
                        : (2ull << 32);
                    unsigned flag = (unsigned)(v >> 32);
                    if (__ballot(flag == 0u)) continue;     // a predecessor not yet published
                    uint64_t incl = __ballot(flag == 2u);
                    int stop = incl ? __ffsll((unsigned long long)incl) - 1 : kWave - 1;
                    unsigned long long mine = lane <= stop ? (v & 0xffffffffull) : 0ull;
                    for (int o = 32; o > 0; o >>= 1) mine += __shfl_xor(mine, o, kWave);
                    prefix += mine;
                    if (incl) break;
                    p -= kWave;
                }
                if (lane == 0)
                    __hip_atomic_store(&status[tile],
                                       (2ull << 32) | (unsigned)(prefix + (unsigned)agg),
                                       __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
            if (lane == 0) s_prefix = (int)prefix;
        }
        __syncthreads();
        int first = le + s_prefix + excl;
        if (valid) {
            if (nch > 0 && first + nch <= node_cap) {
                info[k] = make_int4(first, nch | (L << 8), me.z, me.w);
                for (int c = 0; c < nch; c++)
                    info[first + c] = make_int4(-1, (L + 1) << 8, cs[c], cc[c]);
            } else {
                if (nch > 0) atomicExch(&hdr->overflow, 1);
                info[k] = make_int4(-1, L << 8, me.z, me.w);
            }
        }
        if (tile == ntiles - 1 && t == 0) {
            int end = le + s_prefix + agg;
            if (end > node_cap) end = node_cap;
            hdr->level_start[L + 2] = end;
            hdr->n_nodes = end;
            if (agg > 0) hdr->n_levels_used = L + 2;
        }
        __syncthreads();
    }
}

// Bottom-up mass, centre of mass and body bounding box for the nodes of
// level L (box: 6 doubles lo x, y, z, hi x, y, z per node).
__global__ void __launch_bounds__(256)
k_mass_level(int L, const TreeHeader* __restrict__ hdr,
             const int4* __restrict__ info, const double* __restrict__ sorted,
             int cap, double4* __restrict__ cm, double* __restrict__ sums,
             double* __restrict__ box) {
    const int ls = hdr->level_start[L];
    const int le = hdr->level_start[L + 1];
    const double* sx = sorted;
    const double* sy = sorted + cap;
    const double* sz = sorted + 2 * (size_t)cap;
    const double* sm = sorted + 3 * (size_t)cap;
    const double big = 1.7976931348623157e308;
    for (int k = ls + blockIdx.x * 256 + threadIdx.x; k < le; k += gridDim.x * 256) {
        int4 nd = info[k];
        double M = 0.0, ax = 0.0, ay = 0.0, az = 0.0;
        double lo0 = big, lo1 = big, lo2 = big, hi0 = -big, hi1 = -big, hi2 = -big;
        if (nd.x < 0) {
            for (int q = nd.z; q < nd.z + nd.w; q++) {
                double mb = sm[q];
                double x = sx[q], y = sy[q], z = sz[q];
                M += mb;
                ax += mb * x;
                ay += mb * y;
                az += mb * z;
                lo0 = x < lo0 ? x : lo0; lo1 = y < lo1 ? y : lo1; lo2 = z < lo2 ? z : lo2;
                hi0 = x > hi0 ? x : hi0; hi1 = y > hi1 ? y : hi1; hi2 = z > hi2 ? z : hi2;
            }
        } else {
            int nch = nd.y & 255;
            for (int c = nd.x; c < nd.x + nch; c++) {
                M += cm[c].w;
                ax += sums[3 * (size_t)c];
                ay += sums[3 * (size_t)c + 1];
                az += sums[3 * (size_t)c + 2];
                const double* cb = box + 6 * (size_t)c;
                lo0 = cb[0] < lo0 ? cb[0] : lo0; lo1 = cb[1] < lo1 ? cb[1] : lo1;
                lo2 = cb[2] < lo2 ? cb[2] : lo2; hi0 = cb[3] > hi0 ? cb[3] : hi0;
                hi1 = cb[4] > hi1 ? cb[4] : hi1; hi2 = cb[5] > hi2 ? cb[5] : hi2;
            }
        }
        double* kb = box + 6 * (size_t)k;
        kb[0] = lo0; kb[1] = lo1; kb[2] = lo2; kb[3] = hi0; kb[4] = hi1; kb[5] = hi2;
        sums[3 * (size_t)k] = ax;
        sums[3 * (size_t)k + 1] = ay;
        sums[3 * (size_t)k + 2] = az;
        double4 r;
        if (M > 0.0) { r.x = ax / M; r.y = ay / M; r.z = az / M; }
        else { r.x = 0.0; r.y = 0.0; r.z = 0.0; }
        r.w = M;
        cm[k] = r;
    }
}

__global__ void k_fill_u64(size_t n, unsigned long long* __restrict__ p) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * 256) p[i] = 0ull;
}

// Stable LSD radix sort of n (key, value) pairs over 8-bit digits
// first_pass .. first_pass + passes - 1.  On return *k/*v hold the sorted
// pairs (the buffers are swapped after every pass); hist holds
// 256 * (ceil(n / kSortTile) + 1) words.
void radix_sort_pairs(hipStream_t st, int n, int first_pass, int passes,
                      uint64_t** k, int32_t** v, uint64_t** kt, int32_t** vt,
                      uint32_t* hist) {
    if (n <= 0) return;
    int ntiles = (n + kSortTile - 1) / kSortTile;
    uint32_t* totals = hist + (size_t)256 * ntiles;
    for (int p = first_pass; p < first_pass + passes; p++) {
        int shift = 8 * p;
        hipLaunchKernelGGL(k_sort_hist, dim3(ntiles), dim3(kSortThreads), 0, st, n,
                           shift, ntiles, *k, hist);
        hipLaunchKernelGGL(k_sort_scan, dim3(256), dim3(kSortThreads), 0, st, ntiles,
                           hist, totals);
        hipLaunchKernelGGL(k_sort_scatter, dim3(ntiles), dim3(kSortThreads), 0, st,
                           n, shift, ntiles, *k, *v, hist, totals, *kt, *vt);
        std::swap(*k, *kt);
        std::swap(*v, *vt);
    }
}

// Builds the tree for the current state; leaves keys/order in keys_a/vals_a
// (or keys_b/vals_b when the pass count is odd; *keys_out / *order_out).
int build_tree(Engine& e, const uint64_t** keys_out, const int32_t** order_out) {
    const int n = e.count;
    const int levels = e.cfg.morton_levels;
    hipStream_t st = e.stream;
    int nb = std::min(kBBoxBlocks, std::max(1, (n + 255) / 256));
    hipLaunchKernelGGL(k_bbox, dim3(nb), dim3(256), 0, st, n, e.f[PX], e.f[PY],
                       e.f[PZ], e.act, e.bbox_part, e.bbox_cnt);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(1024), 0, st, nb, e.bbox_part,
                       e.bbox_cnt, e.hdr, e.nd_info);
    int gb = (n + 255) / 256;
    hipLaunchKernelGGL(k_morton, dim3(gb), dim3(256), 0, st, n, levels, e.hdr,
                       e.f[PX], e.f[PY], e.f[PZ], e.act, e.keys_a, e.vals_a);
    if (e.timing) CM_HIP_TRY(e, hipEventRecord(e.ev[1], st));
    int passes = (3 * levels + 1 + 7) / 8;
    uint64_t* ka = e.keys_a; uint64_t* kb = e.keys_b;
    int32_t* va = e.vals_a; int32_t* vb = e.vals_b;
    radix_sort_pairs(st, n, 0, passes, &ka, &va, &kb, &vb, e.hist);
    if (e.timing) CM_HIP_TRY(e, hipEventRecord(e.ev[2], st));
    hipLaunchKernelGGL(k_gather, dim3(gb), dim3(256), 0, st, n, va, e.f[PX],
                       e.f[PY], e.f[PZ], e.f[MASS], e.f[RAD], e.sorted, e.capacity);
    int btiles = (e.node_cap + kBuildTile - 1) / kBuildTile;
    size_t nlb = (size_t)(levels + 1) * btiles;
    hipLaunchKernelGGL(k_fill_u64, dim3(1024), dim3(256), 0, st, nlb, e.lookback);
    for (int L = 0; L <= levels; L++) {
        double maxsz = std::min((double)n, L < 20 ? std::ldexp(1.0, 3 * L) : 1e18);
        int g = (int)std::min(2048.0, std::max(1.0, std::ceil(maxsz / kBuildTile)));
        hipLaunchKernelGGL(k_build_level, dim3(g), dim3(kBuildTile), 0, st, L,
                           levels, e.cfg.leaf_capacity, e.node_cap, btiles, e.hdr,
                           e.nd_info, ka, e.lookback);
    }
    for (int L = levels; L >= 0; L--) {
        double maxsz = std::min((double)n, L < 20 ? std::ldexp(1.0, 3 * L) : 1e18);
        int g = (int)std::min(1024.0, std::max(1.0, std::ceil(maxsz / 256)));
        hipLaunchKernelGGL(k_mass_level, dim3(g), dim3(256), 0, st, L, e.hdr,
                           e.nd_info, e.sorted, e.capacity, e.nd_cm, e.nd_sum, e.nd_box);
    }
    CM_HIP_TRY(e, hipGetLastError());
    *keys_out = ka;
    *order_out = va;
    return CM_OK;
}

}  // namespace cm
