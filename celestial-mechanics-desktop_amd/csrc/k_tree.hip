// k_tree.hip -- octree construction for the Barnes-Hut backend.
//
// Pipeline (all on one stream, no host round trip):
//   1. k_bbox / k_bbox_final: min/max over active bodies and the reference's
//      root cube (BarnesHutBackend.cs:150-170): centre (min+max)/2, half size
//      max extent / 2, clamped to >= 1e-10, times 1.001.
//   2. k_morton: levels x 3-bit Morton key per active body over that cube,
//      x bit most significant in each group, so the top group is the C#
//      octant number (OctreeNode.cs:185-192); inactive bodies get a sentinel
//      that sorts after every real key.  Keys are 32-bit words up to 10
//      levels.  The kernel also counts the first sort pass's digits.
//   3. Stable LSD radix sort of (key, body index), 8-bit digits: per-tile
//      LDS digit histograms, one scan, and a scatter that ranks equal digits
//      with 64-lane ballots in original order (stable).
//   4. k_gather_count: sorted copies of position / mass (and radius, which
//      only the shell theorem reads) for the tree and the traversal, and the
//      per-tile node counts of step 5.
//   5. k_gather_count / k_tree_scan / k_tree_emit: level-order octree from
//      local comparisons of neighbouring keys (below).  A node with
//      count > leaf_capacity below the deepest level has its non-empty
//      children; children of a node are contiguous, every level is
//      contiguous, numbering is deterministic.  Leaves sum their bodies in
//      sorted order while the tile's bodies sit in LDS.
//   6. k_tree_internal (deepest level first): internal sums over the
//      children in order; com = sum(m x) / M.
// Built with -ffp-contract=off: the box, keys and centres of mass are
// bit-identical to the CPU restatement of the same definitions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "cm_internal.h"
#include "cm_testing.h"

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
    const long long stride = (long long)gridDim.x * 256;
    for (long long i0 = (long long)blockIdx.x * 256 + threadIdx.x; i0 < n; i0 += 4 * stride) {
        // the loads of a thread's next four bodies are issued before any is
        // folded; the fold takes them in index order, as a plain loop would
        uint8_t a[4];
        double x[4], y[4], z[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const long long i = i0 + u * stride;
            const bool in = i < n;
            a[u] = in ? act[i] : 0;
            x[u] = in ? px[i] : 0.0;
            y[u] = in ? py[i] : 0.0;
            z[u] = in ? pz[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (!a[u]) continue;
            mn0 = x[u] < mn0 ? x[u] : mn0; mn1 = y[u] < mn1 ? y[u] : mn1;
            mn2 = z[u] < mn2 ? z[u] : mn2;
            mx0 = x[u] > mx0 ? x[u] : mx0; mx1 = y[u] > mx1 ? y[u] : mx1;
            mx2 = z[u] > mx2 ? z[u] : mx2;
            c++;
        }
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
        hdr->n_far_int = 0ull;
        hdr->n_groups = 0;
        for (int q = 0; q < kXcds; q++) hdr->work_next[q] = 0u;
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

// Keys are stored as K: 32-bit words when 3 levels + 1 sentinel bit fit
// (levels <= 10), else 64-bit.  One block per radix-sort tile (kSortTile
// bodies, eight per thread): it also counts the tile's lowest key digit, the
// histogram column of the sort's first pass.
template <typename K>
__global__ void __launch_bounds__(kSortThreads)
k_morton(int n, int levels, int ntiles, const TreeHeader* __restrict__ hdr,
         const double* __restrict__ px, const double* __restrict__ py,
         const double* __restrict__ pz, const double* __restrict__ m,
         const uint8_t* __restrict__ act, K* __restrict__ keys,
         int32_t* __restrict__ vals, double4* __restrict__ pm4,
         int32_t* __restrict__ gmark, int32_t* __restrict__ gtile,
         uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[256];
    const int t = threadIdx.x;
    h[t] = 0u;
    // the target-group counters of this block's group tiles (counted by k_groups)
    if (t < kSortTile / kGroupTile) {
        const int g = blockIdx.x * (kSortTile / kGroupTile) + t;
        if (g * kGroupTile < n) gtile[g] = 0;
    }
    __syncthreads();
    const uint64_t sentinel = (levels >= 21) ? ~0ull : ((2ull << (3 * levels)) - 1ull);
    const double bx = hdr->box[0], by = hdr->box[1], bz = hdr->box[2], bh = hdr->box[3];
    for (int r = 0; r < kSortTile / kSortThreads; r++) {
        const int i = blockIdx.x * kSortTile + r * kSortThreads + t;
        if (i >= n) break;
        gmark[i] = 0;   // target-group marks of the traversal, set after the build
        uint64_t k = sentinel;
        const double x = px[i], y = py[i], z = pz[i];
        if (act[i]) {
            const uint32_t qx = quantise(x, bx, bh, levels);
            const uint32_t qy = quantise(y, by, bh, levels);
            const uint32_t qz = quantise(z, bz, bh, levels);
            k = (spread3(qx) << 2) | (spread3(qy) << 1) | spread3(qz);
        }
        keys[i] = (K)k;
        vals[i] = i;
        // position and mass packed per body, so that the gather after the sort
        // reads one 32-byte record per body instead of four scattered words
        pm4[i] = make_double4(x, y, z, m[i]);
        atomicAdd(&h[(uint32_t)k & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)t * ntiles + blockIdx.x] = h[t];
}

// ---------------------------------------------------------- radix sort ----
template <typename K>
__global__ void __launch_bounds__(kSortThreads)
k_sort_hist(int n, int shift, int ntiles, const K* __restrict__ keys,
            uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const int base = blockIdx.x * kSortTile;
    constexpr int kR = kSortTile / kSortThreads;
    K k[kR];   // all loads of the tile in flight before the first count
#pragma unroll
    for (int r = 0; r < kR; r++) {
        const int i = base + r * kSortThreads + threadIdx.x;
        k[r] = i < n ? keys[i] : (K)0;
    }
#pragma unroll
    for (int r = 0; r < kR; r++)
        if (base + r * kSortThreads + (int)threadIdx.x < n)
            atomicAdd(&h[(uint32_t)(k[r] >> shift) & 255u], 1u);
    __syncthreads();
    hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Inclusive prefix sum over the 64 lanes of a wave (all lanes active): DPP
// row shifts by 1, 2, 4, 8 within each row of 16 lanes, then the broadcasts
// of lane 15 into row 1 and 3 and of lane 31 into rows 2 and 3.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    int x = (int)v;
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);
    return (uint32_t)x;
}

// Block-wide exclusive scan of one value per thread (kSortThreads threads,
// all active); s_tmp holds one word per wave.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_tmp,
                                                    uint32_t* total) {
    constexpr int kWaves = kSortThreads / kWave;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const uint32_t incl = wave_incl_scan(v);
    if (lane == kWave - 1) s_tmp[w] = incl;
    __syncthreads();
    uint32_t off = 0u, tot = 0u;
#pragma unroll
    for (int q = 0; q < kWaves; q++) {
        const uint32_t x = s_tmp[q];
        off += q < w ? x : 0u;
        tot += x;
    }
    *total = tot;
    __syncthreads();
    return off + incl - v;
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

// Stable scatter of one tile.  Wave w owns the contiguous tile positions
// w * 512 .. w * 512 + 511 and ranks them round by round (64 keys per round)
// against digit counters private to the wave, so the ranking needs no
// block barrier: within a round, lanes with equal digits find each other by
// eight ballots and the lowest of them advances the wave's counter.  One
// barrier later every digit's wave offsets and the tile's digit offsets are
// known; the keys are placed in LDS in digit order and written out from
// there, so that consecutive lanes write consecutive addresses of each
// digit's run.  Destination = digit base + earlier tiles' count of the digit
// + rank in input order.
template <typename K>
__global__ void __launch_bounds__(kSortThreads)
k_sort_scatter(int n, int shift, int ntiles, const K* __restrict__ kin,
               const int32_t* __restrict__ vin, const uint32_t* __restrict__ hist,
               const uint32_t* __restrict__ totals, K* __restrict__ kout,
               int32_t* __restrict__ vout) {
    constexpr int kWaves = kSortThreads / kWave;
    constexpr int kR = kSortTile / kSortThreads;   // keys per lane
    constexpr int kSpan = kSortTile / kWaves;      // tile positions per wave
    __shared__ uint32_t base[256];
    __shared__ uint32_t loff[256];
    __shared__ uint32_t s_tmp[kWaves];
    __shared__ uint32_t wcnt[kWaves][256];   // each wave's digit counts, then offsets
    __shared__ K s_key[kSortTile];
    __shared__ int32_t s_val[kSortTile];
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int w = t / kWave;
    const int tile0 = blockIdx.x * kSortTile;
    const int cnt = min(kSortTile, n - tile0);
    const int p0 = w * kSpan;
    K key[kR];
    int32_t val[kR];
#pragma unroll
    for (int r = 0; r < kR; r++) {
        const int p = p0 + r * kWave + lane;
        key[r] = p < cnt ? kin[tile0 + p] : (K)0;
        val[r] = p < cnt ? vin[tile0 + p] : 0;
    }
    for (int q = lane; q < 256; q += kWave) wcnt[w][q] = 0u;
    {   // digit base = all smaller digits' totals + this digit in earlier tiles
        uint32_t tot;
        const uint32_t dbase = block_excl_scan(totals[t], s_tmp, &tot);
        base[t] = dbase + hist[(size_t)t * ntiles + blockIdx.x];
    }
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t lrank[kR];
#pragma unroll
    for (int r = 0; r < kR; r++) {
        const bool valid = p0 + r * kWave + lane < cnt;
        const uint32_t d = (uint32_t)(key[r] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(valid && bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        // the lowest lane of each digit reads and advances the wave's counter
        uint32_t before = 0u;
        if (valid && rank == 0) {
            before = wcnt[w][d];
            wcnt[w][d] = before + (uint32_t)__popcll(peers);
        }
        const int lead = (int)__builtin_ctzll(peers | (1ull << 63));
        lrank[r] = __shfl(before, lead, kWave) + rank;
    }
    __syncthreads();
    uint32_t dcount = 0u;   // thread t owns digit t: wave offsets, tile count
#pragma unroll
    for (int q = 0; q < kWaves; q++) {
        const uint32_t c = wcnt[q][t];
        wcnt[q][t] = dcount;
        dcount += c;
    }
    {
        uint32_t tot;
        const uint32_t o = block_excl_scan(dcount, s_tmp, &tot);
        loff[t] = o;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kR; r++) {
        if (p0 + r * kWave + lane < cnt) {
            const uint32_t d = (uint32_t)(key[r] >> shift) & 255u;
            const uint32_t slot = loff[d] + wcnt[w][d] + lrank[r];
            s_key[slot] = key[r];
            s_val[slot] = val[r];
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kR; r++) {
        const int q = r * kSortThreads + t;
        if (q < cnt) {
            const K k = s_key[q];
            const uint32_t d = (uint32_t)(k >> shift) & 255u;
            const uint32_t pos = base[d] + (uint32_t)q - loff[d];
            kout[pos] = k;
            vout[pos] = s_val[q];
        }
    }
}

// ---------------------------------------------------------- tree build ----
// The level-order octree follows from two local facts of the sorted keys:
//   d[i] = leading 3-bit digits key[i] shares with key[i - 1] (0 for i = 0):
//          body i opens a new cell at every level l > d[i];
//   D[j] = leading digits key[j] shares with key[j + cap] (-1 when j + cap is
//          not an active body): bodies j .. j + cap all lie in one level-m
//          cell, which then holds more than cap bodies, iff D[j] >= m.
// The level-m cell of body i holds more than cap bodies iff one of the
// windows [j, j + cap] with i - cap <= j <= i lies inside it, i.e. iff
// max D[i - cap .. i] >= m.  A cell is a node iff its parent cell holds more
// than cap bodies (a node with more than cap bodies below the deepest level
// is split into its non-empty children, and counts only shrink going down),
// so body i opens exactly the nodes of levels d[i] + 1 .. lam[i] with
// lam[i] = min(levels, 1 + max D[i - cap .. i]); the one of level lam[i] is a
// leaf.  Nodes are numbered level by level in order of their first body (the
// digit order of every parent's children), which is a per-level prefix count
// of opened nodes over the bodies: k_gather_count counts them per tile,
// k_tree_scan turns the counts into offsets, k_tree_emit writes the nodes and
// the leaves' sums, k_tree_internal forms internal nodes bottom-up.
constexpr int kTreeWaves = kTreeTile / kWave;
// bodies staged past each tile: >= the largest leaf capacity (the D window of
// k_gather_count), and large enough that deepest-level leaves, which may hold
// more bodies than the leaf capacity (up to ~120 in the core of a 2^20-body
// Plummer sphere with 30-bit keys), are usually summed from LDS
constexpr int kTreeHalo = 128;

__device__ __forceinline__ int common_digits(uint64_t a, uint64_t b, int levels) {
    uint64_t x = a ^ b;
    if (x == 0ull) return levels;
    int p = 63 - __clzll((long long)x);
    return (3 * levels - 1 - p) / 3;
}

// One tree tile of the sorted order per block: the sorted copies of x, y, z,
// m and, when the shell theorem needs source radii (r != nullptr), of the
// radius for the tile's bodies (of all count bodies, inactive ones last),
// then d, lam and the tile's node count per level over its active bodies.
template <typename K>
__global__ void __launch_bounds__(kTreeTile)
k_gather_count(int count, int levels, int cap, int tstride, const TreeHeader* __restrict__ hdr,
               const K* __restrict__ keys, const int32_t* __restrict__ order,
               const double4* __restrict__ pm4, const double* __restrict__ r,
               double* __restrict__ sorted, int bcap, uchar2* __restrict__ dl,
               uint32_t* __restrict__ tcount) {
    __shared__ signed char sD[kTreeTile + kTreeHalo];
    __shared__ uint32_t s_cnt[kMaxLevels + 1][kTreeWaves];
    const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
    const int t0 = blockIdx.x * kTreeTile;
    if (t0 + t < count) {
        const int k = t0 + t;
        const int o = order[k];
        const double4 v = pm4[o];
        sorted[k] = v.x;
        sorted[(size_t)bcap + k] = v.y;
        sorted[2 * (size_t)bcap + k] = v.z;
        sorted[3 * (size_t)bcap + k] = v.w;
        if (r) sorted[4 * (size_t)bcap + k] = r[o];
    }
    const int n = hdr->n_active;
    // D for j in [t0 - cap, t0 + kTreeTile)
    for (int q = t; q < kTreeTile + cap; q += kTreeTile) {
        int j = t0 - cap + q;
        int D = -1;
        if (j >= 0 && j + cap < n) D = common_digits(keys[j], keys[j + cap], levels);
        sD[q] = (signed char)D;
    }
    __syncthreads();
    const int i = t0 + t;
    int d = 0, lam = 0;
    if (i < n) {
        d = i == 0 ? 0 : common_digits(keys[i - 1], keys[i], levels);
        int m = -1;
        for (int q = t; q <= t + cap; q++) m = max(m, (int)sD[q]);
        lam = min(levels, m + 1);
        dl[i] = make_uchar2((unsigned char)d, (unsigned char)lam);
    }
    for (int l = 1; l <= levels; l++) {
        uint64_t mask = __ballot(d < l && l <= lam);
        if (lane == 0) s_cnt[l][w] = (uint32_t)__popcll(mask);
    }
    __syncthreads();
    if (t >= 1 && t <= levels) {
        uint32_t s = 0u;
        for (int q = 0; q < kTreeWaves; q++) s += s_cnt[t][q];
        tcount[(size_t)t * tstride + blockIdx.x] = s;
    }
}

// One block: per level, exclusive scan of the tile counts (in place), level
// extents, node total, overflow flag and the root's record.
__global__ void __launch_bounds__(1024)
k_tree_scan(int levels, int cap, int ntiles, int tstride, int node_cap,
            TreeHeader* __restrict__ hdr, uint32_t* __restrict__ tcount,
            int4* __restrict__ info) {
    __shared__ uint32_t s_w[kMaxLevels + 1][16];
    __shared__ uint32_t s_carry[kMaxLevels + 1];
    const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
    if (t <= kMaxLevels) s_carry[t] = 0u;
    for (int base = 0; base < ntiles; base += 1024) {
        // every level's count of this thread's tile is loaded at once
        const int q = base + t;
        uint32_t v[kMaxLevels + 1];
#pragma unroll
        for (int l = 1; l <= kMaxLevels; l++)
            v[l] = (l <= levels && q < ntiles) ? tcount[(size_t)l * tstride + q] : 0u;
        uint32_t x[kMaxLevels + 1];
#pragma unroll
        for (int l = 1; l <= kMaxLevels; l++) {
            uint32_t a = v[l];
            if (l <= levels) a = wave_incl_scan(a);   // uniform: all lanes active
            x[l] = a;
            if (lane == kWave - 1) s_w[l][w] = a;
        }
        __syncthreads();
#pragma unroll
        for (int l = 1; l <= kMaxLevels; l++) {
            if (l > levels) continue;
            uint32_t woff = 0u;
            for (int k = 0; k < w; k++) woff += s_w[l][k];
            if (q < ntiles) tcount[(size_t)l * tstride + q] = s_carry[l] + woff + x[l] - v[l];
        }
        __syncthreads();
        if (t >= 1 && t <= levels) {
            uint32_t tot = 0u;
            for (int k = 0; k < 16; k++) tot += s_w[t][k];
            s_carry[t] += tot;
        }
        __syncthreads();
    }
    long long running = 1;   // the root
    int used = 1;
    for (int l = 1; l <= levels; l++) {
        if (t == 0) hdr->level_start[l] = (int)min(running, (long long)INT32_MAX);
        uint32_t total = s_carry[l];
        running += total;
        if (total > 0u) used = l + 1;
    }
    if (t == 0) {
        int end = (int)min(running, (long long)INT32_MAX);
        for (int l = levels + 1; l < kMaxLevels + 3; l++) hdr->level_start[l] = end;
        int over = running > node_cap ? 1 : 0;
        hdr->overflow = over;
        if (over) hdr->overflow_run = 1;
        hdr->n_nodes = over ? node_cap : end;
        hdr->n_levels_used = used;
        int na = hdr->n_active;
        info[0] = make_int4(na > cap ? 1 : -1, 0, 0, na);
    }
}

// The mass and moment sums run in sorted order (bit-identical to the CPU
// restatement); the box is a min / max, which no order changes, taken with
// v_min_f64 / v_max_f64 (one instruction each instead of a compare and two
// selects; they can differ from the compare only in the sign of a zero,
// which changes no box extent or volume computed from it).
__device__ __forceinline__ void leaf_accumulate(double x, double y, double z, double mb,
                                                double& M, double& ax, double& ay,
                                                double& az, double lo[3], double hi[3]) {
    M += mb;
    ax += mb * x;
    ay += mb * y;
    az += mb * z;
    lo[0] = fmin(x, lo[0]); lo[1] = fmin(y, lo[1]); lo[2] = fmin(z, lo[2]);
    hi[0] = fmax(x, hi[0]); hi[1] = fmax(y, hi[1]); hi[2] = fmax(z, hi[2]);
}

__device__ __forceinline__ void store_node_sums(int k, double M, double ax, double ay,
                                                double az, const double lo[3],
                                                const double hi[3], double4* cm,
                                                double* sums, double* box) {
    double* kb = box + 6 * (size_t)k;
    kb[0] = lo[0]; kb[1] = lo[1]; kb[2] = lo[2]; kb[3] = hi[0]; kb[4] = hi[1]; kb[5] = hi[2];
    sums[3 * (size_t)k] = ax;
    sums[3 * (size_t)k + 1] = ay;
    sums[3 * (size_t)k + 2] = az;
    double4 r;
    if (M > 0.0) { r.x = ax / M; r.y = ay / M; r.z = az / M; }
    else { r.x = 0.0; r.y = 0.0; r.z = 0.0; }
    r.w = M;
    cm[k] = r;
}

// Writes every node below the root: start, parent, first child (internal
// nodes; their count, child count and sums come from k_tree_internal), and
// for leaves the count, mass, sum m x, centre of mass and body box, summed
// in sorted order.  Two phases per tile: every body writes the nodes it
// opens and, when it opens a leaf, queues it (in order of its first body);
// then the tile's leaves are taken one per thread, so that a wave sums up to
// 64 leaves instead of the few that start among its 64 bodies (deepest-level
// leaves hold up to ~120 bodies in the core of a 2^20-body Plummer sphere,
// and a wave runs as long as its longest leaf).  The bodies of the tile and
// a halo past it are staged in LDS.
__global__ void __launch_bounds__(kTreeTile)
k_tree_emit(int levels, int node_cap, int tstride, const TreeHeader* __restrict__ hdr,
            const uchar2* __restrict__ dl, const uint32_t* __restrict__ toff,
            const double* __restrict__ sorted, int bcap, int4* __restrict__ info,
            int32_t* __restrict__ parent, double4* __restrict__ cm,
            double* __restrict__ sums, double* __restrict__ box,
            float* __restrict__ quad) {
    constexpr int kS = kTreeTile + kTreeHalo;
    __shared__ uint64_t s_mask[kMaxLevels + 1][kTreeWaves];
    __shared__ uint32_t s_woff[kMaxLevels + 1][kTreeWaves];
    __shared__ int s_ls[kMaxLevels + 3];
    __shared__ double s_x[kS], s_y[kS], s_z[kS], s_m[kS];
    __shared__ unsigned char s_d[kS];
    __shared__ int s_jk[kTreeTile];   // leaf jobs: node index
    __shared__ int s_jb[kTreeTile];   //            first body - t0, level << 16
    __shared__ int s_jw[kTreeWaves];  // jobs queued by earlier waves
    if (hdr->overflow) return;
    const int n = hdr->n_active;
    const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
    const int t0 = blockIdx.x * kTreeTile;
    const double* sx = sorted;
    const double* sy = sorted + bcap;
    const double* sz = sorted + 2 * (size_t)bcap;
    const double* sm = sorted + 3 * (size_t)bcap;
    for (int q = t; q < kS; q += kTreeTile) {
        int j = t0 + q;
        if (j < n) {
            s_d[q] = dl[j].x;
            s_x[q] = sx[j]; s_y[q] = sy[j]; s_z[q] = sz[j]; s_m[q] = sm[j];
        } else {
            s_d[q] = 0;
        }
    }
    if (t < kMaxLevels + 3) s_ls[t] = hdr->level_start[t];
    const int i = t0 + t;
    int d = 0, lam = 0;
    if (i < n) {
        uchar2 v = dl[i];
        d = v.x;
        lam = v.y;
    }
    for (int l = 1; l <= levels; l++) {
        uint64_t mask = __ballot(d < l && l <= lam);
        if (lane == 0) s_mask[l][w] = mask;
    }
    __syncthreads();
    if (t < levels * kTreeWaves) {
        int l = t / kTreeWaves + 1, ww = t % kTreeWaves;
        uint32_t s = toff[(size_t)l * tstride + blockIdx.x];
        for (int q = 0; q < ww; q++) s += (uint32_t)__popcll(s_mask[l][q]);
        s_woff[l][ww] = s;
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
    auto index_at = [&](int l) -> int {
        return s_ls[l] + (int)s_woff[l][w] + __popcll(s_mask[l][w] & below);
    };
    int leaf_k = -1;
    if (i < n) {
        for (int l = d + 1; l <= lam; l++) {
            int k = index_at(l);
            if (k >= node_cap) break;
            // the level l - 1 node holding body i: opened by i itself when
            // d < l - 1, else by the last body before i that opened one
            parent[k] = l == 1 ? 0 : index_at(l - 1) + (d < l - 1 ? 1 : 0) - 1;
            if (l < lam) info[k] = make_int4(index_at(l + 1), l << 8, i, 0);
            else leaf_k = k;
        }
    }
    {   // queue the leaf jobs in order of their first body
        const uint64_t has = __ballot(leaf_k >= 0);
        if (lane == 0) s_jw[w] = __popcll(has);
        __syncthreads();
        int off = 0;
        for (int q = 0; q < w; q++) off += s_jw[q];
        if (leaf_k >= 0) {
            const int jdx = off + __popcll(has & below);
            s_jk[jdx] = leaf_k;
            s_jb[jdx] = (i - t0) | (lam << 16);
        }
    }
    __syncthreads();
    int njobs = 0;
    for (int q = 0; q < kTreeWaves; q++) njobs += s_jw[q];
    const double big = 1.7976931348623157e308;
    const int lim = t0 + kS;
    for (int job = t; job < njobs; job += kTreeTile) {
        const int k = s_jk[job];
        const int b0 = t0 + (s_jb[job] & 0xffff);
        const int l = s_jb[job] >> 16;
        // leaf: bodies b0 .. end - 1, end = first j > b0 that opens a level-l
        // cell (deepest-level leaves may hold more than cap bodies); the
        // candidates are read 8 at a time so their loads overlap: the staged
        // part first (s_d is 0 past the last body, which stops the search),
        // then, for a deepest-level leaf running past the halo, the rest from
        // global memory
        int end = b0 + 1;
        for (bool found = false; !found && end < lim;) {
            int stop = 8;
#pragma unroll
            for (int u = 7; u >= 0; u--) {
                int q = end + u - t0;
                if (q < kS && (int)s_d[q] < l) stop = u;
            }
            found = stop < 8;
            end = min(end + stop, lim);
        }
        if (end == lim) {
            // past the staged part: eight global loads in flight per round
            for (bool found = false; !found && end < n;) {
                int stop = 8;
#pragma unroll
                for (int u = 7; u >= 0; u--) {
                    const int q = end + u;
                    if (q >= n || (int)dl[q].x < l) stop = u;
                }
                found = stop < 8;
                end += stop;
            }
        }
        if (end > n) end = n;
        // sums in sorted order, four bodies loaded per round
        double M = 0.0, ax = 0.0, ay = 0.0, az = 0.0;
        double lo[3] = {big, big, big}, hi[3] = {-big, -big, -big};
        const int mid = min(end, lim);
        for (int j = b0; j < mid; j += 4) {
            double bx[4], by[4], bz[4], bm[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                int q = min(j + u, lim - 1) - t0;
                bx[u] = s_x[q]; by[u] = s_y[q]; bz[u] = s_z[q]; bm[u] = s_m[q];
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (j + u < mid) leaf_accumulate(bx[u], by[u], bz[u], bm[u], M, ax, ay, az, lo, hi);
        }
        // the rest from global memory, four bodies loaded per round (summed
        // in order, so the sums are the same)
        for (int j = mid; j < end; j += 4) {
            double bx[4], by[4], bz[4], bm[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int q = min(j + u, end - 1);
                bx[u] = sx[q]; by[u] = sy[q]; bz[u] = sz[q]; bm[u] = sm[q];
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (j + u < end) leaf_accumulate(bx[u], by[u], bz[u], bm[u], M, ax, ay, az, lo, hi);
        }
        store_node_sums(k, M, ax, ay, az, lo, hi, cm, sums, box);
        info[k] = make_int4(-1, l << 8, b0, end - b0);
        if (quad) {
            // second moments about the centre of mass (learned far-field
            // correction inputs)
            double c0 = 0.0, c1 = 0.0, c2 = 0.0;
            if (M > 0.0) { c0 = ax / M; c1 = ay / M; c2 = az / M; }
            double Q[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            for (int j = b0; j < end; j++) {
                int q = j - t0;
                bool st = j < mid;
                double dx = (st ? s_x[q] : sx[j]) - c0;
                double dy = (st ? s_y[q] : sy[j]) - c1;
                double dz = (st ? s_z[q] : sz[j]) - c2;
                double mb = st ? s_m[q] : sm[j];
                Q[0] += mb * dx * dx; Q[1] += mb * dy * dy; Q[2] += mb * dz * dz;
                Q[3] += mb * dx * dy; Q[4] += mb * dx * dz; Q[5] += mb * dy * dz;
            }
#pragma unroll
            for (int r = 0; r < 6; r++) quad[6 * (size_t)k + r] = (float)Q[r];
        }
    }
}

// Internal nodes of level L (deepest level first): child count, body count,
// mass, sum m x, centre of mass = sum m x / M and body box over the children
// in order; a root that is a leaf sums its bodies.
__global__ void __launch_bounds__(256)
k_tree_internal(int L, const TreeHeader* __restrict__ hdr, int4* __restrict__ info,
                const int32_t* __restrict__ parent, const double* __restrict__ sorted,
                int bcap, double4* __restrict__ cm, double* __restrict__ sums,
                double* __restrict__ box, float* __restrict__ quad) {
    if (hdr->overflow) return;
    const int ls = hdr->level_start[L];
    const int le = hdr->level_start[L + 1];
    const int ce = hdr->level_start[L + 2];
    const double big = 1.7976931348623157e308;
    // eight lanes per node, lane u loads child u; the sums are then folded in
    // child order by every lane of the segment
    const int u = threadIdx.x & 7;
    const int seg = (threadIdx.x & (kWave - 1)) & ~7;
    const int per_block = 256 / 8;
    for (int kb = ls + blockIdx.x * per_block; kb < le; kb += gridDim.x * per_block) {
        const int k = kb + (threadIdx.x >> 3);
        const bool valid = k < le;
        int4 nd = valid ? info[k] : make_int4(-1, 0, 0, 0);
        double M = 0.0, ax = 0.0, ay = 0.0, az = 0.0;
        double lo[3] = {big, big, big}, hi[3] = {-big, -big, -big};
        if (valid && nd.x < 0 && L == 0 && u == 0) {
            // a root that is a leaf (at most cap bodies)
            for (int j = 0; j < nd.w; j++)
                leaf_accumulate(sorted[j], sorted[bcap + j], sorted[2 * (size_t)bcap + j],
                                sorted[3 * (size_t)bcap + j], M, ax, ay, az, lo, hi);
            store_node_sums(k, M, ax, ay, az, lo, hi, cm, sums, box);
            if (quad) {
                double c0 = 0.0, c1 = 0.0, c2 = 0.0;
                if (M > 0.0) { c0 = ax / M; c1 = ay / M; c2 = az / M; }
                double Q[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
                for (int j = 0; j < nd.w; j++) {
                    double mb = sorted[3 * (size_t)bcap + j];
                    double dx = sorted[j] - c0, dy = sorted[bcap + j] - c1;
                    double dz = sorted[2 * (size_t)bcap + j] - c2;
                    Q[0] += mb * dx * dx; Q[1] += mb * dy * dy; Q[2] += mb * dz * dz;
                    Q[3] += mb * dx * dy; Q[4] += mb * dx * dz; Q[5] += mb * dy * dz;
                }
                for (int r = 0; r < 6; r++) quad[6 * (size_t)k + r] = (float)Q[r];
            }
            M = 0.0; ax = 0.0; ay = 0.0; az = 0.0;
        }
        const int c = nd.x + u;
        const bool mine = valid && nd.x >= 0 && c < ce && parent[c] == k;
        double cmw = 0.0, s0 = 0.0, s1 = 0.0, s2 = 0.0;
        double b[6] = {big, big, big, -big, -big, -big};
        int cnt = 0;
        if (mine) {
            cmw = cm[c].w;
            s0 = sums[3 * (size_t)c];
            s1 = sums[3 * (size_t)c + 1];
            s2 = sums[3 * (size_t)c + 2];
            cnt = info[c].w;
            const double* cb = box + 6 * (size_t)c;
#pragma unroll
            for (int q = 0; q < 6; q++) b[q] = cb[q];
        }
        const uint64_t own = __ballot(mine);
        const int nch = __popcll((own >> seg) & 0xffull);
        int count = 0;
        for (int q = 0; q < 8; q++) {
            int src = seg + q;
            double vm = __shfl(cmw, src, kWave);
            double v0 = __shfl(s0, src, kWave);
            double v1 = __shfl(s1, src, kWave);
            double v2 = __shfl(s2, src, kWave);
            int vc = __shfl(cnt, src, kWave);
            double vb[6];
#pragma unroll
            for (int r = 0; r < 6; r++) vb[r] = __shfl(b[r], src, kWave);
            if (q < nch) {
                M += vm; ax += v0; ay += v1; az += v2; count += vc;
                lo[0] = fmin(vb[0], lo[0]); lo[1] = fmin(vb[1], lo[1]); lo[2] = fmin(vb[2], lo[2]);
                hi[0] = fmax(vb[3], hi[0]); hi[1] = fmax(vb[4], hi[1]); hi[2] = fmax(vb[5], hi[2]);
            }
        }
        if (quad) {
            // parallel axis: Q = sum over children of Q_c + m_c (c_c - c)(c_c - c)^T
            double c0 = 0.0, c1 = 0.0, c2 = 0.0;
            if (M > 0.0) { c0 = ax / M; c1 = ay / M; c2 = az / M; }
            double Q[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            if (mine && cmw > 0.0) {
                double dx = s0 / cmw - c0, dy = s1 / cmw - c1, dz = s2 / cmw - c2;
                const float* qc = quad + 6 * (size_t)c;
                Q[0] = (double)qc[0] + cmw * dx * dx; Q[1] = (double)qc[1] + cmw * dy * dy;
                Q[2] = (double)qc[2] + cmw * dz * dz; Q[3] = (double)qc[3] + cmw * dx * dy;
                Q[4] = (double)qc[4] + cmw * dx * dz; Q[5] = (double)qc[5] + cmw * dy * dz;
            }
#pragma unroll
            for (int r = 0; r < 6; r++) {
                Q[r] += __shfl_xor(Q[r], 1, kWave);
                Q[r] += __shfl_xor(Q[r], 2, kWave);
                Q[r] += __shfl_xor(Q[r], 4, kWave);
            }
            if (valid && nd.x >= 0 && u == 0)
                for (int r = 0; r < 6; r++) quad[6 * (size_t)k + r] = (float)Q[r];
        }
        if (valid && nd.x >= 0 && u == 0) {
            store_node_sums(k, M, ax, ay, az, lo, hi, cm, sums, box);
            info[k] = make_int4(nd.x, nch | (L << 8), nd.z, count);
        }
    }
}

// Stable LSD radix sort of n (key, value) pairs over 8-bit digits
// first_pass .. first_pass + passes - 1.  On return *k/*v hold the sorted
// pairs (the buffers are swapped after every pass); hist holds
// 256 * (ceil(n / kSortTile) + 1) words.
// first_counted: the caller has already written the first pass's
// histogram columns.
template <typename K>
static void sort_pairs(hipStream_t st, int n, int first_pass, int passes, K** k,
                       int32_t** v, K** kt, int32_t** vt, uint32_t* hist,
                       bool first_counted = false) {
    if (n <= 0) return;
    int ntiles = (n + kSortTile - 1) / kSortTile;
    uint32_t* totals = hist + (size_t)256 * ntiles;
    for (int p = first_pass; p < first_pass + passes; p++) {
        int shift = 8 * p;
        if (p > first_pass || !first_counted)
            hipLaunchKernelGGL(k_sort_hist<K>, dim3(ntiles), dim3(kSortThreads), 0, st, n,
                               shift, ntiles, *k, hist);
        hipLaunchKernelGGL(k_sort_scan, dim3(256), dim3(kSortThreads), 0, st, ntiles,
                           hist, totals);
        hipLaunchKernelGGL(k_sort_scatter<K>, dim3(ntiles), dim3(kSortThreads), 0, st,
                           n, shift, ntiles, *k, *v, hist, totals, *kt, *vt);
        std::swap(*k, *kt);
        std::swap(*v, *vt);
    }
}

void radix_sort_pairs(hipStream_t st, int n, int first_pass, int passes,
                      uint64_t** k, int32_t** v, uint64_t** kt, int32_t** vt,
                      uint32_t* hist) {
    sort_pairs<uint64_t>(st, n, first_pass, passes, k, v, kt, vt, hist);
}

void radix_sort_pairs32(hipStream_t st, int n, int first_pass, int passes,
                        uint32_t** k, int32_t** v, uint32_t** kt, int32_t** vt,
                        uint32_t* hist) {
    sort_pairs<uint32_t>(st, n, first_pass, passes, k, v, kt, vt, hist);
}

// Keys, sort and node counts of the build for K-wide keys.
template <typename K>
static int build_keys_and_counts(Engine& e, int n, int levels, const int32_t** order,
                                 const K** keys) {
    hipStream_t st = e.stream;
    const int stiles = (n + kSortTile - 1) / kSortTile;
    K* ka = reinterpret_cast<K*>(e.keys_a);
    K* kb = reinterpret_cast<K*>(e.keys_b);
    hipLaunchKernelGGL(k_morton<K>, dim3(stiles), dim3(kSortThreads), 0, st, n, levels,
                       stiles, e.hdr, e.f[PX], e.f[PY], e.f[PZ], e.f[MASS], e.act, ka,
                       e.vals_a, e.pm4, e.gmark, e.gtile, e.hist);
    if (e.detail_timing()) CM_HIP_TRY(e, hipEventRecord(e.ev[1], st));
    const int passes = (3 * levels + 1 + 7) / 8;
    int32_t* va = e.vals_a; int32_t* vb = e.vals_b;
    sort_pairs<K>(st, n, 0, passes, &ka, &va, &kb, &vb, e.hist, true);
    if (e.detail_timing()) CM_HIP_TRY(e, hipEventRecord(e.ev[2], st));
    // source radii enter only the shell theorem's interior branch
    const int ntiles = std::max(1, (n + kTreeTile - 1) / kTreeTile);
    hipLaunchKernelGGL(k_gather_count<K>, dim3(ntiles), dim3(kTreeTile), 0, st, n, levels,
                       e.cfg.leaf_capacity, tree_tiles(e.capacity), e.hdr, ka, va, e.pm4,
                       e.cfg.shell_theorem ? e.f[RAD] : nullptr, e.sorted, e.capacity,
                       e.tree_dl, e.tree_tcount);
    *order = va;
    *keys = ka;
    return CM_OK;
}

// Builds the tree for the current state; *keys_out / *order_out are the sorted
// keys (32-bit words when tree_keys32(e), else 64-bit) and body order.
int build_tree(Engine& e, const void** keys_out, const int32_t** order_out) {
    const int n = e.count;
    const int levels = e.cfg.morton_levels;
    hipStream_t st = e.stream;
    int nb = std::min(kBBoxBlocks, std::max(1, (n + 255) / 256));
    hipLaunchKernelGGL(k_bbox, dim3(nb), dim3(256), 0, st, n, e.f[PX], e.f[PY],
                       e.f[PZ], e.act, e.bbox_part, e.bbox_cnt);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(1024), 0, st, nb, e.bbox_part,
                       e.bbox_cnt, e.hdr, e.nd_info);
    const int32_t* va = nullptr;
    int rc;
    if (tree_keys32(e)) {
        const uint32_t* k32 = nullptr;
        rc = build_keys_and_counts<uint32_t>(e, n, levels, &va, &k32);
        *keys_out = k32;
    } else {
        const uint64_t* k64 = nullptr;
        rc = build_keys_and_counts<uint64_t>(e, n, levels, &va, &k64);
        *keys_out = k64;
    }
    if (rc != CM_OK) return rc;
    const int cap = e.cfg.leaf_capacity;
    const int ntiles = std::max(1, (n + kTreeTile - 1) / kTreeTile);
    const int tstride = tree_tiles(e.capacity);
    hipLaunchKernelGGL(k_tree_scan, dim3(1), dim3(1024), 0, st, levels, cap, ntiles,
                       tstride, e.node_cap, e.hdr, e.tree_tcount, e.nd_info);
    hipLaunchKernelGGL(k_tree_emit, dim3(ntiles), dim3(kTreeTile), 0, st, levels,
                       e.node_cap, tstride, e.hdr, e.tree_dl, e.tree_tcount, e.sorted,
                       e.capacity, e.nd_info, e.nd_parent, e.nd_cm, e.nd_sum, e.nd_box,
                       e.cfg.learned_correction ? e.nd_quad : nullptr);
    // deepest internal nodes live at level levels - 1 (one launch per level: a
    // single block for the top four levels measured 25 us slower per build)
    for (int L = levels - 1; L >= 0; L--) {
        double maxsz = std::min((double)n, L < 20 ? std::ldexp(1.0, 3 * L) : 1e18);
        int g = (int)std::min(2048.0, std::max(1.0, std::ceil(maxsz / 32)));
        hipLaunchKernelGGL(k_tree_internal, dim3(g), dim3(256), 0, st, L, e.hdr,
                           e.nd_info, e.nd_parent, e.sorted, e.capacity, e.nd_cm,
                           e.nd_sum, e.nd_box, e.cfg.learned_correction ? e.nd_quad : nullptr);
    }
    CM_HIP_TRY(e, hipGetLastError());
    *order_out = va;
    return CM_OK;
}

}  // namespace cm

// ---- test-only export (csrc/cm_testing.h) ----------------------------------

extern "C" int32_t cm_test_tree(void* h, int32_t node_cap, uint64_t* keys, int32_t* order,
                                double* box4, int32_t* n_active, int32_t* n_nodes,
                                int32_t* info, double* com_mass, int32_t* level_start) {
    using namespace cm;
    Engine* e = static_cast<Engine*>(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (node_cap < 0) return set_error(*e, CM_ERR_INVALID_ARGUMENT, "negative node_cap");
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    const int n = e->count;
    if (n == 0) {
        if (n_active) *n_active = 0;
        if (n_nodes) *n_nodes = 0;
        return CM_OK;
    }
    const void* k = nullptr;
    const int32_t* o = nullptr;
    int rc = build_tree(*e, &k, &o);
    if (rc != CM_OK) return rc;
    TreeHeader hh;
    CM_HIP_TRY(*e, hipMemcpyAsync(&hh, e->hdr, sizeof(hh), hipMemcpyDeviceToHost, e->stream));
    CM_HIP_TRY(*e, hipStreamSynchronize(e->stream));
    if (hh.overflow) return set_error(*e, CM_ERR_POOL_EXHAUSTED, "octree node pool exhausted");
    if (n_active) *n_active = hh.n_active;
    if (n_nodes) *n_nodes = hh.n_nodes;
    if (box4) for (int q = 0; q < 4; q++) box4[q] = hh.box[q];
    if (level_start) for (int q = 0; q < kMaxLevels + 3; q++) level_start[q] = hh.level_start[q];
    std::vector<uint32_t> k32;
    if (keys && tree_keys32(*e)) {
        k32.resize((size_t)n);
        CM_HIP_TRY(*e, hipMemcpyAsync(k32.data(), k, (size_t)n * 4, hipMemcpyDeviceToHost,
                                      e->stream));
    } else if (keys) {
        CM_HIP_TRY(*e, hipMemcpyAsync(keys, k, (size_t)n * 8, hipMemcpyDeviceToHost, e->stream));
    }
    if (order)
        CM_HIP_TRY(*e, hipMemcpyAsync(order, o, (size_t)n * 4, hipMemcpyDeviceToHost, e->stream));
    const int nn = std::min(node_cap, hh.n_nodes);
    std::vector<int4> inf((size_t)std::max(nn, 1));
    std::vector<double4> cm4((size_t)std::max(nn, 1));
    if (nn > 0) {
        CM_HIP_TRY(*e, hipMemcpyAsync(inf.data(), e->nd_info, (size_t)nn * sizeof(int4),
                                      hipMemcpyDeviceToHost, e->stream));
        CM_HIP_TRY(*e, hipMemcpyAsync(cm4.data(), e->nd_cm, (size_t)nn * sizeof(double4),
                                      hipMemcpyDeviceToHost, e->stream));
    }
    CM_HIP_TRY(*e, hipStreamSynchronize(e->stream));
    for (size_t q = 0; q < k32.size(); q++) keys[q] = k32[q];
    for (int q = 0; q < nn; q++) {
        if (info) {
            int32_t* r = info + 5 * (size_t)q;
            r[0] = inf[q].x;
            r[1] = inf[q].y & 255;
            r[2] = inf[q].y >> 8;
            r[3] = inf[q].z;
            r[4] = inf[q].w;
        }
        if (com_mass) {
            double* r = com_mass + 4 * (size_t)q;
            r[0] = cm4[q].x; r[1] = cm4[q].y; r[2] = cm4[q].z; r[3] = cm4[q].w;
        }
    }
    return CM_OK;
}
