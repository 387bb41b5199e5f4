// k_traverse.hip -- target-group formation and the group-MAC Barnes-Hut
// traversal, one 64-lane wavefront per target group.
//
// Target groups (k_groups) follow the octree: for every node holding more
// than group_size bodies, runs of consecutive children that each hold at most
// group_size bodies are packed, in child order, into groups while a group
// stays within group_size and stays compact (it is not packed further when
// the squared diagonal of its box would exceed 4x the larger part's, or its
// padded volume 4x the parts' summed volume).  A group is a contiguous range
// of Morton-sorted bodies inside one parent cell, so its box never spans a
// Morton jump into a distant, denser region.
//
// Traversal (k_traverse): persistent wavefronts take groups from eight device
// queues, one per XCD, each an eighth of the groups in body order (so each
// XCD's L2 serves one compact region).  A queue runs first the groups that
// did at least twice the mean work in the previous traversal (k_group_order);
// the order decides only which wavefront computes a group and when, never its
// result.  Each group walks the tree
// from the root with an LDS stack,
// 64 nodes per step (fewer when the stack is nearly full, so it never
// overflows): nodes of zero mass are skipped (BarnesHutBackend.cs:402); a node
// is accepted as a monopole when d2 > 0 and (4 hL) hL < theta^2 d2, d2 being
// the squared distance from its centre of mass to the group's box and
// hL = h 2^-level (BarnesHutBackend.cs:431-433 with the body replaced by the
// nearest point of the box); otherwise a leaf contributes its bodies and an
// internal node pushes its children.  Accepted nodes and leaf bodies go to two
// LDS interaction lists (64 entries; 128 in the FP32 split kernel, whose far
// cells go to a third); a full list is gathered and staged once (one entry per
// lane and 64-entry half) and every member lane sums all its entries (FP32
// split: a group of at most 32 bodies spreads each target over several
// lanes, see below).
// FP64 (CM_PREC_FP64): monopole terms use the reference's monopole arithmetic
// (BarnesHutBackend.cs:438-445), body terms GravityKernel's
// (GravityKernel.cs:9-33); each lane keeps the two sums apart and adds them at
// the end: bit-identical to the CPU restatement of this traversal.
// FP64_FAST: the same with a hardware reciprocal square root + Newton steps.
// FP32_SPLIT: offsets from the group's box centre are formed in double and
// rounded to float once per list entry / target; the pair terms run in float
// (v_rsq_f32, FMA) and each flush's float partial is folded into double.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bh_device.h"
#include "cm_internal.h"
#include "cm_testing.h"

namespace cm {

// ------------------------------------------------------------ grouping ----
__device__ __forceinline__ double diag2_of(const double* lo, const double* hi) {
    double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return dx * dx + dy * dy + dz * dz;
}

__device__ __forceinline__ double vol_of(const double* lo, const double* hi, double pad) {
    return ((hi[0] - lo[0]) + pad) * ((hi[1] - lo[1]) + pad) * ((hi[2] - lo[2]) + pad);
}

// A group is recorded at its first body (gmark[s] = e) and counted in the
// tile of kGroupTile bodies that holds it; k_group_emit then lists the groups
// in body (Morton) order, so that consecutive groups are spatial neighbours.
// A thread's groups are in body order, so it counts them per tile and adds
// each tile's count once (the counter is cleared by k_morton).
struct GroupCounter {
    int32_t* gtile;
    int tile = -1, cnt = 0;
    __device__ void flush() {
        if (cnt > 0) atomicAdd(&gtile[tile], cnt);
        cnt = 0;
    }
    __device__ void emit(int32_t* gmark, int s, int e) {
        gmark[s] = e;
        const int t = s / kGroupTile;
        if (t != tile) { flush(); tile = t; }
        cnt++;
    }
};

// box_split with the three boxes given directly: the bin [s, mid) as lo1/hi1,
// the next child [mid, e) as lo2/hi2; the whole box is their union.  Node
// boxes are exact min/max over their bodies, so this equals box_split on the
// bodies.
__device__ bool box_split_nodes(const double* lo1, const double* hi1,
                                const double* lo2, const double* hi2) {
    double lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
        lo[k] = lo1[k] < lo2[k] ? lo1[k] : lo2[k];
        hi[k] = hi1[k] > hi2[k] ? hi1[k] : hi2[k];
    }
    double d = diag2_of(lo, hi), d1 = diag2_of(lo1, hi1), d2 = diag2_of(lo2, hi2);
    double dm = d1 > d2 ? d1 : d2;
    double pad = 1e-3 * ((hi[0] - lo[0]) + (hi[1] - lo[1]) + (hi[2] - lo[2]));
    double v = vol_of(lo, hi, pad), v1 = vol_of(lo1, hi1, pad), v2 = vol_of(lo2, hi2, pad);
    return d > 4.0 * dm || v > 4.0 * (v1 + v2);
}

__global__ void __launch_bounds__(256)
k_groups(TreeHeader* __restrict__ hdr, const int4* __restrict__ info,
         const double* __restrict__ nbox, int gsize, int32_t* __restrict__ gmark,
         int32_t* __restrict__ gtile) {
    // a tree that outgrew the node pool is incomplete: no groups, no
    // traversal (the evaluation reports CM_ERR_POOL_EXHAUSTED)
    if (hdr->overflow) return;
    const int nn = hdr->n_nodes;
    GroupCounter gc{gtile};
    for (int k = blockIdx.x * 256 + threadIdx.x; k < nn; k += gridDim.x * 256) {
        int4 nd = info[k];
        if (k == 0 && nd.w > 0 && nd.w <= gsize) gc.emit(gmark, 0, nd.w);
        if (nd.w <= gsize) continue;
        if (nd.x < 0) {
            // a deepest-level leaf holding more than gsize bodies: consecutive
            // slices of gsize
            for (int q = nd.z; q < nd.z + nd.w; q += gsize)
                gc.emit(gmark, q, min(q + gsize, nd.z + nd.w));
            continue;
        }
        int nch = nd.y & 255;
        int bs = -1, be = -1;
        double blo[3], bhi[3];   // box of the open bin
        for (int c = nd.x; c < nd.x + nch; c++) {
            int4 cn = info[c];
            int cs = cn.z, ce = cn.z + cn.w;
            const double* cb = nbox + 6 * (size_t)c;
            bool close = false, open = false;
            if (cn.w > gsize) {
                close = true;
            } else if (bs < 0) {
                open = true;
            } else if (ce - bs <= gsize && !box_split_nodes(blo, bhi, cb, cb + 3)) {
                be = ce;
                for (int q = 0; q < 3; q++) {
                    blo[q] = blo[q] < cb[q] ? blo[q] : cb[q];
                    bhi[q] = bhi[q] > cb[3 + q] ? bhi[q] : cb[3 + q];
                }
            } else {
                close = true;
                open = true;
            }
            if (close && bs >= 0) {
                gc.emit(gmark, bs, be);
                bs = -1;
            }
            if (open) {
                bs = cs; be = ce;
                for (int q = 0; q < 3; q++) { blo[q] = cb[q]; bhi[q] = cb[3 + q]; }
            }
        }
        if (bs >= 0) gc.emit(gmark, bs, be);
    }
    gc.flush();
}

// Groups in body order: k_groups counts the groups of every tile of
// kGroupTile bodies; up to kGroupPrefixTiles tiles, each block of k_group_emit
// sums the counts of the tiles before its own, else one scan kernel turns the
// counts into offsets first; each tile's groups are then written at its
// offset (ballot ranks).
constexpr int kGroupPrefixTiles = 4096;

__global__ void __launch_bounds__(1024)
k_group_scan(int ntiles, TreeHeader* __restrict__ hdr, int32_t* __restrict__ tcount) {
    __shared__ int s_w[16];
    __shared__ int s_carry;
    const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
    if (t == 0) s_carry = 0;
    __syncthreads();
    for (int base = 0; base < ntiles; base += 1024) {
        const int q = base + t;
        const int v = q < ntiles ? tcount[q] : 0;
        int x = v;
        for (int o = 1; o < kWave; o <<= 1) {
            int y = __shfl_up(x, o, kWave);
            if (lane >= o) x += y;
        }
        if (lane == kWave - 1) s_w[w] = x;
        __syncthreads();
        int woff = 0, tot = 0;
        for (int k = 0; k < 16; k++) {
            if (k < w) woff += s_w[k];
            tot += s_w[k];
        }
        const int carry = s_carry;
        if (q < ntiles) tcount[q] = carry + woff + x - v;
        __syncthreads();
        if (t == 0) s_carry = carry + tot;
        __syncthreads();
    }
    if (t == 0) hdr->n_groups = s_carry;
}

// scanned: tcount holds each tile's offset and hdr->n_groups the total
// (k_group_scan ran); else tcount holds each tile's count.
__global__ void __launch_bounds__(kGroupTile)
k_group_emit(int ntiles, int scanned, TreeHeader* __restrict__ hdr,
             const int32_t* __restrict__ gmark, const int32_t* __restrict__ tcount,
             const int32_t* __restrict__ order, const int32_t* __restrict__ bcost,
             int2* __restrict__ groups, int32_t* __restrict__ gcost,
             long long* __restrict__ twork) {
    constexpr int kW = kGroupTile / kWave;
    __shared__ int s_w[kW];   // marks per wave of this tile
    __shared__ int s_b[kW];   // per wave: counts of the tiles before this one
    __shared__ int s_a[kW];   // per wave: counts of all tiles
    __shared__ long long s_c[kW];   // per wave: work of its groups
    const int n = hdr->n_active;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const int i = blockIdx.x * kGroupTile + threadIdx.x;
    const int e = i < n ? gmark[i] : 0;
    // the group's work in the last traversal (at its first body), and this
    // tile's total (each group at least 1) for k_group_order
    const int c = e > 0 ? bcost[order[i]] : 0;
    long long cw = e > 0 ? max(1, c) : 0;
    int before = 0, all = 0;
    if (!scanned) {
        for (int q = threadIdx.x; q < ntiles; q += kGroupTile) {
            const int c = tcount[q];
            before += q < (int)blockIdx.x ? c : 0;
            all += c;
        }
        for (int o = kWave / 2; o > 0; o >>= 1) {
            before += __shfl_xor(before, o, kWave);
            all += __shfl_xor(all, o, kWave);
        }
    }
    for (int o = kWave / 2; o > 0; o >>= 1) cw += __shfl_xor(cw, o, kWave);
    const uint64_t m = __ballot(e > 0);
    if (lane == 0) {
        s_w[w] = __popcll(m);
        s_b[w] = before;
        s_a[w] = all;
        s_c[w] = cw;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long bw = 0;
        for (int k = 0; k < kW; k++) bw += s_c[k];
        twork[blockIdx.x] = bw;
    }
    int base = 0;
    if (scanned) {
        base = tcount[blockIdx.x];
    } else {
        int tot = 0;
        for (int k = 0; k < kW; k++) { base += s_b[k]; tot += s_a[k]; }
        if (blockIdx.x == 0 && threadIdx.x == 0) hdr->n_groups = tot;
    }
    if (e > 0) {
        int off = base;
        for (int k = 0; k < w; k++) off += s_w[k];
        off += __popcll(m & ((1ull << lane) - 1ull));
        groups[off] = make_int2(i, e);
        gcost[off] = c;
    }
}

// Traversal order.  Eighth q of the groups (in body order) is XCD q's queue.
// Within it the heavy groups run first: those whose work in the previous
// traversal, read at their first body from bcost (kept by the caller's body
// index, so it survives the re-sort), is at least twice the mean over all
// groups; the rest follow in body order.  In a dense region whose deepest
// leaves hold many bodies a group does several times the mean work, and one
// drawn late kept its wavefront running after the others had finished.  Only
// the heavy groups move: a queue ordered entirely by descending work was
// slower, because it ends on a run of light groups whose latency-bound walks
// then have no arithmetic of other wavefronts to hide behind.  The partition
// is stable, so each class keeps body order, and any bcost content (zeros for
// a fresh engine: body order) gives a valid permutation; the order decides
// which wavefront computes a group and when, never its result.
//
// Fat groups (split != 0: the bit-exact FP64 path, whose sums do not depend
// on which lanes or wavefronts form them): a group whose work is at least
// kFatFactor times the mean enters its queue, ahead of the heavy ones, as
// P = min(kMaxParts, bodies, work / (2 mean)) entries, part p holding the
// targets [x + p n / P, x + (p + 1) n / P) of its n bodies.  Each part walks
// the whole group's box, so it finds the group's lists, and computes only its
// targets (with the lanes of the other targets): a group spanning a wide box
// (a sparse region of a thin disk) no longer sets the kernel's length on one
// wavefront.  Entries are (group start, group end, targets start, targets
// end); queue q fills [qbase[q], qend[q]) of the queue array, qbase[q] =
// q ng / 8 + q B with B = (capacity - ng) / 8 spare entries per queue; a queue
// whose parts would need more keeps its fat groups whole (as heavy ones).
constexpr int kHeavyFactor = 2;
constexpr int kFatFactor = 4;
constexpr int kMaxParts = 8;
constexpr int kOrderThreads = 1024;
constexpr int kOrderWaves = kOrderThreads / kWave;
// opened leaves of at most this many bodies are copied into the body list by
// their own lanes; a walk step that opens a larger one expands its leaves 64
// list positions at a time instead
constexpr int kDirectLeaf = 16;

__device__ __forceinline__ int group_parts(long long c, long long mean, int nm, int split) {
    if (!split || mean <= 0 || c < kFatFactor * mean || nm < 2) return 1;
    return (int)min((long long)min(kMaxParts, nm), c / (2 * mean));
}

// one block per eighth; gcost: the work of each group in body order, twork
// the totals per tile of bodies (both written by k_group_emit)
__global__ void __launch_bounds__(kOrderThreads)
k_group_order(int ntiles, int split, int qcap, TreeHeader* __restrict__ hdr,
              const int2* __restrict__ groups, const int32_t* __restrict__ gcost,
              const long long* __restrict__ twork, int4* __restrict__ gq) {
    __shared__ int s_cnt[kOrderWaves][3];   // per wave: part entries, heavy, light; count, then base
    __shared__ int s_fat[kOrderWaves];      // per wave: groups entered as parts
    __shared__ long long s_w[kOrderWaves];
    __shared__ int s_split;
    const int ng = hdr->n_groups;
    const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
    // the mean work over all groups (each counted as at least 1)
    long long sum = 0;
    for (int k = t; k < ntiles; k += kOrderThreads) sum += twork[k];
    for (int o = kWave / 2; o > 0; o >>= 1) sum += __shfl_xor(sum, o, kWave);
    if (lane == 0) s_w[w] = sum;
    __syncthreads();
    long long total = 0;
    for (int k = 0; k < kOrderWaves; k++) total += s_w[k];
    const long long mean = ng > 0 ? total / ng : 0;
    const long long heavy = kHeavyFactor * mean;
    // wave w partitions a contiguous slice of the eighth
    const int lo = (int)((long long)blockIdx.x * ng / kXcds);
    const int len = (int)((long long)(blockIdx.x + 1) * ng / kXcds) - lo;
    const int spare = split ? max(0, (qcap - ng) / kXcds) : 0;
    const int qb = lo + (int)blockIdx.x * spare;
    const int ws = lo + (int)((long long)len * w / kOrderWaves);
    const int we = lo + (int)((long long)len * (w + 1) / kOrderWaves);
    int np = 0, nf = 0, nh = 0;
    for (int i = ws + lane; i < we; i += kWave) {
        const int2 gv = groups[i];
        const int P = group_parts(gcost[i], mean, gv.y - gv.x, split);
        if (P > 1) { np += P; nf++; }
        else nh += gcost[i] >= heavy ? 1 : 0;
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        np += __shfl_xor(np, o, kWave);
        nf += __shfl_xor(nf, o, kWave);
        nh += __shfl_xor(nh, o, kWave);
    }
    if (lane == 0) {
        s_cnt[w][0] = np; s_cnt[w][1] = nh; s_cnt[w][2] = (we - ws) - nh - nf;
        s_fat[w] = nf;
    }
    __syncthreads();
    if (t == 0) {
        int extra = 0;
        for (int u = 0; u < kOrderWaves; u++) extra += s_cnt[u][0] - s_fat[u];
        const int ok = extra <= spare;
        if (!ok)   // not enough room: the fat groups stay whole, as heavy ones
            for (int u = 0; u < kOrderWaves; u++) {
                s_cnt[u][1] += s_fat[u];
                s_cnt[u][0] = 0;
            }
        s_split = ok;
        // bases: the parts of every slice, then its heavy groups, then the
        // light ones
        int run = qb;
        for (int k = 0; k < 3; k++)
            for (int u = 0; u < kOrderWaves; u++) {
                const int c = s_cnt[u][k];
                s_cnt[u][k] = run;
                run += c;
            }
        hdr->qbase[blockIdx.x] = qb;
        hdr->qend[blockIdx.x] = run;
    }
    __syncthreads();
    const int sp = s_split ? split : 0;
    int pb = s_cnt[w][0], hb = s_cnt[w][1], lb = s_cnt[w][2];
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    for (int b = ws; b < we; b += kWave) {
        const int i = b + lane;
        const bool ok = i < we;
        int2 gv = make_int2(0, 0);
        int P = 1;
        if (ok) {
            gv = groups[i];
            P = group_parts(gcost[i], mean, gv.y - gv.x, sp);
        }
        const bool fat = ok && P > 1;
        const bool hv = ok && !fat && gcost[i] >= heavy;
        const uint64_t mh = __ballot(hv), ml = __ballot(ok && !fat && !hv);
        // parts: an exclusive prefix sum of P over the fat lanes
        int pin = fat ? P : 0;
        for (int o = 1; o < kWave; o <<= 1) {
            const int y = __shfl_up(pin, o, kWave);
            if (lane >= o) pin += y;
        }
        const int ptot = __shfl(pin, kWave - 1, kWave);
        if (fat) {
            const int nm = gv.y - gv.x;
            const int at = pb + pin - P;
            for (int p = 0; p < P; p++)
                gq[at + p] = make_int4(gv.x, gv.y, gv.x + (int)((long long)p * nm / P),
                                       gv.x + (int)((long long)(p + 1) * nm / P));
        } else if (ok) {
            gq[hv ? hb + __popcll(mh & lt) : lb + __popcll(ml & lt)] =
                make_int4(gv.x, gv.y, gv.x, gv.y);
        }
        pb += ptot;
        hb += __popcll(mh);
        lb += __popcll(ml);
    }
}

// ----------------------------------------------------------- traversal ----
struct TravArgs {
    TreeHeader* hdr;
    const int4* info;
    const double4* cm;
    const int32_t* order;
    const int4* queue;       // queue entries in traversal order (k_group_order)
    int32_t* bcost;          // per body (caller's index): its group's work units
    const double* sorted;    // x, y, z, m, r blocks of `cap`
    int cap;
    int n;                   // bodies (active first, then inactive, in sorted order)
    double* ax;
    double* ay;
    double* az;
    double eps2;
    double theta2;
    int levels;
    int shell;
    unsigned long long* counters;   // [0] cell, [1] body, [2] accumulated terms,
                                    // [3] far cells, [4] accumulated far cells
    double far_ratio2;              // kappa^2 of the far-field expansion (0: off)
    double* rows;                   // potential mode: 0.5 m_i phi_i per body (original index)
    // learned far-field correction (k_traverse<FP32_SPLIT, true> only)
    const uint16_t* mlp_w;          // bf16 weights [out][in] per layer
    const float* mlp_p;             // FP32 parameters (biases)
    const float* quad;              // 6 second moments per node
};

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ double wmin(double v) {
    for (int o = 32; o > 0; o >>= 1) {
        double t = __shfl_xor(v, o, kWave);
        v = t < v ? t : v;
    }
    return v;
}
__device__ __forceinline__ double wmax(double v) {
    for (int o = 32; o > 0; o >>= 1) {
        double t = __shfl_xor(v, o, kWave);
        v = t > v ? t : v;
    }
    return v;
}
// Inclusive prefix sum over the 64 lanes of a wavefront (all lanes active):
// Hillis-Steele within each row of 16 lanes by DPP row shifts of 1, 2, 4 and 8
// (lanes shifted in from outside the row contribute 0), then the row totals
// carried forward by the row broadcasts of lanes 15 and 31.  Integer adds, so
// the result is exact; it replaces six LDS permutes.
__device__ __forceinline__ int wscan_incl(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);   // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);   // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);   // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);   // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);   // row_bcast:15
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);   // row_bcast:31
    return v;
}

// Inclusive running maximum over the 64 lanes (values >= -1; all lanes
// active), by the same DPP pattern as wscan_incl with -1 shifted in.
__device__ __forceinline__ int wmax_scan(int v) {
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x111, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x112, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x114, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x118, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x142, 0xa, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x143, 0xc, 0xf, false));
    return v;
}

__device__ __forceinline__ double gk_coeff(double dist_sq, double m, double r,
                                           double eps2, bool shell) {
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

// Hardware reciprocal square root of a double refined by two Newton steps.
__device__ __forceinline__ double rsqrt_fast(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double hx = 0.5 * x;
    y = y * __builtin_fma(-hx * y, y, 1.5);
    y = y * __builtin_fma(-hx * y, y, 1.5);
    return y;
}

__device__ __forceinline__ float gk_coeff_f(float dist_sq, float m, float r,
                                            float eps2, bool shell) {
    if (shell && r > 0.0f) {
        float dist = __builtin_sqrtf(dist_sq);
        if (dist < r) {
            float denom = __builtin_fmaf(r, r, eps2);
            float inv = __builtin_amdgcn_rsqf(denom);
            return m * inv / denom;
        }
    }
    float inv = __builtin_amdgcn_rsqf(dist_sq + eps2);
    return m * inv * inv * inv;
}

template <int PREC, int N>
struct Staging { double4 v[N]; double r[N]; int id[N]; };
template <int N>
struct Staging<CM_PREC_FP32_SPLIT, N> { float4 v[N]; float r[N]; int id[N]; };

// Occupancy of the traversal: the FP32 split kernel is held to the register
// budget of four wavefronts per SIMD (16 per CU, as many as its LDS allows;
// five or six per SIMD with a smaller stack measured slower); the others
// take what they need.  The FP64 paths keep the stack the CPU restatement of
// this traversal models (kStackCap), which fixes the order of their sums.
// The FP32 split kernel, whose sums are judged by tolerance, takes a stack of
// 1,024 entries (the CPU model of the walk at N = 2^20: 0.7 % more walk
// steps than with 1,536, deepest stack 856) and spends the LDS saved on
// interaction lists of 128 entries (bodies, accepted cells and far cells):
// each gather of list entries from global memory then feeds 128 entries
// instead of 64, halving those round trips.
template <int PREC, bool LEARNED>
struct TravShape {
    static constexpr bool split = PREC == CM_PREC_FP32_SPLIT && !LEARNED;
    static constexpr int stack = split ? 1024 : kStackCap;
    static constexpr int waves = split ? 4 : 1;   // minimum per SIMD
    static constexpr int list = split ? 2 * kWave : kWave;   // interaction-list capacity
};

// POT (FP64 only): the same walk sums the potential phi_i = -sum m / sqrt(d^2
// + eps^2) over the cells and bodies of each body's lists instead of its
// acceleration, as compute_potential_bh_kernel does over its own tree
// (engine/src/cuda/potential_energy_kernel.cu:29-111), and writes
// 0.5 m_i phi_i, body i's share of the potential energy.
template <int PREC, bool LEARNED, bool POT = false>
__global__ void __launch_bounds__(kWave, (TravShape<PREC, LEARNED>::waves))
k_traverse(TravArgs a) {
    static_assert(!POT || (PREC == CM_PREC_FP64 && !LEARNED), "potential mode is FP64");
    constexpr int kCap = TravShape<PREC, LEARNED>::stack;
    constexpr int kPL = TravShape<PREC, LEARNED>::list;
    __shared__ int s_stack[kCap];
    __shared__ int s_cell[kPL + kWave];   // a walk step adds at most 64
    constexpr bool kFar = PREC == CM_PREC_FP32_SPLIT && !LEARNED;
    __shared__ int s_far[kFar ? kPL + kWave : 1];
    __shared__ int s_part[kPL];
    __shared__ int s_lpref[kWave];
    __shared__ int s_lstart[kWave];
    __shared__ int s_mark[kWave];
    __shared__ int s_g;
    __shared__ Staging<PREC, kPL> st;
    // MLP inputs, bf16 [cell][16]; read as 8-byte operands
    __shared__ __align__(16) uint16_t s_mx[LEARNED ? kWave * kMlpIn : 4];
    __shared__ float s_my[LEARNED ? kWave * kMlpUsed : 1];   // MLP outputs per cell

    const int lane = lane_id();
    const double h = a.hdr->box[3];
    const int cap = a.cap;
    const double* sx = a.sorted;
    const double* sy = a.sorted + cap;
    const double* sz = a.sorted + 2 * (size_t)cap;
    const double* sm = a.sorted + 3 * (size_t)cap;
    const double* sr = a.sorted + 4 * (size_t)cap;
    const bool shell = a.shell != 0;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int reserve = 8 * (a.levels + 2);
    unsigned long long n_cell = 0ull, n_body = 0ull, n_far = 0ull;

    // eighth q of the groups (in body order; heavy ones first) is queue q,
    // served first by the wavefronts of XCD q so that each XCD's L2 holds one
    // compact region of the tree; a wavefront whose queue is empty helps the
    // others.  A queue's bounds are read once per queue a wavefront serves.
    const int xcd = blockIdx.x % kXcds;
    int qi = 0;
    int qlo = a.hdr->qbase[xcd], qhi = a.hdr->qend[xcd];
    for (;;) {
        const int q = (xcd + qi) % kXcds;
        if (lane == 0) s_g = qlo + (int)atomicAdd(&a.hdr->work_next[q], 1u);
        __syncthreads();
        // group index, counts and list lengths are wave-uniform: keeping them
        // in scalar registers keeps the inner loops' control and LDS
        // addresses off the vector ALU
        const int g = __builtin_amdgcn_readfirstlane(s_g);
        __syncthreads();
        if (g >= qhi) {
            if (++qi == kXcds) break;
            const int qn = (xcd + qi) % kXcds;
            qlo = a.hdr->qbase[qn];
            qhi = a.hdr->qend[qn];
            continue;
        }
        // the group's bodies [gs, ge) set the box (and so the lists); the
        // entry computes the targets [s, e) among them (all of them unless
        // k_group_order split the group)
        const int4 gr = a.queue[g];
        const int gs = gr.x, ge = gr.y, s = gr.z, e = gr.w;
        const bool whole = s == gs && e == ge;
        const int nmem = e - s;
        const int my = s + lane;
        const bool member = lane < nmem;
        const bool inbox = lane < ge - gs;
        double bxv = 0.0, byv = 0.0, bzv = 0.0;
        if (inbox) { bxv = sx[gs + lane]; byv = sy[gs + lane]; bzv = sz[gs + lane]; }
        double xi = 0.0, yi = 0.0, zi = 0.0;
        if (whole) { xi = bxv; yi = byv; zi = bzv; }
        else if (member) { xi = sx[my]; yi = sy[my]; zi = sz[my]; }

        const double big = 1.7976931348623157e308;
        double lo0 = wmin(inbox ? bxv : big), lo1 = wmin(inbox ? byv : big),
               lo2 = wmin(inbox ? bzv : big);
        double hi0 = wmax(inbox ? bxv : -big), hi1 = wmax(inbox ? byv : -big),
               hi2 = wmax(inbox ? bzv : -big);
        const double cxg = (lo0 + hi0) * 0.5;
        const double cyg = (lo1 + hi1) * 0.5;
        const double czg = (lo2 + hi2) * 0.5;
        const float txf = (float)(xi - cxg), tyf = (float)(yi - cyg), tzf = (float)(zi - czg);
        // far field (FP32 split): an accepted cell whose centre of mass r
        // lies beyond R / kappa of the box centre c (R the box half
        // diagonal) is summed for the whole group as the second-order Taylor
        // expansion about c of its softened monopole field
        //   a(c + d) = T0 + T1 d + (1/2) T2 d d,
        //   T0 = M r s^-3/2,  T1_ab = 3 M r_a r_b s^-5/2 - d_ab M s^-3/2,
        //   T2_abc = 15 M r_a r_b r_c s^-7/2
        //            - 3 M (d_ab r_c + d_ac r_b + d_bc r_a) s^-5/2,
        // s = |r|^2 + eps^2.  Each lane sums the coefficients of its share of
        // the far cells; the wave adds them up once per group.
        double far_lim = 1.7976931348623157e308;
        if constexpr (kFar) {
            if (a.far_ratio2 > 0.0) {
                const double ex = hi0 - lo0, ey = hi1 - lo1, ez = hi2 - lo2;
                far_lim = 0.25 * (ex * ex + ey * ey + ez * ez) / a.far_ratio2;
            }
        }
        float f0x = 0.0f, f0y = 0.0f, f0z = 0.0f;                       // T0
        float f1xx = 0.0f, f1yy = 0.0f, f1zz = 0.0f, f1xy = 0.0f, f1xz = 0.0f, f1yz = 0.0f;
        float f2xxx = 0.0f, f2yyy = 0.0f, f2zzz = 0.0f, f2xxy = 0.0f, f2xxz = 0.0f;
        float f2yyx = 0.0f, f2yyz = 0.0f, f2zzx = 0.0f, f2zzy = 0.0f, f2xyz = 0.0f;
        int any_far = 0;
        // FP32 split: a small group gives each target kl lanes (a power of two,
        // kl * nmem <= 64); lane l sums entries sub, sub + kl, ... of every
        // list for target l % nmem, and the kl partial sums are added in sub
        // order at the end.  The staged list is stored so that each sub's
        // entries are contiguous: entry j sits in slot (j % kl) * W + j / kl,
        // W = C / kl for a list of C slots (128; 64 in the FP64 kernels),
        // so lane sub reads slots sub * W + t, t = 0, 1, ...
        // The FP64 baseline (FP64_FAST) splits small groups the same way.
        // The bit-exact FP64 path (kOrdered) spreads a target's entries over
        // kl lanes too, but only their terms: lane (t, q) forms the term of
        // entry k kl + q in trip k, and target t's own lane subtracts the kl
        // terms of the trip one after another, in entry order, so every sum
        // takes the same operations in the same order as one lane alone.
        constexpr bool kSplit = (PREC == CM_PREC_FP32_SPLIT && !LEARNED) ||
                                (PREC == CM_PREC_FP64_FAST && !POT);
        constexpr bool kOrdered = PREC == CM_PREC_FP64 && !POT;
        constexpr bool kLanes = kSplit || kOrdered;
        int kl = 1, klog = 0;
        if constexpr (kLanes) {
            while (2 * kl * nmem <= kWave) { kl *= 2; klog++; }
        }
        // the group's work in units of one list entry per lane (13 VALU
        // instructions): entries / kl per flush, 14 per walk step, 5 per 64
        // far cells; recorded for the next traversal's order
        int wcost = 0;
        const int tl = lane % nmem, sub = lane / nmem;
        const bool worker = lane < kl * nmem;
        float txw = txf, tyw = tyf, tzw = tzf;
        double xw = xi, yw = yi, zw = zi;   // FP64 paths: the lane's target
        if constexpr (kLanes) {
            if (worker && sub > 0) {
                const int mt = s + tl;
                if constexpr (PREC == CM_PREC_FP32_SPLIT) {
                    txw = (float)(sx[mt] - cxg);
                    tyw = (float)(sy[mt] - cyg);
                    tzw = (float)(sz[mt] - czg);
                } else {
                    xw = sx[mt]; yw = sy[mt]; zw = sz[mt];
                }
            }
        }
        double acx = 0.0, acy = 0.0, acz = 0.0;   // monopole sums
        double apx = 0.0, apy = 0.0, apz = 0.0;   // body sums (FP32: folded partials)
        // learned correction summed over this lane's cells (lane = list
        // entry): residual at the group centre and its gradient (xx, yy, zz,
        // xy, xz, yz), in the original axes
        float lres[3] = {0.0f, 0.0f, 0.0f};
        float lgrad[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};

        // MLP over the staged cells (lane j = cell j < cnt), one 16-cell tile
        // at a time; hidden activations stay in registers: the transposed
        // layer Z^T = W X^T leaves lane l with Z[cell l % 16][4 (l / 16) + r],
        // which is exactly the B operand the next layer needs
        auto learned_flush = [&](int cnt) {
          if constexpr (LEARNED && PREC == CM_PREC_FP32_SPLIT) {
            const int lr = lane & 15, lg = lane >> 4;
            float sres = 0.0f, sgrad = 0.0f;
            float x[kMlpIn], f[9];
            for (int k = 0; k < kMlpIn; k++) x[k] = 0.0f;
            for (int k = 0; k < 9; k++) f[k] = 0.0f;
            if (lane < cnt) {
                const int node = s_cell[lane];
                const int4 ni = a.info[node];
                if (ni.w >= 2) {   // one-body cells have no residual
                    const float4 v = st.v[lane];
                    const float hl = (float)ldexp(h, -(ni.y >> 8));
                    mlp_inputs(v.x, v.y, v.z, v.w, hl, a.quad + 6 * (size_t)node, x, sres,
                               sgrad, f);
                }
            }
            for (int k = 0; k < kMlpIn; k++) s_mx[lane * kMlpIn + k] = f2bf(x[k]);
            __syncthreads();
            const uint16_t* W = a.mlp_w;
            const float* P = a.mlp_p;
            for (int tile = 0; tile < 4; tile++) {
                if (tile * 16 >= cnt) break;
                v4s hin[4], hout[4];
                // layer 1 (K = 16)
                {
                    v4s b = *(const v4s*)&s_mx[(tile * 16 + lr) * kMlpIn + 4 * lg];
                    for (int ot = 0; ot < 4; ot++) {
                        v4s w = *(const v4s*)&W[kBW1 + (ot * 16 + lr) * kMlpIn + 4 * lg];
                        v4f z = mfma16(w, b, v4f{0.0f, 0.0f, 0.0f, 0.0f});
                        for (int r = 0; r < 4; r++) {
                            float t = z[r] + P[kPB1 + ot * 16 + 4 * lg + r];
                            hin[ot][r] = (short)f2bf(t > 0.0f ? t : 0.0f);
                        }
                    }
                }
                // layers 2 and 3 (K = 64)
                for (int layer = 0; layer < 2; layer++) {
                    const int wo = layer == 0 ? kBW2 : kBW3;
                    const int bo = layer == 0 ? kPB2 : kPB3;
                    for (int ot = 0; ot < 4; ot++) {
                        v4f z = {0.0f, 0.0f, 0.0f, 0.0f};
                        for (int kt = 0; kt < 4; kt++) {
                            v4s w = *(const v4s*)&W[wo + (ot * 16 + lr) * kMlpHid + kt * 16 + 4 * lg];
                            z = mfma16(w, hin[kt], z);
                        }
                        for (int r = 0; r < 4; r++) {
                            float t = z[r] + P[bo + ot * 16 + 4 * lg + r];
                            hout[ot][r] = (short)f2bf(t > 0.0f ? t : 0.0f);
                        }
                    }
                    for (int ot = 0; ot < 4; ot++) hin[ot] = hout[ot];
                }
                // output layer (16 outputs, 8 used)
                v4f y = {0.0f, 0.0f, 0.0f, 0.0f};
                for (int kt = 0; kt < 4; kt++) {
                    v4s w = *(const v4s*)&W[kBW4 + lr * kMlpHid + kt * 16 + 4 * lg];
                    y = mfma16(w, hin[kt], y);
                }
                // lane l holds outputs 4 lg + r of cell tile * 16 + lr: hand
                // the 8 used ones to the cell's lane through LDS
                if (lg < 2)
                    for (int r = 0; r < 4; r++)
                        s_my[(tile * 16 + lr) * kMlpUsed + 4 * lg + r] = y[r] + P[kPB4 + 4 * lg + r];
            }
            __syncthreads();
            if (lane < cnt && sres != 0.0f) {
                const float* yo = &s_my[lane * kMlpUsed];
                // res = F^T res' sres; J = F^T J' F sgrad
                const float rp[3] = {yo[0] * sres, yo[1] * sres, yo[2] * sres};
                const float jxx = yo[3], jyy = yo[4], jxy = yo[5], jxz = yo[6], jyz = yo[7];
                const float jp[9] = {jxx, jxy, jxz, jxy, jyy, jyz, jxz, jyz, -jxx - jyy};
                for (int i = 0; i < 3; i++)
                    lres[i] += f[i] * rp[0] + f[3 + i] * rp[1] + f[6 + i] * rp[2];
                float t[9];   // J' F
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++)
                        t[3 * i + j] = jp[3 * i] * f[j] + jp[3 * i + 1] * f[3 + j] + jp[3 * i + 2] * f[6 + j];
                auto jg = [&](int i, int j) {   // (F^T J' F)_ij
                    return (f[i] * t[j] + f[3 + i] * t[3 + j] + f[6 + i] * t[6 + j]) * sgrad;
                };
                lgrad[0] += jg(0, 0); lgrad[1] += jg(1, 1); lgrad[2] += jg(2, 2);
                lgrad[3] += jg(0, 1); lgrad[4] += jg(0, 2); lgrad[5] += jg(1, 2);
            }
            __syncthreads();
          }
        };

        auto flush = [&](int cnt, bool cells) {
            wcost += kLanes ? (cnt + kl - 1) >> klog : cnt;
            // the list's kPL slots; W = kPL / kl
            const int wlog = (kPL == kWave ? 6 : 7) - klog;
            const int kw = kPL >> klog;
            const int rbase = worker ? sub * kw : 0;
            // lists of 128 entries: both halves' gathers are issued before
            // either is converted, so one round trip serves them
            double4 g[kPL / kWave];
            float gr[kPL / kWave];
#pragma unroll
            for (int hh = 0; hh < kPL / kWave; hh++) {
                const int j = hh * kWave + lane;
                if (j < cnt) {
                    if (cells) {
                        g[hh] = a.cm[s_cell[j]];
                    } else {
                        const int q = s_part[j];
                        g[hh] = make_double4(sx[q], sy[q], sz[q], sm[q]);
                        gr[hh] = 0.0f;
                        if constexpr (PREC == CM_PREC_FP32_SPLIT) {
                            if (shell) gr[hh] = (float)sr[q];
                        }
                    }
                }
            }
#pragma unroll
            for (int hh = 0; hh < kPL / kWave; hh++) {
                const int j = hh * kWave + lane;
                const int slot = kLanes ? (((j & (kl - 1)) << wlog) | (j >> klog)) : j;
                if constexpr (PREC == CM_PREC_FP32_SPLIT) {
                    // the slots past cnt hold massless far entries, whose terms
                    // are exactly zero, so the loops below may run past cnt
                    if (j >= cnt) st.v[slot] = make_float4(1e18f, 1e18f, 1e18f, 0.0f);
                } else if constexpr (kLanes) {
                    if (j >= cnt) st.v[slot] = make_double4(1e18, 1e18, 1e18, 0.0);
                }
                if (j < cnt) {
                    // a cell's centre of mass or a body: (x, y, z, m)
                    const double4 v = g[hh];
                    if constexpr (PREC == CM_PREC_FP32_SPLIT) {
                        st.v[slot] = make_float4((float)(v.x - cxg), (float)(v.y - cyg),
                                                 (float)(v.z - czg), (float)v.w);
                    } else {
                        st.v[slot] = v;
                    }
                    if (!cells) {
                        const int q = s_part[j];
                        if constexpr (PREC == CM_PREC_FP32_SPLIT) st.r[j] = gr[hh];
                        else st.r[j] = shell ? sr[q] : 0.0;
                        st.id[j] = q;
                    }
                }
            }
            __syncthreads();
            if (cells) {
                n_cell += (unsigned long long)cnt * nmem;
            } else {
#pragma unroll
                for (int hh = 0; hh < kPL / kWave; hh++) {
                    const int j = hh * kWave + lane;
                    uint64_t selfm = __ballot(j < cnt && s_part[j] >= s && s_part[j] < e);
                    n_body += (unsigned long long)min(max(cnt - hh * kWave, 0), kWave) * nmem -
                              (unsigned long long)__popcll(selfm);
                }
            }
            if constexpr (kOrdered) {
              if (kl > 1) {
                // every lane takes part (the terms are handed over by
                // shuffles); lanes past kl * nmem and entries past cnt
                // contribute exact zeros, as does a body's own entry, which
                // the one-lane loop skips: x - (+0) == x for every x the sums
                // can hold (they start at +0, so never hold -0).  Four trips'
                // terms are formed before any is handed over, so the
                // shuffles' latency is paid once per four trips.
                const int trips = (cnt + kl - 1) >> klog;
                const int mt = s + tl;
                for (int t0 = 0; t0 < trips; t0 += 4) {
                    double tx[4], ty[4], tz[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const int t = t0 + u;
                        const int j = t * kl + sub;
                        tx[u] = 0.0; ty[u] = 0.0; tz[u] = 0.0;
                        if (worker && t < trips && j < cnt) {
                            const double4 v = st.v[rbase + t];
                            const double dx = xw - v.x, dy = yw - v.y, dz = zw - v.z;
                            const double dist2 = dx * dx + dy * dy + dz * dz;
                            double f = 0.0;
                            if (cells) {
                                const double soft = dist2 + a.eps2;
                                const double inv = 1.0 / sqrt(soft);
                                const double inv3 = inv * inv * inv;
                                f = v.w * inv3;
                            } else if (st.id[j] != mt) {
                                f = gk_coeff(dist2, v.w, st.r[j], a.eps2, shell);
                            }
                            tx[u] = f * dx; ty[u] = f * dy; tz[u] = f * dz;
                        }
                    }
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        if (t0 + u >= trips) break;
                        for (int q = 0; q < kl; q++) {
                            double vx = tx[u], vy = ty[u], vz = tz[u];
                            if (q > 0) {
                                const int src = lane + q * nmem;
                                const int sl = src < kWave ? src : lane;
                                vx = __shfl(tx[u], sl, kWave);
                                vy = __shfl(ty[u], sl, kWave);
                                vz = __shfl(tz[u], sl, kWave);
                            }
                            if (cells) { acx -= vx; acy -= vy; acz -= vz; }
                            else { apx -= vx; apy -= vy; apz -= vz; }
                        }
                    }
                }
              } else if (member) {
                // one lane per target: the entries in order
                if (cells) {
                    for (int j = 0; j < cnt; j++) {
                        double4 v = st.v[j];
                        double dx = xi - v.x, dy = yi - v.y, dz = zi - v.z;
                        double dist2 = dx * dx + dy * dy + dz * dz;
                        double soft = dist2 + a.eps2;
                        double inv = 1.0 / sqrt(soft);
                        double inv3 = inv * inv * inv;
                        double f = v.w * inv3;
                        acx -= f * dx; acy -= f * dy; acz -= f * dz;
                    }
                } else {
                    for (int j = 0; j < cnt; j++) {
                        if (st.id[j] == my) continue;
                        double4 v = st.v[j];
                        double dx = xi - v.x, dy = yi - v.y, dz = zi - v.z;
                        double dist2 = dx * dx + dy * dy + dz * dz;
                        double f = gk_coeff(dist2, v.w, st.r[j], a.eps2, shell);
                        apx -= f * dx; apy -= f * dy; apz -= f * dz;
                    }
                }
              }
            } else if (kSplit ? worker : member) {
                if constexpr (PREC == CM_PREC_FP32_SPLIT) {
                    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
                    const float e2 = (float)a.eps2;
                    if (cells || (!shell && a.eps2 >= kSelfTermMinEps2)) {
                        // with this softening the self term (dx = 0) is
                        // exactly zero; entries past cnt are massless and far
                        auto term = [&](const float4 v) {
                            float dx = txw - v.x, dy = tyw - v.y, dz = tzw - v.z;
                            float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, e2)));
                            float inv = __builtin_amdgcn_rsqf(r2);
                            float f = v.w * (inv * inv * inv);
                            gx = __builtin_fmaf(-f, dx, gx);
                            gy = __builtin_fmaf(-f, dy, gy);
                            gz = __builtin_fmaf(-f, dz, gz);
                        };
                        // entries sub, sub + kl, ... below cnt: trips of them
                        // for sub = 0, fewer (the rest padding) for the others
                        const int trips = (cnt + kl - 1) >> klog;
                        const float4* sv = &st.v[rbase];
                        if (kw >= 4) {
                            const int t4 = min(kw, (trips + 3) & ~3);
                            for (int t = 0; t < t4; t += 4) {
                                const float4 v0 = sv[t], v1 = sv[t + 1], v2 = sv[t + 2],
                                             v3 = sv[t + 3];
                                term(v0); term(v1); term(v2); term(v3);
                            }
                        } else {
                            for (int t = 0; t < trips; t++) term(sv[t]);
                        }
                    } else {
                        const int mt = s + tl;
                        for (int t = 0, j = sub; j < cnt; t++, j += kl) {
                            if (st.id[j] == mt) continue;
                            float4 v = st.v[rbase + t];
                            float dx = txw - v.x, dy = tyw - v.y, dz = tzw - v.z;
                            float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                            float f = gk_coeff_f(d2, v.w, st.r[j], e2, shell);
                            gx = __builtin_fmaf(-f, dx, gx);
                            gy = __builtin_fmaf(-f, dy, gy);
                            gz = __builtin_fmaf(-f, dz, gz);
                        }
                    }
                    apx += (double)gx; apy += (double)gy; apz += (double)gz;
                } else if (POT) {
                    // potential: phi -= m / sqrt(d^2 + eps^2), cells and
                    // bodies (self skipped) summed apart
                    for (int j = 0; j < cnt; j++) {
                        if (!cells && st.id[j] == my) continue;
                        double4 v = st.v[j];
                        double dx = xi - v.x, dy = yi - v.y, dz = zi - v.z;
                        double dist2 = dx * dx + dy * dy + dz * dz;
                        double dist = sqrt(dist2 + a.eps2);
                        if (cells) acx -= v.w / dist;
                        else apx -= v.w / dist;
                    }
                } else if constexpr (PREC == CM_PREC_FP64_FAST) {
                    // the FP64 baseline: FMA-contracted pair terms (it is
                    // judged against the IEEE path by tolerance, not bit for
                    // bit), lanes split over small groups as in FP32 split
                    double gx = 0.0, gy = 0.0, gz = 0.0;
                    auto term = [&](const double4 v, double f_shell_r) {
                        const double dx = xw - v.x, dy = yw - v.y, dz = zw - v.z;
                        double f;
                        if (f_shell_r < 0.0) {
                            const double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy,
                                              __builtin_fma(dz, dz, a.eps2)));
                            const double inv = rsqrt_fast(r2);
                            f = v.w * (inv * inv * inv);
                        } else {
                            f = gk_coeff(dx * dx + dy * dy + dz * dz, v.w, f_shell_r, a.eps2, true);
                        }
                        gx = __builtin_fma(-f, dx, gx);
                        gy = __builtin_fma(-f, dy, gy);
                        gz = __builtin_fma(-f, dz, gz);
                    };
                    const int trips = (cnt + kl - 1) >> klog;
                    if (cells || (!shell && a.eps2 >= kSelfTermMinEps2)) {
                        // the self term (dx = 0) is exactly zero at this
                        // softening; entries past cnt are massless and far
                        for (int t = 0; t < trips; t++) term(st.v[rbase + t], -1.0);
                    } else {
                        const int mt = s + tl;
                        for (int t = 0, j = sub; j < cnt; t++, j += kl) {
                            if (st.id[j] == mt) continue;
                            term(st.v[rbase + t], shell ? st.r[j] : -1.0);
                        }
                    }
                    apx += gx; apy += gy; apz += gz;
                }
            }
            if constexpr (LEARNED) {
                if (cells) learned_flush(cnt);
            }
            __syncthreads();
        };

        // one far cell per lane and half list: add its expansion coefficients
        // (both halves' cells are loaded first, one round trip for 128)
        auto far_flush = [&](int cnt) {
          if constexpr (kFar) {
            double4 fv[kPL / kWave];
#pragma unroll
            for (int hh = 0; hh < kPL / kWave; hh++)
                if (hh * kWave + lane < cnt) fv[hh] = a.cm[s_far[hh * kWave + lane]];
#pragma unroll
            for (int hh = 0; hh < kPL / kWave; hh++) {
              if (hh * kWave + lane < cnt) {
                const double4 v = fv[hh];
                const float rx = (float)(v.x - cxg), ry = (float)(v.y - cyg),
                            rz = (float)(v.z - czg), M = (float)v.w;
                const float s2 = __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry,
                                 __builtin_fmaf(rz, rz, (float)a.eps2)));
                const float ir = __builtin_amdgcn_rsqf(s2), ir2 = ir * ir;
                const float i3 = M * (ir * ir2), i5 = i3 * ir2, i7 = i5 * ir2;
                f0x = __builtin_fmaf(i3, rx, f0x);
                f0y = __builtin_fmaf(i3, ry, f0y);
                f0z = __builtin_fmaf(i3, rz, f0z);
                const float u = 3.0f * i5;
                const float ux = u * rx, uy = u * ry, uz = u * rz;
                f1xx += __builtin_fmaf(ux, rx, -i3);
                f1yy += __builtin_fmaf(uy, ry, -i3);
                f1zz += __builtin_fmaf(uz, rz, -i3);
                f1xy = __builtin_fmaf(ux, ry, f1xy);
                f1xz = __builtin_fmaf(ux, rz, f1xz);
                f1yz = __builtin_fmaf(uy, rz, f1yz);
                const float w = 15.0f * i7;
                const float qx = w * rx * rx - u, qy = w * ry * ry - u, qz = w * rz * rz - u;
                const float u2 = 2.0f * u;
                f2xxx = __builtin_fmaf(qx - u2, rx, f2xxx);   // 15 M x^3 s^-7/2 - 9 M x s^-5/2
                f2yyy = __builtin_fmaf(qy - u2, ry, f2yyy);
                f2zzz = __builtin_fmaf(qz - u2, rz, f2zzz);
                f2xxy = __builtin_fmaf(qx, ry, f2xxy);         // 15 M x^2 y s^-7/2 - 3 M y s^-5/2
                f2xxz = __builtin_fmaf(qx, rz, f2xxz);
                f2yyx = __builtin_fmaf(qy, rx, f2yyx);
                f2yyz = __builtin_fmaf(qy, rz, f2yyz);
                f2zzx = __builtin_fmaf(qz, rx, f2zzx);
                f2zzy = __builtin_fmaf(qz, ry, f2zzy);
                f2xyz = __builtin_fmaf(w * rx, ry * rz, f2xyz);
              }
            }
            n_far += (unsigned long long)cnt;
            any_far = 1;
            wcost += 5 * (kPL / kWave);
            __syncthreads();
          }
        };

        int S = 1, ncell = 0, npart = 0, nfar = 0;
        if (lane == 0) s_stack[0] = 0;
        __syncthreads();
        while (S > 0) {
            wcost += 14;
            int lim = (kCap - reserve - S) / 7;
            if (lim < 1) lim = 1;
            int take = S < kWave ? S : kWave;
            if (take > lim) take = lim;
            const int base = S - take;
            const int node = lane < take ? s_stack[base + lane] : -1;
            S = base;
            __syncthreads();
            int kind = 0;   // 1 cell, 2 leaf, 3 open, 4 far cell
            int4 nd = make_int4(-1, 0, 0, 0);
            if (node >= 0) {
                nd = a.info[node];
                const double4 c = a.cm[node];
                kind = mac_kind(c, nd, lo0, lo1, lo2, hi0, hi1, hi2, h, a.theta2, shell);
                if constexpr (kFar) {
                    const double rx = c.x - cxg, ry = c.y - cyg, rz = c.z - czg;
                    if (kind == 1 && rx * rx + ry * ry + rz * rz > far_lim) kind = 4;
                }
            }
            if constexpr (kFar) {
                uint64_t fm = __ballot(kind == 4);
                if (kind == 4) s_far[nfar + __popcll(fm & lt)] = node;
                nfar += __popcll(fm);
            }
            // accepted cells, in lane order
            uint64_t cm_ = __ballot(kind == 1);
            if (kind == 1) s_cell[ncell + __popcll(cm_ & lt)] = node;
            ncell += __popcll(cm_);
            // opened internal nodes push their children, in lane order
            int nop = kind == 3 ? (nd.y & 255) : 0;
            int oin = wscan_incl(nop);
            int otot = __builtin_amdgcn_readlane(oin, kWave - 1);
            if (kind == 3)
                for (int c = 0; c < nop; c++) s_stack[S + oin - nop + c] = nd.x + c;
            S += otot;
            // opened leaves: body ranges, expanded below in lane order
            int nl = kind == 2 ? nd.w : 0;
            int lin = wscan_incl(nl);
            int ltot = __builtin_amdgcn_readlane(lin, kWave - 1);
            // when every opened leaf holds at most kDirectLeaf bodies (all but
            // the deepest level's oversized leaves), each lane copies its own
            // leaf's bodies into their list positions
            const bool direct = __ballot(nl > kDirectLeaf) == 0;
            if (!direct) {
                s_lpref[lane] = lin - nl;
                s_lstart[lane] = nd.z;
            }
            __syncthreads();
            if constexpr (kFar) {
                if (nfar >= kPL) {
                    far_flush(kPL);
                    int rest = nfar - kPL;
                    int v = lane < rest ? s_far[kPL + lane] : 0;
                    __syncthreads();
                    if (lane < rest) s_far[lane] = v;
                    nfar = rest;
                    __syncthreads();
                }
            }
            while (ncell >= kPL) {
                flush(kPL, true);
                int rest = ncell - kPL;
                int v = lane < rest ? s_cell[kPL + lane] : 0;
                __syncthreads();
                if (lane < rest) s_cell[lane] = v;
                ncell = rest;
                __syncthreads();
            }
            if (direct && ltot > 0) {
                // positions [done, done + tk) of this step's leaf bodies fill
                // list slots [npart, npart + tk); a full list is flushed
                const int pref = lin - nl;
                int done = 0;
                while (done < ltot) {
                    const int tk = min(ltot - done, kPL - npart);
                    const int c0 = max(0, done - pref), c1 = min(nl, done + tk - pref);
                    const int at = npart + pref - done;
                    for (int c = c0; c < c1; c++) s_part[at + c] = nd.z + c;
                    npart += tk;
                    done += tk;
                    __syncthreads();
                    if (npart == kPL) { flush(kPL, false); npart = 0; }
                }
            }
            // otherwise position q of the opened leaves' bodies belongs to the last lane
            // whose leaf starts at or before q: per window of 64 positions,
            // each leaf marks its first position with its lane, and a running
            // maximum over the window (with the owner of the position before
            // it carried in) gives every position its leaf
            int done = direct ? ltot : 0, carry = -1;
            while (done < ltot) {
                // a window covers at most 64 positions (one per lane)
                const int room = min(kPL - npart, kWave);
                int tk = ltot - done < room ? ltot - done : room;
                s_mark[lane] = -1;
                __syncthreads();
                const int pref = lin - nl;
                if (nl > 0 && pref >= done && pref < done + kWave) s_mark[pref - done] = lane;
                __syncthreads();
                int own = s_mark[lane];
                if (lane == 0 && own < 0) own = carry;
                own = wmax_scan(own);
                if (lane < tk) {
                    int q = done + lane;
                    s_part[npart + lane] = s_lstart[own] + (q - s_lpref[own]);
                }
                carry = __builtin_amdgcn_readlane(own, tk - 1);
                npart += tk;
                done += tk;
                __syncthreads();
                if (npart == kPL) { flush(kPL, false); npart = 0; }
            }
        }
        if (ncell > 0) flush(ncell, true);
        if (npart > 0) flush(npart, false);
        if constexpr (kFar) {
            if (nfar > 0) far_flush(nfar);
        }

        if constexpr (LEARNED) {
            // group totals of residual and gradient, then a = res + J (x - centre)
            for (int o = 32; o > 0; o >>= 1) {
                for (int r = 0; r < 3; r++) lres[r] += __shfl_xor(lres[r], o, kWave);
                for (int r = 0; r < 6; r++) lgrad[r] += __shfl_xor(lgrad[r], o, kWave);
            }
            if (member) {
                apx += (double)(lres[0] + lgrad[0] * txf + lgrad[3] * tyf + lgrad[4] * tzf);
                apy += (double)(lres[1] + lgrad[3] * txf + lgrad[1] * tyf + lgrad[5] * tzf);
                apz += (double)(lres[2] + lgrad[4] * txf + lgrad[5] * tyf + lgrad[2] * tzf);
            }
        }

        if constexpr (kSplit) {
            // add the kl partial sums of each target in sub order
            for (int q = 1; q < kl; q++) {
                const int src = lane + q * nmem;
                const double bx = __shfl(apx, src < kWave ? src : lane, kWave);
                const double by = __shfl(apy, src < kWave ? src : lane, kWave);
                const double bz = __shfl(apz, src < kWave ? src : lane, kWave);
                if (member) { apx += bx; apy += by; apz += bz; }
            }
        }
        if constexpr (kFar) {
            if (any_far) {
                // group sums of the coefficients (in double), then each
                // member evaluates the expansion at its offset d from c
                double t[19] = {f0x, f0y, f0z, f1xx, f1yy, f1zz, f1xy, f1xz, f1yz,
                                f2xxx, f2yyy, f2zzz, f2xxy, f2xxz, f2yyx, f2yyz, f2zzx,
                                f2zzy, f2xyz};
                for (int o = 32; o > 0; o >>= 1)
#pragma unroll
                    for (int k = 0; k < 19; k++) t[k] += __shfl_xor(t[k], o, kWave);
                if (member) {
                    const double dx = txf, dy = tyf, dz = tzf;
                    const double xx = dx * dx, yy = dy * dy, zz = dz * dz;
                    const double xy = dx * dy, xz = dx * dz, yz = dy * dz;
                    apx += t[0] + (t[3] * dx + t[6] * dy + t[7] * dz)
                           + 0.5 * (t[9] * xx + t[14] * yy + t[16] * zz)
                           + (t[12] * xy + t[13] * xz + t[18] * yz);
                    apy += t[1] + (t[6] * dx + t[4] * dy + t[8] * dz)
                           + 0.5 * (t[12] * xx + t[10] * yy + t[17] * zz)
                           + (t[14] * xy + t[18] * xz + t[15] * yz);
                    apz += t[2] + (t[7] * dx + t[8] * dy + t[5] * dz)
                           + 0.5 * (t[13] * xx + t[15] * yy + t[11] * zz)
                           + (t[18] * xy + t[16] * xz + t[17] * yz);
                }
            }
        }
        if (member) {
            const int o = a.order[my];
            // a part records the whole group's work (the parts' share of
            // the targets scaled up), which the next order reads
            a.bcost[o] = whole ? wcost : (int)min((long long)wcost * (ge - gs) / nmem,
                                                   (long long)INT32_MAX);
            if constexpr (POT) {
                a.rows[o] = 0.5 * (sm[my] * (acx + apx));
            } else if constexpr (PREC == CM_PREC_FP32_SPLIT) {
                a.ax[o] = apx; a.ay[o] = apy; a.az[o] = apz;
            } else {
                a.ax[o] = acx + apx; a.ay[o] = acy + apy; a.az[o] = acz + apz;
            }
        }
    }
    // inactive bodies sit after the active ones in sorted order: zero their
    // acc (IPhysicsComputeBackend.cs:21-33)
    for (int k = a.hdr->n_active + blockIdx.x * kWave + lane; k < a.n; k += gridDim.x * kWave) {
        const int o = a.order[k];
        if constexpr (POT) {
            a.rows[o] = 0.0;
        } else {
            a.ax[o] = 0.0; a.ay[o] = 0.0; a.az[o] = 0.0;
        }
    }
    if (lane == 0) {
        atomicAdd(&a.counters[0], n_cell);
        atomicAdd(&a.counters[1], n_body);
        atomicAdd(&a.counters[2], n_cell + n_body);
        if (n_far) {
            atomicAdd(&a.counters[3], n_far);
            atomicAdd(&a.counters[4], n_far);
        }
    }
}

int form_groups(Engine& e, const int32_t* order) {
    const int n = e.count;
    const int gb = std::min(1024, std::max(1, (n + 255) / 256));
    const int tiles = std::max(1, (n + kGroupTile - 1) / kGroupTile);
    // group tile counters are cleared by k_morton
    hipLaunchKernelGGL(k_groups, dim3(gb), dim3(256), 0, e.stream, e.hdr, e.nd_info,
                       e.nd_box, e.cfg.group_size, e.gmark, e.gtile);
    const int scanned = tiles > kGroupPrefixTiles ? 1 : 0;
    if (scanned)
        hipLaunchKernelGGL(k_group_scan, dim3(1), dim3(1024), 0, e.stream, tiles, e.hdr,
                           e.gtile);
    hipLaunchKernelGGL(k_group_emit, dim3(tiles), dim3(kGroupTile), 0, e.stream, tiles,
                       scanned, e.hdr, e.gmark, e.gtile, order, e.bcost, e.groups, e.gcost,
                       e.gwork);
    // parts of fat groups only for the bit-exact FP64 path, whose sums they
    // leave unchanged
    const int split = e.cfg.precision == CM_PREC_FP64 ? 1 : 0;
    hipLaunchKernelGGL(k_group_order, dim3(kXcds), dim3(kOrderThreads), 0, e.stream, tiles,
                       split, e.gq_cap, e.hdr, e.groups, e.gcost, e.gwork, e.gq);
    CM_HIP_TRY(e, hipGetLastError());
    return CM_OK;
}

int launch_barnes_hut(Engine& e, double softening) {
    const int n = e.count;
    e.stats = cm_force_stats{};
    if (n == 0) return CM_OK;
    hipStream_t st = e.stream;
    if (e.detail_timing()) CM_HIP_TRY(e, hipEventRecord(e.ev[0], st));
    e.mark(0);
    const void* keys = nullptr;
    const int32_t* order = nullptr;
    const bool learned = e.cfg.learned_correction != 0;
    if (learned && (e.cfg.precision != CM_PREC_FP32_SPLIT || !mlp_ready(e)))
        return set_error(e, learned && e.cfg.precision != CM_PREC_FP32_SPLIT
                                ? CM_ERR_INVALID_ARGUMENT : CM_ERR_NOT_INITIALIZED,
                         "the learned correction needs FP32_SPLIT and a network "
                         "(cm_mlp_init / cm_mlp_set_params / cm_mlp_train)");
    int rc = build_tree(e, &keys, &order);
    if (rc != CM_OK) return rc;
    rc = form_groups(e, order);
    if (rc != CM_OK) return rc;
    if (e.detail_timing()) CM_HIP_TRY(e, hipEventRecord(e.ev[3], st));
    e.mark(1);
    TravArgs a;
    a.hdr = e.hdr;
    a.info = e.nd_info;
    a.cm = e.nd_cm;
    a.order = order;
    a.queue = e.gq;
    a.bcost = e.bcost;
    a.sorted = e.sorted;
    a.cap = e.capacity;
    a.n = n;
    a.ax = e.f[AX]; a.ay = e.f[AY]; a.az = e.f[AZ];
    a.eps2 = softening * softening;
    double th = e.cfg.theta > 0.2 ? e.cfg.theta : 0.2;   // BarnesHutBackend.cs:129
    a.theta2 = th * th;
    a.levels = e.cfg.morton_levels;
    a.shell = e.cfg.shell_theorem;
    a.counters = &e.hdr->n_cell_int;
    a.far_ratio2 = e.cfg.far_field_ratio * e.cfg.far_field_ratio;
    a.mlp_w = learned ? mlp_weights(e) : nullptr;
    a.mlp_p = learned ? mlp_params(e) : nullptr;
    a.quad = e.nd_quad;
    // persistent wavefronts: as many as fit on the 256 CUs (4 SIMDs each),
    // never more than bodies / 8
    const bool split = e.cfg.precision == CM_PREC_FP32_SPLIT && !learned;
    const int per_simd = split ? TravShape<CM_PREC_FP32_SPLIT, false>::waves : 4;
    static_assert(TravShape<CM_PREC_FP32_SPLIT, false>::waves == 4, "grid sized for 4 per SIMD");
    int nblocks = std::max(1, std::min(1024 * per_simd, (n + 7) / 8));
    switch (e.cfg.precision) {
        case CM_PREC_FP64:
            hipLaunchKernelGGL((k_traverse<CM_PREC_FP64, false>), dim3(nblocks), dim3(kWave), 0, st, a);
            break;
        case CM_PREC_FP64_FAST:
            hipLaunchKernelGGL((k_traverse<CM_PREC_FP64_FAST, false>), dim3(nblocks), dim3(kWave), 0, st, a);
            break;
        default:
            if (learned)
                hipLaunchKernelGGL((k_traverse<CM_PREC_FP32_SPLIT, true>), dim3(nblocks), dim3(kWave), 0, st, a);
            else
                hipLaunchKernelGGL((k_traverse<CM_PREC_FP32_SPLIT, false>), dim3(nblocks), dim3(kWave), 0, st, a);
            break;
    }
    CM_HIP_TRY(e, hipGetLastError());
    if (e.detail_timing()) CM_HIP_TRY(e, hipEventRecord(e.ev[5], st));
    e.mark(2);
    if (e.stats_eval)
        CM_HIP_TRY(e, hipMemcpyAsync(e.hdr_host, e.hdr, sizeof(TreeHeader),
                                 hipMemcpyDeviceToHost, st));
    return CM_OK;
}

// Barnes-Hut potential of every body for the current state: rows[i] =
// 0.5 m_i phi_i (0 for inactive bodies), with the backend's theta (clamped to
// >= 0.2) and this softening, through the FP64 group traversal on a fresh
// tree; the force statistics are left alone.
int launch_bh_potential(Engine& e, double softening, double* rows) {
    const int n = e.count;
    if (n == 0) return CM_OK;
    const void* keys = nullptr;
    const int32_t* order = nullptr;
    int rc = build_tree(e, &keys, &order);
    if (rc != CM_OK) return rc;
    rc = form_groups(e, order);
    if (rc != CM_OK) return rc;
    TravArgs a{};
    a.hdr = e.hdr;
    a.info = e.nd_info;
    a.cm = e.nd_cm;
    a.order = order;
    a.queue = e.gq;
    a.bcost = e.bcost;
    a.sorted = e.sorted;
    a.cap = e.capacity;
    a.n = n;
    a.rows = rows;
    a.eps2 = softening * softening;
    double th = e.cfg.theta > 0.2 ? e.cfg.theta : 0.2;
    a.theta2 = th * th;
    a.levels = e.cfg.morton_levels;
    a.shell = 0;   // the potential kernel has no shell-theorem branch
    a.counters = e.hdr->pot_int;
    int nblocks = std::max(1, std::min(1024 * 4, (n + 7) / 8));
    hipLaunchKernelGGL((k_traverse<CM_PREC_FP64, false, true>), dim3(nblocks), dim3(kWave), 0,
                       e.stream, a);
    CM_HIP_TRY(e, hipGetLastError());
    return CM_OK;
}

}  // namespace cm

extern "C" int32_t cm_test_groups(void* h, int32_t cap, int32_t* groups, int32_t* queue,
                                  int32_t* qbounds, int32_t* cost, int32_t* order,
                                  int32_t* n_groups) {
    using namespace cm;
    Engine* e = static_cast<Engine*>(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (cap < 0) return set_error(*e, CM_ERR_INVALID_ARGUMENT, "negative cap");
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    const int n = e->count;
    if (n_groups) *n_groups = 0;
    if (n == 0) return CM_OK;
    const void* k = nullptr;
    const int32_t* o = nullptr;
    int rc = build_tree(*e, &k, &o);
    if (rc != CM_OK) return rc;
    rc = form_groups(*e, o);
    if (rc != CM_OK) return rc;
    TreeHeader hh;
    CM_HIP_TRY(*e, hipMemcpyAsync(&hh, e->hdr, sizeof(hh), hipMemcpyDeviceToHost, e->stream));
    CM_HIP_TRY(*e, hipStreamSynchronize(e->stream));
    if (hh.overflow) return set_error(*e, CM_ERR_POOL_EXHAUSTED, "octree node pool exhausted");
    if (n_groups) *n_groups = hh.n_groups;
    const size_t m = (size_t)std::min(cap, hh.n_groups);
    if (groups && m)
        CM_HIP_TRY(*e, hipMemcpyAsync(groups, e->groups, m * sizeof(int2), hipMemcpyDeviceToHost,
                                      e->stream));
    // the queue: entries [0, qend[7]) of which queue q holds [qbase[q], qend[q])
    const size_t mq = (size_t)std::min(cap, std::min(hh.qend[kXcds - 1], e->gq_cap));
    if (queue && mq)
        CM_HIP_TRY(*e, hipMemcpyAsync(queue, e->gq, mq * sizeof(int4), hipMemcpyDeviceToHost,
                                      e->stream));
    if (qbounds)
        for (int q = 0; q < kXcds; q++) {
            qbounds[q] = hh.qbase[q];
            qbounds[kXcds + q] = hh.qend[q];
        }
    if (cost)
        CM_HIP_TRY(*e, hipMemcpyAsync(cost, e->bcost, (size_t)n * 4, hipMemcpyDeviceToHost,
                                      e->stream));
    if (order)
        CM_HIP_TRY(*e, hipMemcpyAsync(order, o, (size_t)n * 4, hipMemcpyDeviceToHost, e->stream));
    CM_HIP_TRY(*e, hipStreamSynchronize(e->stream));
    return CM_OK;
}
