/*
 * cm_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see cm_oracle.h).
 *
 * Compile with -ffp-contract=off: every a*b+c below must round twice, as the
 * .NET JIT does for the reference's C# (it never contracts to FMA).
 *
 * Citations are to the reference tree (mfscpayload-690/celestial-mechanics-desktop):
 *   Solvers/...      = src/CelestialMechanics.Physics/Solvers/...
 *   Integrators/...  = src/CelestialMechanics.Physics/Integrators/...
 *   etc.
 */
#include "cm_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* ======================================================================== */
/* Gravity                                                                  */
/* ======================================================================== */

/* GravityKernel.AccelerationCoeffFromSource (Solvers/GravityKernel.cs:9-33):
 * m/(r^2+eps^2)^{3/2}, or the linear interior value m/(R^2+eps^2)^{3/2} when
 * the shell theorem is on and r < R.  m*inv*inv*inv associates left to right. */
double cmo_gravity_coeff(double dist_sq, double src_mass, double src_radius,
                         double eps2, int shell) {
    if (!shell || src_radius <= 0.0) {
        double inv = 1.0 / sqrt(dist_sq + eps2);
        return src_mass * inv * inv * inv;
    }
    double dist = sqrt(dist_sq);
    if (dist < src_radius) {
        double denom = src_radius * src_radius + eps2;
        double inv_denom_sqrt = 1.0 / sqrt(denom);
        return src_mass * inv_denom_sqrt / denom;
    }
    double inv = 1.0 / sqrt(dist_sq + eps2);
    return src_mass * inv * inv * inv;
}

/* CpuSingleThreadBackend.ComputeForces, non-shell branch
 * (Solvers/CpuSingleThreadBackend.cs:50-121): Newton-3 over i<j; body j
 * receives += factor_ji*dx from every earlier i, then its own row sum is
 * added at the end of its i-iteration. */
void cmo_direct_n3(int n, const double* px, const double* py, const double* pz,
                   const double* m, const uint8_t* act, double eps,
                   double* ax, double* ay, double* az) {
    double eps2 = eps * eps;
    for (int i = 0; i < n; i++) { ax[i] = 0.0; ay[i] = 0.0; az[i] = 0.0; }
    for (int i = 0; i < n; i++) {
        if (!act[i]) continue;
        double xi = px[i], yi = py[i], zi = pz[i], mi = m[i];
        double axi = 0.0, ayi = 0.0, azi = 0.0;
        for (int j = i + 1; j < n; j++) {
            if (!act[j]) continue;
            double dx = xi - px[j];
            double dy = yi - py[j];
            double dz = zi - pz[j];
            double dist2 = dx * dx + dy * dy + dz * dz + eps2;
            double inv = 1.0 / sqrt(dist2);
            double inv3 = inv * inv * inv;
            double fij = m[j] * inv3;
            double fji = mi * inv3;
            axi -= fij * dx; ayi -= fij * dy; azi -= fij * dz;
            ax[j] += fji * dx; ay[j] += fji * dy; az[j] += fji * dz;
        }
        ax[i] += axi; ay[i] += ayi; az[i] += azi;
    }
}

/* CpuParallelBackend.ComputeForces (Solvers/CpuParallelBackend.cs:69-142) and
 * the shell-theorem loop of CpuSingleThreadBackend (:124-158): full j sweep,
 * j ascending, GravityKernel coefficient, result assigned (not added). */
void cmo_direct_gk(int n, const double* px, const double* py, const double* pz,
                   const double* m, const double* rad, const uint8_t* act,
                   double eps, int shell, double* ax, double* ay, double* az) {
    double eps2 = eps * eps;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        ax[i] = 0.0; ay[i] = 0.0; az[i] = 0.0;
        if (!act[i]) continue;
        double xi = px[i], yi = py[i], zi = pz[i];
        double axi = 0.0, ayi = 0.0, azi = 0.0;
        for (int j = 0; j < n; j++) {
            if (j == i || !act[j]) continue;
            double dx = xi - px[j];
            double dy = yi - py[j];
            double dz = zi - pz[j];
            double raw = dx * dx + dy * dy + dz * dz;
            double c = cmo_gravity_coeff(raw, m[j], rad[j], eps2, shell);
            axi -= c * dx; ayi -= c * dy; azi -= c * dz;
        }
        ax[i] = axi; ay[i] = ayi; az[i] = azi;
    }
}

/* FP64 direct sum restricted to a target subset: the force-accuracy judge of
 * ForceErrorAnalyzer.CompareForces (Validation/ForceErrorAnalyzer.cs:67-92)
 * evaluated only on sampled bodies (non-shell GravityKernel order). */
void cmo_direct_targets(int n, const double* px, const double* py,
                        const double* pz, const double* m, const uint8_t* act,
                        double eps, int nt, const int32_t* targets,
                        double* ax, double* ay, double* az, int nthreads) {
    double eps2 = eps * eps;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 4)
    for (int t = 0; t < nt; t++) {
        int i = targets[t];
        double xi = px[i], yi = py[i], zi = pz[i];
        double axi = 0.0, ayi = 0.0, azi = 0.0;
        if (act[i]) {
            for (int j = 0; j < n; j++) {
                if (j == i || !act[j]) continue;
                double dx = xi - px[j];
                double dy = yi - py[j];
                double dz = zi - pz[j];
                double raw = dx * dx + dy * dy + dz * dz;
                double c = cmo_gravity_coeff(raw, m[j], 0.0, eps2, 0);
                axi -= c * dx; ayi -= c * dy; azi -= c * dz;
            }
        }
        ax[t] = axi; ay[t] = ayi; az[t] = azi;
    }
}

/* ======================================================================== */
/* Reference Barnes-Hut: C# insertion octree                                */
/* ======================================================================== */

typedef struct {
    double cx, cy, cz, half;
    double mass, comx, comy, comz;
    int32_t body;
    int32_t leaf;
    int32_t child[8];
} csnode;

typedef struct {
    csnode* nodes;
    int64_t count, cap;
    double min_node_size;
} cspool;

/* OctreePool.Allocate + OctreeNode.Init (BarnesHut/OctreePool.cs:37-80,
 * BarnesHut/OctreeNode.cs:150-175); the pool doubles when exhausted. */
static int32_t cs_alloc(cspool* p, double cx, double cy, double cz, double h) {
    if (p->count >= p->cap) {
        int64_t nc = p->cap * 2;
        p->nodes = (csnode*)realloc(p->nodes, (size_t)nc * sizeof(csnode));
        p->cap = nc;
    }
    csnode* nd = &p->nodes[p->count];
    nd->cx = cx; nd->cy = cy; nd->cz = cz; nd->half = h;
    nd->mass = 0.0; nd->comx = 0.0; nd->comy = 0.0; nd->comz = 0.0;
    nd->body = -1; nd->leaf = 0;
    for (int k = 0; k < 8; k++) nd->child[k] = -1;
    return (int32_t)(p->count++);
}

/* OctreeNode.OctantFor: x->4, y->2, z->1 (BarnesHut/OctreeNode.cs:185-192). */
static int cs_octant(const csnode* nd, double x, double y, double z) {
    int o = 0;
    if (x >= nd->cx) o |= 4;
    if (y >= nd->cy) o |= 2;
    if (z >= nd->cz) o |= 1;
    return o;
}

/* BarnesHutBackend.UpdateMassFromInsertion (Solvers/BarnesHutBackend.cs:359-371). */
static void cs_update_mass(csnode* nd, double bm, double bx, double by, double bz) {
    double nm = nd->mass + bm;
    if (nm > 0.0) {
        double inv = 1.0 / nm;
        nd->comx = (nd->comx * nd->mass + bx * bm) * inv;
        nd->comy = (nd->comy * nd->mass + by * bm) * inv;
        nd->comz = (nd->comz * nd->mass + bz * bm) * inv;
    }
    nd->mass = nm;
}

static void cs_insert(cspool* p, int32_t ni, int32_t b, double bx, double by,
                      double bz, double bm);

/* BarnesHutBackend.InsertIntoChild (Solvers/BarnesHutBackend.cs:332-349). */
static void cs_insert_child(cspool* p, int32_t pi, int32_t b, double bx,
                            double by, double bz, double bm) {
    int o = cs_octant(&p->nodes[pi], bx, by, bz);
    int32_t ci = p->nodes[pi].child[o];
    if (ci == -1) {
        /* OctreeNode.ChildCenter (BarnesHut/OctreeNode.cs:199-205) */
        double q = p->nodes[pi].half * 0.5;
        double cx = p->nodes[pi].cx + ((o & 4) ? q : -q);
        double cy = p->nodes[pi].cy + ((o & 2) ? q : -q);
        double cz = p->nodes[pi].cz + ((o & 1) ? q : -q);
        ci = cs_alloc(p, cx, cy, cz, p->nodes[pi].half * 0.5);
        p->nodes[pi].child[o] = ci;
    }
    cs_insert(p, ci, b, bx, by, bz, bm);
}

/* BarnesHutBackend.InsertBody (Solvers/BarnesHutBackend.cs:258-326). */
static void cs_insert(cspool* p, int32_t ni, int32_t b, double bx, double by,
                      double bz, double bm) {
    csnode* nd = &p->nodes[ni];
    if (nd->mass == 0.0 && nd->body == -1) {
        nd->body = b; nd->leaf = 1; nd->mass = bm;
        nd->comx = bx; nd->comy = by; nd->comz = bz;
        return;
    }
    if (nd->leaf) {
        if (nd->half * 0.5 < p->min_node_size) {
            /* coincident aggregation (:279-292) -- uses division, not a
             * reciprocal, as the reference does here */
            double tm = nd->mass + bm;
            nd->comx = (nd->comx * nd->mass + bx * bm) / tm;
            nd->comy = (nd->comy * nd->mass + by * bm) / tm;
            nd->comz = (nd->comz * nd->mass + bz * bm) / tm;
            nd->mass = tm;
            return;
        }
        int32_t eb = nd->body;
        double em = nd->mass, ex = nd->comx, ey = nd->comy, ez = nd->comz;
        nd->body = -1; nd->leaf = 0; nd->mass = 0.0;
        nd->comx = 0.0; nd->comy = 0.0; nd->comz = 0.0;
        cs_insert_child(p, ni, eb, ex, ey, ez, em);
        cs_insert_child(p, ni, b, bx, by, bz, bm);
        cs_update_mass(&p->nodes[ni], em, ex, ey, ez);
        cs_update_mass(&p->nodes[ni], bm, bx, by, bz);
        return;
    }
    cs_insert_child(p, ni, b, bx, by, bz, bm);
    cs_update_mass(&p->nodes[ni], bm, bx, by, bz);
}

/* BarnesHutBackend.ComputeForceOnBody (Solvers/BarnesHutBackend.cs:392-459). */
static void cs_force(const csnode* nodes, int32_t ni, int32_t b, double bx,
                     double by, double bz, double eps2, double theta2,
                     const double* rad, int shell, double* axi, double* ayi,
                     double* azi, int64_t* nint) {
    const csnode* nd = &nodes[ni];
    if (nd->mass == 0.0) return;
    double dx = bx - nd->comx;
    double dy = by - nd->comy;
    double dz = bz - nd->comz;
    double dist2 = dx * dx + dy * dy + dz * dz;
    if (nd->leaf) {
        if (nd->body == b) return;
        double sr = nd->body >= 0 ? rad[nd->body] : 0.0;
        double f = cmo_gravity_coeff(dist2, nd->mass, sr, eps2, shell);
        *axi -= f * dx; *ayi -= f * dy; *azi -= f * dz;
        (*nint)++;
        return;
    }
    double size2 = 4.0 * nd->half * nd->half;
    if (dist2 > 0.0 && size2 < theta2 * dist2) {
        double soft = dist2 + eps2;
        double inv = 1.0 / sqrt(soft);
        double inv3 = inv * inv * inv;
        double f = nd->mass * inv3;
        *axi -= f * dx; *ayi -= f * dy; *azi -= f * dz;
        (*nint)++;
        return;
    }
    for (int k = 0; k < 8; k++)
        if (nd->child[k] != -1)
            cs_force(nodes, nd->child[k], b, bx, by, bz, eps2, theta2, rad,
                     shell, axi, ayi, azi, nint);
}

/* BarnesHutBackend.ComputeForces (Solvers/BarnesHutBackend.cs:118-236).
 * The build is serial as in the reference; the traversal runs on nthreads
 * OpenMP threads (static schedule), standing in for Parallel.For. */
int cmo_bh_csharp(int n, const double* px, const double* py, const double* pz,
                  const double* m, const double* rad, const uint8_t* act,
                  double eps, double theta, int shell, double min_node_size,
                  int nthreads, double* ax, double* ay, double* az,
                  cmo_bh_stats* stats) {
    if (n == 0) return 0;
    double t0 = now_ms();
    double eps2 = eps * eps;
    double th = theta > 0.2 ? theta : 0.2;
    double theta2 = th * th;
    for (int i = 0; i < n; i++) { ax[i] = 0.0; ay[i] = 0.0; az[i] = 0.0; }
    double box[4];
    cmo_root_box(n, px, py, pz, act, box);
    cspool pool;
    pool.cap = 4 * (int64_t)n + 64;
    pool.nodes = (csnode*)malloc((size_t)pool.cap * sizeof(csnode));
    if (!pool.nodes) return -1;
    pool.count = 0;
    pool.min_node_size = min_node_size;
    int32_t root = cs_alloc(&pool, box[0], box[1], box[2], box[3]);
    for (int i = 0; i < n; i++) {
        if (!act[i]) continue;
        cs_insert(&pool, root, i, px[i], py[i], pz[i], m[i]);
    }
    double t1 = now_ms();
    int64_t total = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    (void)nthreads;
#pragma omp parallel for schedule(static) reduction(+ : total)
    for (int i = 0; i < n; i++) {
        if (!act[i]) continue;
        double axi = 0.0, ayi = 0.0, azi = 0.0;
        int64_t ni = 0;
        cs_force(pool.nodes, root, i, px[i], py[i], pz[i], eps2, theta2, rad,
                 shell, &axi, &ayi, &azi, &ni);
        ax[i] = axi; ay[i] = ayi; az[i] = azi;
        total += ni;
    }
    double t2 = now_ms();
    if (stats) {
        stats->node_count = pool.count;
        stats->interactions = total;
        stats->build_ms = t1 - t0;
        stats->traverse_ms = t2 - t1;
    }
    free(pool.nodes);
    return 0;
}

/* ======================================================================== */
/* Verlet                                                                   */
/* ======================================================================== */

/* SoAVerletIntegrator.Step phases 1-2 (Integrators/SoAVerletIntegrator.cs:74-90). */
void cmo_verlet_kick_drift(int n, const uint8_t* act, double dt,
                           double* px, double* py, double* pz,
                           double* vx, double* vy, double* vz,
                           const double* oax, const double* oay, const double* oaz) {
    double half = 0.5 * dt;
    for (int i = 0; i < n; i++) {
        if (!act[i]) continue;
        vx[i] += half * oax[i];
        vy[i] += half * oay[i];
        vz[i] += half * oaz[i];
    }
    for (int i = 0; i < n; i++) {
        if (!act[i]) continue;
        px[i] += dt * vx[i];
        py[i] += dt * vy[i];
        pz[i] += dt * vz[i];
    }
}

/* SoAVerletIntegrator.Step phases 4-5 (:99-115): the rotation OldAcc=Acc runs
 * over every body, active or not. */
void cmo_verlet_kick_rotate(int n, const uint8_t* act, double dt,
                            double* vx, double* vy, double* vz,
                            const double* ax, const double* ay, const double* az,
                            double* oax, double* oay, double* oaz) {
    double half = 0.5 * dt;
    for (int i = 0; i < n; i++) {
        if (!act[i]) continue;
        vx[i] += half * ax[i];
        vy[i] += half * ay[i];
        vz[i] += half * az[i];
    }
    for (int i = 0; i < n; i++) {
        oax[i] = ax[i]; oay[i] = ay[i]; oaz[i] = az[i];
    }
}

/* ======================================================================== */
/* Build-defined octree: root cube, Morton keys, stable sort, level-order   */
/* bucket octree, bottom-up centre of mass, group-MAC traversal.            */
/* ======================================================================== */

/* Root cube exactly as BarnesHutBackend.ComputeForces (Solvers/BarnesHutBackend.cs:150-170). */
void cmo_root_box(int n, const double* px, const double* py, const double* pz,
                  const uint8_t* act, double* box4) {
    double minx = DBL_MAX, miny = DBL_MAX, minz = DBL_MAX;
    double maxx = -DBL_MAX, maxy = -DBL_MAX, maxz = -DBL_MAX;
    for (int i = 0; i < n; i++) {
        if (!act[i]) continue;
        if (px[i] < minx) minx = px[i];
        if (py[i] < miny) miny = py[i];
        if (pz[i] < minz) minz = pz[i];
        if (px[i] > maxx) maxx = px[i];
        if (py[i] > maxy) maxy = py[i];
        if (pz[i] > maxz) maxz = pz[i];
    }
    double cx = (minx + maxx) * 0.5;
    double cy = (miny + maxy) * 0.5;
    double cz = (minz + maxz) * 0.5;
    double ex = maxx - minx, ey = maxy - miny, ez = maxz - minz;
    double e = ey > ez ? ey : ez;
    e = ex > e ? ex : e;
    double h = e * 0.5;
    h = h > 1e-10 ? h : 1e-10;   /* Math.Max(halfSize, 1e-10) */
    h *= 1.001;
    box4[0] = cx; box4[1] = cy; box4[2] = cz; box4[3] = h;
}

/* Quantise one coordinate to [0, 2^levels - 1] over [c-h, c+h). */
static uint32_t quantise(double x, double c, double h, int levels) {
    double cells = ldexp(1.0, levels);
    double scale = cells / (2.0 * h);
    double u = (x - (c - h)) * scale;
    if (!(u >= 0.0)) u = 0.0;
    if (u > cells - 1.0) u = cells - 1.0;
    return (uint32_t)u;
}

/* Morton key, x bit most significant in each 3-bit group, so the top group is
 * the C# octant number (x->4, y->2, z->1; BarnesHut/OctreeNode.cs:185-192). */
uint64_t cmo_morton_key(double x, double y, double z, const double* box4,
                        int levels) {
    uint32_t qx = quantise(x, box4[0], box4[3], levels);
    uint32_t qy = quantise(y, box4[1], box4[3], levels);
    uint32_t qz = quantise(z, box4[2], box4[3], levels);
    uint64_t k = 0;
    for (int b = levels - 1; b >= 0; b--) {
        k = (k << 3) | ((uint64_t)((qx >> b) & 1u) << 2) |
            ((uint64_t)((qy >> b) & 1u) << 1) | (uint64_t)((qz >> b) & 1u);
    }
    return k;
}

void cmo_morton_keys(int n, const double* px, const double* py,
                     const double* pz, const uint8_t* act, const double* box4,
                     int levels, uint64_t* keys) {
    for (int i = 0; i < n; i++)
        keys[i] = act[i] ? cmo_morton_key(px[i], py[i], pz[i], box4, levels)
                         : UINT64_MAX;
}

/* Stable sort of (key, original index): counting sort, one byte per pass,
 * least significant first.  The result is the unique stable order. */
void cmo_stable_sort_keys(int n, const uint64_t* keys, uint64_t* keys_sorted,
                          int32_t* order) {
    uint64_t* k0 = (uint64_t*)malloc((size_t)n * 8 + 8);
    uint64_t* k1 = (uint64_t*)malloc((size_t)n * 8 + 8);
    int32_t* o0 = (int32_t*)malloc((size_t)n * 4 + 4);
    int32_t* o1 = (int32_t*)malloc((size_t)n * 4 + 4);
    for (int i = 0; i < n; i++) { k0[i] = keys[i]; o0[i] = i; }
    for (int pass = 0; pass < 8; pass++) {
        int sh = pass * 8;
        size_t cnt[257];
        memset(cnt, 0, sizeof(cnt));
        for (int i = 0; i < n; i++) cnt[((k0[i] >> sh) & 255u) + 1]++;
        for (int d = 0; d < 256; d++) cnt[d + 1] += cnt[d];
        for (int i = 0; i < n; i++) {
            size_t pos = cnt[(k0[i] >> sh) & 255u]++;
            k1[pos] = k0[i]; o1[pos] = o0[i];
        }
        uint64_t* tk = k0; k0 = k1; k1 = tk;
        int32_t* to = o0; o0 = o1; o1 = to;
    }
    memcpy(keys_sorted, k0, (size_t)n * 8);
    memcpy(order, o0, (size_t)n * 4);
    free(k0); free(k1); free(o0); free(o1);
}

/* Level-order bucket octree over the sorted keys of the active bodies.
 * Node 0 is the root (level 0) covering all active bodies.  A node with
 * count <= leaf_cap, or at the deepest level, is a leaf; otherwise its
 * non-empty children (digit 0..7 of the next 3 key bits) are appended, in
 * digit order, after all nodes of its own level.  Children of one node are
 * therefore contiguous and every level is a contiguous index range.
 * Mass and centre of mass are then formed bottom-up: a leaf sums m and m*x
 * over its bodies in sorted order, an internal node sums its children's m and
 * m*x in child order, and com = (sum m*x) / (sum m) (0 when the mass is 0). */
typedef struct {
    int32_t level, start, count, first, nchild;
    double mass, sx, sy, sz, cx, cy, cz;
} tnode;

static int tree_build_internal(int n, const double* px, const double* py,
                               const double* pz, const double* m,
                               const uint8_t* act, int levels, int leaf_cap,
                               double* box4, int* n_active_out,
                               uint64_t** keys_out, int32_t** order_out,
                               tnode** nodes_out, int* nn_out,
                               int32_t* level_start) {
    cmo_root_box(n, px, py, pz, act, box4);
    uint64_t* keys = (uint64_t*)malloc((size_t)n * 8 + 8);
    uint64_t* ks = (uint64_t*)malloc((size_t)n * 8 + 8);
    int32_t* order = (int32_t*)malloc((size_t)n * 4 + 4);
    cmo_morton_keys(n, px, py, pz, act, box4, levels, keys);
    cmo_stable_sort_keys(n, keys, ks, order);
    free(keys);
    int na = 0;
    for (int i = 0; i < n; i++) na += act[i] ? 1 : 0;
    int cap = 2 * na + 16;
    tnode* nd = (tnode*)malloc((size_t)cap * sizeof(tnode));
    int nn = 1;
    memset(&nd[0], 0, sizeof(tnode));
    nd[0].level = 0; nd[0].start = 0; nd[0].count = na;
    int ls = 0, le = 1;
    if (level_start) { level_start[0] = 0; level_start[1] = 1; }
    for (int L = 0; L <= levels && ls < le; L++) {
        for (int k = ls; k < le; k++) {
            tnode* t = &nd[k];
            if (t->count <= leaf_cap || L == levels) {
                t->first = -1; t->nchild = 0;
                continue;
            }
            int sh = 3 * (levels - L - 1);
            t->first = nn; t->nchild = 0;
            int s = t->start, e = t->start + t->count;
            int i = s;
            while (i < e) {
                int d = (int)((ks[i] >> sh) & 7u);
                int j = i;
                while (j < e && (int)((ks[j] >> sh) & 7u) == d) j++;
                if (nn >= cap) {
                    cap *= 2;
                    nd = (tnode*)realloc(nd, (size_t)cap * sizeof(tnode));
                    t = &nd[k];
                }
                memset(&nd[nn], 0, sizeof(tnode));
                nd[nn].level = L + 1; nd[nn].start = i; nd[nn].count = j - i;
                nn++;
                t->nchild++;
                i = j;
            }
        }
        ls = le; le = nn;
        if (level_start) level_start[L + 2] = nn;
    }
    /* bottom-up mass and centre of mass: children always have larger
     * indices than their parent, so a reverse sweep sees children first */
    for (int k = nn - 1; k >= 0; k--) {
        tnode* t = &nd[k];
        double M = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
        if (t->first < 0) {
            for (int q = t->start; q < t->start + t->count; q++) {
                int b = order[q];
                double mb = m[b];
                M += mb;
                sx += mb * px[b]; sy += mb * py[b]; sz += mb * pz[b];
            }
        } else {
            for (int c = t->first; c < t->first + t->nchild; c++) {
                M += nd[c].mass;
                sx += nd[c].sx; sy += nd[c].sy; sz += nd[c].sz;
            }
        }
        t->mass = M; t->sx = sx; t->sy = sy; t->sz = sz;
        if (M > 0.0) { t->cx = sx / M; t->cy = sy / M; t->cz = sz / M; }
        else { t->cx = 0.0; t->cy = 0.0; t->cz = 0.0; }
    }
    *n_active_out = na;
    *keys_out = ks;
    *order_out = order;
    *nodes_out = nd;
    *nn_out = nn;
    return 0;
}

int cmo_tree_build(int n, const double* px, const double* py, const double* pz,
                   const double* m, const uint8_t* act, int levels, int leaf_cap,
                   int node_cap, double* box4, int32_t* n_active,
                   uint64_t* keys_sorted, int32_t* order,
                   int32_t* nd_level, int32_t* nd_start, int32_t* nd_count,
                   int32_t* nd_first, int32_t* nd_nchild, double* nd_mass,
                   double* nd_com, int32_t* level_start) {
    int na, nn;
    uint64_t* ks;
    int32_t* ord;
    tnode* nd;
    tree_build_internal(n, px, py, pz, m, act, levels, leaf_cap, box4, &na,
                        &ks, &ord, &nd, &nn, level_start);
    *n_active = na;
    memcpy(keys_sorted, ks, (size_t)n * 8);
    memcpy(order, ord, (size_t)n * 4);
    if (nn > node_cap) { free(ks); free(ord); free(nd); return -nn; }
    for (int k = 0; k < nn; k++) {
        nd_level[k] = nd[k].level; nd_start[k] = nd[k].start;
        nd_count[k] = nd[k].count; nd_first[k] = nd[k].first;
        nd_nchild[k] = nd[k].nchild; nd_mass[k] = nd[k].mass;
        nd_com[3 * k] = nd[k].cx; nd_com[3 * k + 1] = nd[k].cy;
        nd_com[3 * k + 2] = nd[k].cz;
    }
    free(ks); free(ord); free(nd);
    return nn;
}

/* Group-MAC Barnes-Hut over the level-order octree.
 *
 * Targets are taken in groups of group_size consecutive sorted active bodies.
 * For each group the tree is walked breadth-first (FIFO): a node of mass 0 is
 * skipped (BarnesHutBackend.cs:402); otherwise with d2 the squared distance
 * from its centre of mass to the group's bounding box and hL = h * 2^-level,
 * it is accepted as a monopole when d2 > 0 and (4*hL)*hL < theta^2 * d2 -- the
 * reference's (2h)^2 < theta^2 d^2 test (BarnesHutBackend.cs:431-433) with the
 * body replaced by the nearest point of the group box -- except that leaves
 * are never approximated when the shell theorem is on.  An opened leaf
 * contributes its bodies, an opened internal node its children.  Each body
 * then sums, in list order, the accepted monopoles (BarnesHutBackend.cs:438-445
 * arithmetic) and, separately, the listed bodies other than itself
 * (GravityKernel arithmetic, BarnesHutBackend.cs:415-425); a = monopoles + bodies.
 * theta is clamped to >= 0.2 (BarnesHutBackend.cs:129). */
/* walk statistics of the last cmo_bh_tree calls: sub-groups, walk steps,
 * cell-list flush slots, body-list flush slots (64-entry flushes) */
static int64_t g_walk[4];
void cmo_walk_stats(int64_t* out, int reset) {
    for (int k = 0; k < 4; k++) { out[k] = g_walk[k]; if (reset) g_walk[k] = 0; }
}

static void group_box(const int32_t* ord, const double* px, const double* py,
                      const double* pz, int s, int e, double* lo, double* hi) {
    lo[0] = lo[1] = lo[2] = DBL_MAX;
    hi[0] = hi[1] = hi[2] = -DBL_MAX;
    for (int q = s; q < e; q++) {
        int b = ord[q];
        if (px[b] < lo[0]) lo[0] = px[b];
        if (py[b] < lo[1]) lo[1] = py[b];
        if (pz[b] < lo[2]) lo[2] = pz[b];
        if (px[b] > hi[0]) hi[0] = px[b];
        if (py[b] > hi[1]) hi[1] = py[b];
        if (pz[b] > hi[2]) hi[2] = pz[b];
    }
}

static double box_diag2(const double* lo, const double* hi) {
    double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return dx * dx + dy * dy + dz * dz;
}

static double box_vol(const double* lo, const double* hi, double pad) {
    return ((hi[0] - lo[0]) + pad) * ((hi[1] - lo[1]) + pad) * ((hi[2] - lo[2]) + pad);
}

/* Split the sorted range [s, e) of one target group at its highest differing
 * key bit while the two halves are far apart: the split is taken when the
 * squared diagonal of the whole box exceeds 4x the larger half's, or its
 * volume exceeds 4x the halves' summed volumes (extents padded by 1e-3 of the
 * whole box's summed extents), up to max_depth times.  Emitted sub-ranges are
 * appended to out (pairs). */
static int split_group(const uint64_t* ks, const int32_t* ord, const double* px,
                       const double* py, const double* pz, int s, int e,
                       int depth, int max_depth, int32_t* out, int nout) {
    if (depth < max_depth && e - s >= 2 && ks[s] != ks[e - 1]) {
        uint64_t x = ks[s] ^ ks[e - 1];
        int p = 63 - __builtin_clzll(x);
        int mid = s;
        while (mid < e && !((ks[mid] >> p) & 1u)) mid++;
        double lo[3], hi[3], lo1[3], hi1[3], lo2[3], hi2[3];
        group_box(ord, px, py, pz, s, e, lo, hi);
        group_box(ord, px, py, pz, s, mid, lo1, hi1);
        group_box(ord, px, py, pz, mid, e, lo2, hi2);
        double d = box_diag2(lo, hi), d1 = box_diag2(lo1, hi1), d2 = box_diag2(lo2, hi2);
        double dm = d1 > d2 ? d1 : d2;
        double pad = 1e-3 * ((hi[0] - lo[0]) + (hi[1] - lo[1]) + (hi[2] - lo[2]));
        double v = box_vol(lo, hi, pad), v1 = box_vol(lo1, hi1, pad), v2 = box_vol(lo2, hi2, pad);
        if (d > 4.0 * dm || v > 4.0 * (v1 + v2)) {
            nout = split_group(ks, ord, px, py, pz, s, mid, depth + 1, max_depth, out, nout);
            return split_group(ks, ord, px, py, pz, mid, e, depth + 1, max_depth, out, nout);
        }
    }
    out[2 * nout] = s;
    out[2 * nout + 1] = e;
    return nout + 1;
}

static int cmp_range(const void* a, const void* b) {
    int x = ((const int32_t*)a)[0], y = ((const int32_t*)b)[0];
    return (x > y) - (x < y);
}

/* Partition of the na sorted active bodies into target groups of at most
 * group_size.  policy >= 0: consecutive chunks of group_size, each split by
 * split_group with max depth = policy.  policy == -1: the maximal tree nodes
 * holding at most group_size bodies.  policy == -2: those nodes, with
 * consecutive ones merged (left to right, greedily) while the merged count
 * fits and split_group's rule would not separate them again.
 * Returns a malloc'd array of (start, end) pairs in ascending order. */
/* 1 when the sorted range [s, e) should stay split at mid: the whole box's
 * squared diagonal exceeds 4x the larger part's, or its padded volume
 * exceeds 4x the parts' summed padded volumes. */
static int box_split(const int32_t* ord, const double* px, const double* py,
                     const double* pz, int s, int mid, int e) {
    double lo[3], hi[3], lo1[3], hi1[3], lo2[3], hi2[3];
    group_box(ord, px, py, pz, s, e, lo, hi);
    group_box(ord, px, py, pz, s, mid, lo1, hi1);
    group_box(ord, px, py, pz, mid, e, lo2, hi2);
    double d = box_diag2(lo, hi), d1 = box_diag2(lo1, hi1), d2 = box_diag2(lo2, hi2);
    double dm = d1 > d2 ? d1 : d2;
    double pad = 1e-3 * ((hi[0] - lo[0]) + (hi[1] - lo[1]) + (hi[2] - lo[2]));
    double v = box_vol(lo, hi, pad), v1 = box_vol(lo1, hi1, pad), v2 = box_vol(lo2, hi2, pad);
    return d > 4.0 * dm || v > 4.0 * (v1 + v2);
}

static int common_level(uint64_t a, uint64_t b, int levels) {
    if (a == b) return levels;
    int bits = 3 * levels - (64 - __builtin_clzll(a ^ b));   /* shared top bits */
    return bits / 3;
}

static int32_t* make_groups(int na, const uint64_t* ks, const int32_t* ord,
                            const double* px, const double* py, const double* pz,
                            const tnode* nd, int nn, int group_size, int policy,
                            int levels, int* nr_out) {
    int cap = na / 2 + 64;
    int32_t* r = (int32_t*)malloc((size_t)cap * 2 * 4);
    int nr = 0;
    if (policy >= 0) {
        int32_t sub[2 * 64];
        for (int g0 = 0; g0 < na; g0 += group_size) {
            int g1 = g0 + group_size < na ? g0 + group_size : na;
            int ns = split_group(ks, ord, px, py, pz, g0, g1, 0, policy, sub, 0);
            for (int k = 0; k < ns; k++) {
                if (nr == cap) { cap *= 2; r = (int32_t*)realloc(r, (size_t)cap * 8); }
                r[2 * nr] = sub[2 * k]; r[2 * nr + 1] = sub[2 * k + 1]; nr++;
            }
        }
    } else {
        for (int k = 0; k < nn; k++) {
            const tnode* t = &nd[k];
            if (t->count == 0) continue;
            int take = 0;
            if (k == 0) take = t->count <= group_size;
            if (t->first >= 0 && t->count > group_size) {
                for (int c = t->first; c < t->first + t->nchild; c++) {
                    if (nd[c].count <= group_size) {
                        if (nr == cap) { cap *= 2; r = (int32_t*)realloc(r, (size_t)cap * 8); }
                        r[2 * nr] = nd[c].start; r[2 * nr + 1] = nd[c].start + nd[c].count; nr++;
                    }
                }
            } else if (take) {
                r[2 * nr] = t->start; r[2 * nr + 1] = t->start + t->count; nr++;
            } else if (t->first < 0 && t->count > group_size) {
                for (int q = t->start; q < t->start + t->count; q += group_size) {
                    if (nr == cap) { cap *= 2; r = (int32_t*)realloc(r, (size_t)cap * 8); }
                    int qe = q + group_size < t->start + t->count ? q + group_size
                                                                  : t->start + t->count;
                    r[2 * nr] = q; r[2 * nr + 1] = qe; nr++;
                }
            }
        }
        if (policy == -3) {
            /* sibling packing: for every node with more than group_size
             * bodies, runs of consecutive children that hold at most
             * group_size bodies are packed in child order into bins while the
             * bin stays within group_size and the box rule keeps it whole */
            nr = 0;
            if (nd[0].count <= group_size && nd[0].count > 0) {
                r[0] = 0; r[1] = nd[0].count; nr = 1;
            }
            for (int k = 0; k < nn; k++) {
                const tnode* t = &nd[k];
                if (t->count <= group_size) continue;
                if (t->first < 0) {
                    /* deepest-level leaf with more than group_size bodies:
                     * consecutive slices of group_size */
                    for (int q = t->start; q < t->start + t->count; q += group_size) {
                        if (nr == cap) { cap *= 2; r = (int32_t*)realloc(r, (size_t)cap * 8); }
                        int qe = q + group_size < t->start + t->count ? q + group_size
                                                                      : t->start + t->count;
                        r[2 * nr] = q; r[2 * nr + 1] = qe; nr++;
                    }
                    continue;
                }
                int bs = -1, be = -1;
                for (int c = t->first; c < t->first + t->nchild; c++) {
                    int cs = nd[c].start, ce = nd[c].start + nd[c].count;
                    int close = 0, open = 0;
                    if (nd[c].count > group_size) {
                        close = 1;
                    } else if (bs < 0) {
                        open = 1;
                    } else if (ce - bs <= group_size && !box_split(ord, px, py, pz, bs, cs, ce)) {
                        be = ce;
                    } else {
                        close = 1; open = 1;
                    }
                    if (close && bs >= 0) {
                        if (nr == cap) { cap *= 2; r = (int32_t*)realloc(r, (size_t)cap * 8); }
                        r[2 * nr] = bs; r[2 * nr + 1] = be; nr++;
                        bs = -1;
                    }
                    if (open) { bs = cs; be = ce; }
                }
                if (bs >= 0) {
                    if (nr == cap) { cap *= 2; r = (int32_t*)realloc(r, (size_t)cap * 8); }
                    r[2 * nr] = bs; r[2 * nr + 1] = be; nr++;
                }
            }
        }
        qsort(r, (size_t)nr, 8, cmp_range);
        if (policy == -2 && nr > 0) {
            int w = 0;
            for (int k = 1; k < nr; k++) {
                int s = r[2 * w], mid = r[2 * k], e = r[2 * k + 1];
                int merge = 0;
                /* tightest key cell of a range: common prefix of its end keys */
                int lvl_all = common_level(ks[s], ks[e - 1], levels);
                int lvl_a = common_level(ks[s], ks[mid - 1], levels);
                int lvl_b = common_level(ks[mid], ks[e - 1], levels);
                int lvl_min = lvl_a > lvl_b ? lvl_a : lvl_b;   /* the finer one */
                if (e - s <= group_size && lvl_all >= lvl_min - 1) {
                    double lo[3], hi[3], lo1[3], hi1[3], lo2[3], hi2[3];
                    group_box(ord, px, py, pz, s, e, lo, hi);
                    group_box(ord, px, py, pz, s, mid, lo1, hi1);
                    group_box(ord, px, py, pz, mid, e, lo2, hi2);
                    double d = box_diag2(lo, hi), d1 = box_diag2(lo1, hi1), d2 = box_diag2(lo2, hi2);
                    double dm = d1 > d2 ? d1 : d2;
                    double pad = 1e-3 * ((hi[0] - lo[0]) + (hi[1] - lo[1]) + (hi[2] - lo[2]));
                    double v = box_vol(lo, hi, pad), v1 = box_vol(lo1, hi1, pad), v2 = box_vol(lo2, hi2, pad);
                    merge = !(d > 4.0 * dm || v > 4.0 * (v1 + v2));
                }
                if (merge) r[2 * w + 1] = e;
                else { w++; r[2 * w] = mid; r[2 * w + 1] = e; }
            }
            nr = w + 1;
        }
    }
    *nr_out = nr;
    return r;
}

/* Far-field expansion of a cell about a group centre c (this build's
 * definition; no reference counterpart).  The softened monopole field of
 * mass M at c + r, seen at c + d, is f(d) = M (r - d) / s(r - d)^(3/2) with
 * s(y) = |y|^2 + eps^2; its Taylor coefficients at d = 0 are
 *   T0_a   = M r_a s^-3/2
 *   T1_ab  = 3 M r_a r_b s^-5/2 - delta_ab M s^-3/2
 *   T2_abc = 15 M r_a r_b r_c s^-7/2 - 3 M (delta_ab r_c + delta_ac r_b + delta_bc r_a) s^-5/2
 * (s = |r|^2 + eps^2), and f(d) ~ T0 + T1 d + (1/2) T2 d d. */
void cmo_far_coeffs(double M, double rx, double ry, double rz, double eps2, double* t) {
    double s = rx * rx + ry * ry + rz * rz + eps2;
    double ir = 1.0 / sqrt(s), ir2 = ir * ir;
    double i3 = M * ir * ir2, i5 = i3 * ir2, i7 = i5 * ir2;
    double r[3] = {rx, ry, rz};
    for (int a = 0; a < 3; a++) t[a] += i3 * r[a];
    /* T1: xx, yy, zz, xy, xz, yz */
    static const int p1[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
    for (int k = 0; k < 6; k++) {
        int a = p1[k][0], b = p1[k][1];
        t[3 + k] += 3.0 * i5 * r[a] * r[b] - (a == b ? i3 : 0.0);
    }
    /* T2: xxx, yyy, zzz, xxy, xxz, yyx, yyz, zzx, zzy, xyz */
    static const int p2[10][3] = {{0, 0, 0}, {1, 1, 1}, {2, 2, 2}, {0, 0, 1}, {0, 0, 2},
                                  {1, 1, 0}, {1, 1, 2}, {2, 2, 0}, {2, 2, 1}, {0, 1, 2}};
    for (int k = 0; k < 10; k++) {
        int a = p2[k][0], b = p2[k][1], c = p2[k][2];
        double dl = (a == b ? r[c] : 0.0) + (a == c ? r[b] : 0.0) + (b == c ? r[a] : 0.0);
        t[9 + k] += 15.0 * i7 * r[a] * r[b] * r[c] - 3.0 * i5 * dl;
    }
}

void cmo_far_eval(const double* t, double dx, double dy, double dz, double* out) {
    /* full symmetric tensors from the packed coefficients */
    double T1[3][3] = {{t[3], t[6], t[7]}, {t[6], t[4], t[8]}, {t[7], t[8], t[5]}};
    double T2[3][3][3];
    static const int p2[10][3] = {{0, 0, 0}, {1, 1, 1}, {2, 2, 2}, {0, 0, 1}, {0, 0, 2},
                                  {1, 1, 0}, {1, 1, 2}, {2, 2, 0}, {2, 2, 1}, {0, 1, 2}};
    for (int k = 0; k < 10; k++) {
        int i = p2[k][0], j = p2[k][1], l = p2[k][2];
        T2[i][j][l] = T2[i][l][j] = T2[j][i][l] = T2[j][l][i] = T2[l][i][j] = T2[l][j][i] = t[9 + k];
    }
    double d[3] = {dx, dy, dz};
    for (int a = 0; a < 3; a++) {
        double o = t[a];
        for (int b = 0; b < 3; b++) {
            o += T1[a][b] * d[b];
            for (int c = 0; c < 3; c++) o += 0.5 * T2[a][b][c] * d[b] * d[c];
        }
        out[a] = o;
    }
}

/* The group traversal; with rows != NULL it sums each body's potential
 * instead of its acceleration (see cmo_bh_potential). */
static int bh_tree_core(int n, const double* px, const double* py, const double* pz,
                        const double* m, const double* rad, const uint8_t* act,
                        double eps, double theta, int shell, int levels, int leaf_cap,
                        int group_size, int split_depth, int lifo, int nthreads,
                        double far_ratio, double* ax, double* ay, double* az,
                        int32_t* n_cells, int32_t* n_parts, int32_t* max_queue,
                        int64_t* far_total, double* rows) {
    double box[4];
    int na, nn;
    uint64_t* ks;
    int32_t* ord;
    tnode* nd;
    tree_build_internal(n, px, py, pz, m, act, levels, leaf_cap, box, &na, &ks,
                        &ord, &nd, &nn, NULL);
    for (int i = 0; i < n; i++) {
        if (rows) rows[i] = 0.0;
        else { ax[i] = 0.0; ay[i] = 0.0; az[i] = 0.0; }
        if (n_cells) n_cells[i] = 0;
        if (n_parts) n_parts[i] = 0;
    }
    double th = theta > 0.2 ? theta : 0.2;
    double theta2 = th * th;
    double eps2 = eps * eps;
    double h = box[3];
    int nr = 0;
    int32_t* rng = make_groups(na, ks, ord, px, py, pz, nd, nn, group_size,
                               split_depth, levels, &nr);
    __atomic_fetch_add(&g_walk[0], nr, __ATOMIC_RELAXED);
    {   /* the groups must partition [0, na) */
        int at = 0, ok = 1;
        for (int k = 0; k < nr && ok; k++) {
            if (rng[2 * k] != at || rng[2 * k + 1] <= at || rng[2 * k + 1] - at > group_size) ok = 0;
            at = rng[2 * k + 1];
        }
        if (!ok || at != na) {
            free(ks); free(ord); free(nd); free(rng);
            return -2;
        }
    }
    int mq = 0;
    int64_t nfar_all = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    (void)nthreads;
#pragma omp parallel reduction(max : mq) reduction(+ : nfar_all)
    {
        int qcap = 1024, ccap = 1024, pcap = 1024;
        int32_t* q = (int32_t*)malloc((size_t)qcap * 4);
        int32_t* cl = (int32_t*)malloc((size_t)ccap * 4);
        int32_t* pl = (int32_t*)malloc((size_t)pcap * 4);
#pragma omp for schedule(dynamic, 16)
        for (int r = 0; r < nr; r++) {
          {
            int gs = rng[2 * r], ge = rng[2 * r + 1];
            double glo[3], ghi[3];
            group_box(ord, px, py, pz, gs, ge, glo, ghi);
            double lox = glo[0], loy = glo[1], loz = glo[2];
            double hix = ghi[0], hiy = ghi[1], hiz = ghi[2];
            /* far field: an accepted cell whose centre of mass lies beyond
             * R / far_ratio of the box centre (R the half diagonal) joins the
             * group's Taylor expansion instead of the per-body cell list */
            double gcx = (lox + hix) * 0.5, gcy = (loy + hiy) * 0.5, gcz = (loz + hiz) * 0.5;
            double far_lim = DBL_MAX;
            if (far_ratio > 0.0) {
                double ex = hix - lox, ey = hiy - loy, ez = hiz - loz;
                far_lim = 0.25 * (ex * ex + ey * ey + ez * ez) / (far_ratio * far_ratio);
            }
            double tfar[19] = {0};
            int nfar = 0;
            int qh = 0, qt = 0, nc = 0, np = 0;
            q[qt++] = 0;
            while (qh < qt) {
                if (qt - qh > mq) mq = qt - qh;
                int k;
                if (lifo) {
                    /* chunked LIFO over a stack of lifo entries: the top
                     * `take` entries are taken in push order, one per lane;
                     * take = min(64, S, max(1, (cap - R - S) / 7)) with S the
                     * stack size and R = 8 (levels + 2), so S never exceeds
                     * cap (a popped node pushes at most 8 children, and
                     * single-node steps add at most 7 per level descended) */
                    k = -1;
                    int S = qt - qh;
                    int lim = (lifo - 8 * (levels + 2) - S) / 7;
                    if (lim < 1) lim = 1;
                    int take = S < 64 ? S : 64;
                    __atomic_fetch_add(&g_walk[1], 1, __ATOMIC_RELAXED);
                    if (take > lim) take = lim;
                    /* serialise the chunk: process q[qt-take .. qt-1] in order,
                     * children pushed after the chunk is removed */
                    int base = qt - take;
                    int32_t chunk[64];
                    for (int c = 0; c < take; c++) chunk[c] = q[base + c];
                    qt = base;
                    for (int c = 0; c < take; c++) {
                        int kk = chunk[c];
                        const tnode* t = &nd[kk];
                        if (t->mass == 0.0) continue;
                        double ddx = t->cx < lox ? lox - t->cx : (t->cx > hix ? t->cx - hix : 0.0);
                        double ddy = t->cy < loy ? loy - t->cy : (t->cy > hiy ? t->cy - hiy : 0.0);
                        double ddz = t->cz < loz ? loz - t->cz : (t->cz > hiz ? t->cz - hiz : 0.0);
                        double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                        double hl = ldexp(h, -t->level);
                        double size2 = 4.0 * hl * hl;
                        int leaf = t->first < 0;
                        int far = 0;
                        if (!(shell && leaf) && d2 > 0.0 && size2 < theta2 * d2) {
                            double rx = t->cx - gcx, ry = t->cy - gcy, rz = t->cz - gcz;
                            far = rx * rx + ry * ry + rz * rz > far_lim;
                        }
                        if (far) {
                            cmo_far_coeffs(t->mass, t->cx - gcx, t->cy - gcy, t->cz - gcz, eps2, tfar);
                            nfar++;
                        } else if (!(shell && leaf) && d2 > 0.0 && size2 < theta2 * d2) {
                            if (nc == ccap) { ccap *= 2; cl = (int32_t*)realloc(cl, (size_t)ccap * 4); }
                            cl[nc++] = kk;
                        } else if (leaf) {
                            if (np + t->count > pcap) {
                                while (np + t->count > pcap) pcap *= 2;
                                pl = (int32_t*)realloc(pl, (size_t)pcap * 4);
                            }
                            for (int s2 = t->start; s2 < t->start + t->count; s2++) pl[np++] = s2;
                        } else {
                            while (qt + t->nchild > qcap) { qcap *= 2; q = (int32_t*)realloc(q, (size_t)qcap * 4); }
                            for (int c2 = t->first; c2 < t->first + t->nchild; c2++) q[qt++] = c2;
                        }
                    }
                    continue;
                }
                k = q[qh++];
                const tnode* t = &nd[k];
                if (t->mass == 0.0) continue;
                double ddx = t->cx < lox ? lox - t->cx : (t->cx > hix ? t->cx - hix : 0.0);
                double ddy = t->cy < loy ? loy - t->cy : (t->cy > hiy ? t->cy - hiy : 0.0);
                double ddz = t->cz < loz ? loz - t->cz : (t->cz > hiz ? t->cz - hiz : 0.0);
                double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                double hl = ldexp(h, -t->level);
                double size2 = 4.0 * hl * hl;
                int leaf = t->first < 0;
                int far = 0;
                if (!(shell && leaf) && d2 > 0.0 && size2 < theta2 * d2) {
                    double rx = t->cx - gcx, ry = t->cy - gcy, rz = t->cz - gcz;
                    far = rx * rx + ry * ry + rz * rz > far_lim;
                }
                if (far) {
                    cmo_far_coeffs(t->mass, t->cx - gcx, t->cy - gcy, t->cz - gcz, eps2, tfar);
                    nfar++;
                } else if (!(shell && leaf) && d2 > 0.0 && size2 < theta2 * d2) {
                    if (nc == ccap) { ccap *= 2; cl = (int32_t*)realloc(cl, (size_t)ccap * 4); }
                    cl[nc++] = k;
                } else if (leaf) {
                    if (np + t->count > pcap) {
                        while (np + t->count > pcap) pcap *= 2;
                        pl = (int32_t*)realloc(pl, (size_t)pcap * 4);
                    }
                    for (int s = t->start; s < t->start + t->count; s++) pl[np++] = s;
                } else {
                    if (qt + t->nchild > qcap) {
                        /* compact the FIFO, then grow it */
                        memmove(q, q + qh, (size_t)(qt - qh) * 4);
                        qt -= qh; qh = 0;
                        while (qt + t->nchild > qcap) qcap *= 2;
                        q = (int32_t*)realloc(q, (size_t)qcap * 4);
                    }
                    for (int c = t->first; c < t->first + t->nchild; c++) q[qt++] = c;
                }
            }
            for (int s = gs; s < ge; s++) {
                int i = ord[s];
                double xi = px[i], yi = py[i], zi = pz[i];
                if (rows) {
                    /* potential: phi -= m / sqrt(d^2 + eps^2) over the cells,
                     * then over the bodies (self skipped); the share of the
                     * energy is 1/2 m_i phi_i (potential_energy_kernel.cu:
                     * 29-111, 160-165) */
                    double pc = 0.0, pp = 0.0;
                    for (int c = 0; c < nc; c++) {
                        const tnode* t = &nd[cl[c]];
                        double dx = xi - t->cx, dy = yi - t->cy, dz = zi - t->cz;
                        double dist2 = dx * dx + dy * dy + dz * dz;
                        double dist = sqrt(dist2 + eps2);
                        pc -= t->mass / dist;
                    }
                    int self = 0;
                    for (int c = 0; c < np; c++) {
                        int sj = pl[c];
                        if (sj == s) { self = 1; continue; }
                        int j = ord[sj];
                        double dx = xi - px[j], dy = yi - py[j], dz = zi - pz[j];
                        double dist2 = dx * dx + dy * dy + dz * dz;
                        double dist = sqrt(dist2 + eps2);
                        pp -= m[j] / dist;
                    }
                    rows[i] = 0.5 * (m[i] * (pc + pp));
                    if (n_cells) n_cells[i] = nc;
                    if (n_parts) n_parts[i] = np - self;
                    continue;
                }
                double cxs = 0.0, cys = 0.0, czs = 0.0;
                for (int c = 0; c < nc; c++) {
                    const tnode* t = &nd[cl[c]];
                    double dx = xi - t->cx;
                    double dy = yi - t->cy;
                    double dz = zi - t->cz;
                    double dist2 = dx * dx + dy * dy + dz * dz;
                    double soft = dist2 + eps2;
                    double inv = 1.0 / sqrt(soft);
                    double inv3 = inv * inv * inv;
                    double f = t->mass * inv3;
                    cxs -= f * dx; cys -= f * dy; czs -= f * dz;
                }
                double pxs = 0.0, pys = 0.0, pzs = 0.0;
                int self = 0;
                for (int c = 0; c < np; c++) {
                    int sj = pl[c];
                    if (sj == s) { self = 1; continue; }
                    int j = ord[sj];
                    double dx = xi - px[j];
                    double dy = yi - py[j];
                    double dz = zi - pz[j];
                    double dist2 = dx * dx + dy * dy + dz * dz;
                    double f = cmo_gravity_coeff(dist2, m[j], rad[j], eps2, shell);
                    pxs -= f * dx; pys -= f * dy; pzs -= f * dz;
                }
                if (nfar) {
                    double o[3];
                    cmo_far_eval(tfar, xi - gcx, yi - gcy, zi - gcz, o);
                    cxs += o[0]; cys += o[1]; czs += o[2];
                }
                ax[i] = cxs + pxs; ay[i] = cys + pys; az[i] = czs + pzs;
                if (n_cells) n_cells[i] = nc;
                if (n_parts) n_parts[i] = np - self;
            }
            nfar_all += nfar;
            __atomic_fetch_add(&g_walk[2], (int64_t)((nc + 63) / 64) * 64, __ATOMIC_RELAXED);
            __atomic_fetch_add(&g_walk[3], (int64_t)((np + 63) / 64) * 64, __ATOMIC_RELAXED);
          }
        }
        free(q); free(cl); free(pl);
    }
    if (max_queue) *max_queue = mq;
    if (far_total) *far_total = nfar_all;
    free(ks); free(ord); free(nd); free(rng);
    return nn;
}

int cmo_bh_tree(int n, const double* px, const double* py, const double* pz,
                const double* m, const double* rad, const uint8_t* act,
                double eps, double theta, int shell, int levels, int leaf_cap,
                int group_size, int split_depth, int lifo, int nthreads,
                double far_ratio, double* ax, double* ay, double* az,
                int32_t* n_cells, int32_t* n_parts, int32_t* max_queue,
                int64_t* far_total) {
    return bh_tree_core(n, px, py, pz, m, rad, act, eps, theta, shell, levels, leaf_cap,
                        group_size, split_depth, lifo, nthreads, far_ratio, ax, ay, az,
                        n_cells, n_parts, max_queue, far_total, NULL);
}

/* Barnes-Hut potential through the group traversal (this build's
 * definition of the engine's compute_potential_bh_kernel,
 * engine/src/cuda/potential_energy_kernel.cu:29-111): rows[i] =
 * 1/2 m_i phi_i, phi_i = -sum m / sqrt(d^2 + eps^2) over body i's cells and
 * bodies (no shell theorem, no far field, theta clamped as the force path). */
int cmo_bh_potential(int n, const double* px, const double* py, const double* pz,
                     const double* m, const uint8_t* act, double eps, double theta,
                     int levels, int leaf_cap, int group_size, int split_depth, int lifo,
                     int nthreads, double* rows, int32_t* n_cells, int32_t* n_parts) {
    return bh_tree_core(n, px, py, pz, m, NULL, act, eps, theta, 0, levels, leaf_cap,
                        group_size, split_depth, lifo, nthreads, 0.0, NULL, NULL, NULL,
                        n_cells, n_parts, NULL, NULL, rows);
}

/* compute_potential_bh_kernel (engine/src/cuda/potential_energy_kernel.cu:
 * 29-111) restated over the C# insertion octree: zero-mass nodes skipped,
 * a leaf (one body) gives -m / sqrt(d^2 + eps^2) unless it is the body
 * itself, an internal node with d^2 > 0 and (2h)^2 < theta^2 d^2 its
 * monopole, else its children 0..7 in order; theta is used as given. */
static void cs_potential(const csnode* nodes, int32_t ni, int32_t b, double bx, double by,
                         double bz, double eps2, double theta2, double* phi) {
    const csnode* nd = &nodes[ni];
    if (nd->mass <= 0.0) return;
    double dx = bx - nd->comx, dy = by - nd->comy, dz = bz - nd->comz;
    double dist2 = dx * dx + dy * dy + dz * dz;
    if (nd->leaf) {
        if (nd->body == b) return;
        *phi -= nd->mass / sqrt(dist2 + eps2);
        return;
    }
    double size2 = 4.0 * nd->half * nd->half;
    if (dist2 > 0.0 && size2 < theta2 * dist2) {
        *phi -= nd->mass / sqrt(dist2 + eps2);
        return;
    }
    for (int k = 0; k < 8; k++)
        if (nd->child[k] != -1) cs_potential(nodes, nd->child[k], b, bx, by, bz, eps2, theta2, phi);
}

int cmo_bh_potential_reference(int n, const double* px, const double* py, const double* pz,
                               const double* m, const uint8_t* act, double eps, double theta,
                               int nthreads, double* rows) {
    for (int i = 0; i < n; i++) rows[i] = 0.0;
    if (n == 0) return 0;
    double box[4];
    cmo_root_box(n, px, py, pz, act, box);
    cspool pool;
    pool.cap = 4 * (int64_t)n + 64;
    pool.nodes = (csnode*)malloc((size_t)pool.cap * sizeof(csnode));
    if (!pool.nodes) return -1;
    pool.count = 0;
    pool.min_node_size = 1e-12;
    int32_t root = cs_alloc(&pool, box[0], box[1], box[2], box[3]);
    for (int i = 0; i < n; i++)
        if (act[i]) cs_insert(&pool, root, i, px[i], py[i], pz[i], m[i]);
    double eps2 = eps * eps, theta2 = theta * theta;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    (void)nthreads;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        if (!act[i]) continue;
        double phi = 0.0;
        cs_potential(pool.nodes, root, i, px[i], py[i], pz[i], eps2, theta2, &phi);
        rows[i] = 0.5 * (m[i] * phi);
    }
    free(pool.nodes);
    return 0;
}

/* ======================================================================== */
/* Diagnostics                                                              */
/* ======================================================================== */

/* EnergyCalculator (Validation/EnergyCalculator.cs:16-81) with the
 * NewtonianGravity potential (Forces/NewtonianGravity.cs:68-77): naive FP64
 * sums in index order; PE over i<j with r = pos_j - pos_i. */
void cmo_energy_naive(int n, const double* px, const double* py,
                      const double* pz, const double* vx, const double* vy,
                      const double* vz, const double* m, const uint8_t* act,
                      double eps, int with_pe, cmo_energy_t* out) {
    double ke = 0.0, pe = 0.0;
    double mx = 0.0, my = 0.0, mz = 0.0, lx = 0.0, ly = 0.0, lz = 0.0;
    for (int i = 0; i < n; i++) {
        if (!act[i]) continue;
        ke += 0.5 * m[i] * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
    }
    if (with_pe) {
        double e2 = eps * eps;
        for (int i = 0; i < n; i++) {
            if (!act[i]) continue;
            for (int j = i + 1; j < n; j++) {
                if (!act[j]) continue;
                double rx = px[j] - px[i];
                double ry = py[j] - py[i];
                double rz = pz[j] - pz[i];
                double d2 = rx * rx + ry * ry + rz * rz + e2;
                double d = sqrt(d2);
                pe += -1.0 * m[i] * m[j] / d;
            }
        }
    }
    for (int i = 0; i < n; i++) {
        if (!act[i]) continue;
        double mvx = vx[i] * m[i], mvy = vy[i] * m[i], mvz = vz[i] * m[i];
        mx += mvx; my += mvy; mz += mvz;
        lx += py[i] * mvz - pz[i] * mvy;
        ly += pz[i] * mvx - px[i] * mvz;
        lz += px[i] * mvy - py[i] * mvx;
    }
    out->ke = ke; out->pe = pe;
    out->px = mx; out->py = my; out->pz = mz;
    out->lx = lx; out->ly = ly; out->lz = lz;
    out->l_mag = sqrt(lx * lx + ly * ly + lz * lz);
}

/* The same quantities summed in long double (64-bit significand): the
 * near-exact values that the GPU's compensated reductions are judged against. */
void cmo_energy_exact(int n, const double* px, const double* py,
                      const double* pz, const double* vx, const double* vy,
                      const double* vz, const double* m, const uint8_t* act,
                      double eps, int with_pe, int nthreads, cmo_energy_t* out) {
    long double ke = 0, pe = 0, mx = 0, my = 0, mz = 0, lx = 0, ly = 0, lz = 0;
    for (int i = 0; i < n; i++) {
        if (!act[i]) continue;
        long double mi = m[i];
        ke += 0.5L * mi * ((long double)vx[i] * vx[i] + (long double)vy[i] * vy[i] +
                           (long double)vz[i] * vz[i]);
        long double mvx = mi * vx[i], mvy = mi * vy[i], mvz = mi * vz[i];
        mx += mvx; my += mvy; mz += mvz;
        lx += py[i] * mvz - pz[i] * mvy;
        ly += pz[i] * mvx - px[i] * mvz;
        lz += px[i] * mvy - py[i] * mvx;
    }
    if (with_pe) {
        long double e2 = (long double)eps * eps;
#ifdef _OPENMP
        if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
        (void)nthreads;
#pragma omp parallel
        {
            long double part = 0;
#pragma omp for schedule(dynamic, 64)
            for (int i = 0; i < n; i++) {
                if (!act[i]) continue;
                for (int j = i + 1; j < n; j++) {
                    if (!act[j]) continue;
                    long double rx = (long double)px[j] - px[i];
                    long double ry = (long double)py[j] - py[i];
                    long double rz = (long double)pz[j] - pz[i];
                    long double d = sqrtl(rx * rx + ry * ry + rz * rz + e2);
                    part -= (long double)m[i] * m[j] / d;
                }
            }
#pragma omp critical
            pe += part;
        }
    }
    out->ke = (double)ke; out->pe = (double)pe;
    out->px = (double)mx; out->py = (double)my; out->pz = (double)mz;
    out->lx = (double)lx; out->ly = (double)ly; out->lz = (double)lz;
    out->l_mag = (double)sqrtl(lx * lx + ly * ly + lz * lz);
}

/* Centre of mass and total mass of the active bodies, as the engine's
 * gpu_reduce_com finalises them (engine/src/cuda/gpu_reduction.cu:481-517,
 * per-body terms m x, m y, m z, m at :232-262): sum m r / sum m, or 0 when the
 * total mass is not positive.  Each sum is rounded once to double, then the
 * quotient.  out4 = (cx, cy, cz, total mass). */
void cmo_center_of_mass(int n, const double* px, const double* py, const double* pz,
                        const double* m, const uint8_t* act, double* out4) {
    /* Neumaier-compensated long double sums: a plain long double sum of 1e6
     * equal masses already drifts by ~1e-14 relative, more than the device's
     * compensated double sums. */
    long double s[4] = {0, 0, 0, 0}, c[4] = {0, 0, 0, 0};
    for (int i = 0; i < n; i++) {
        if (!act[i]) continue;
        long double t[4] = {(long double)m[i] * px[i], (long double)m[i] * py[i],
                            (long double)m[i] * pz[i], m[i]};
        for (int k = 0; k < 4; k++) {
            long double u = s[k] + t[k];
            c[k] += fabsl(s[k]) >= fabsl(t[k]) ? (s[k] - u) + t[k] : (t[k] - u) + s[k];
            s[k] = u;
        }
    }
    double M = (double)(s[3] + c[3]);
    out4[3] = M;
    for (int k = 0; k < 3; k++) out4[k] = M > 0.0 ? (double)(s[k] + c[k]) / M : 0.0;
}

/* Engine::compute_max_acceleration's CPU path (engine/src/sim/engine.cpp:
 * 490-499) with the GPU kernel's per-body magnitude (k_reduce_max_accel,
 * engine/src/cuda/gpu_reduction.cu:269-283): max over active bodies of
 * sqrt(ax*ax + ay*ay + az*az), 0 when none is active. */
double cmo_max_acceleration(int n, const double* ax, const double* ay, const double* az,
                            const uint8_t* act) {
    double best = 0.0;
    for (int i = 0; i < n; i++) {
        if (!act[i]) continue;
        double a = sqrt(ax[i] * ax[i] + ay[i] * ay[i] + az[i] * az[i]);
        if (a > best) best = a;
    }
    return best;
}

static int cmp_double(const void* a, const void* b) {
    double x = *(const double*)a, y = *(const double*)b;
    return (x > y) - (x < y);
}

/* ForceErrorAnalyzer.CompareForces statistics (Validation/ForceErrorAnalyzer.cs:100-135):
 * per-body |a_test - a_exact| / |a_exact| over active bodies with
 * |a_exact| >= 1e-30; mean, max, RMS, plus the median the metric uses. */
void cmo_force_error(int n, const uint8_t* act, const double* ex,
                     const double* ey, const double* ez, const double* tx,
                     const double* ty, const double* tz, cmo_ferr_t* out) {
    double* errs = (double*)malloc((size_t)n * sizeof(double) + 8);
    double s = 0.0, s2 = 0.0, mx = 0.0;
    int cnt = 0;
    for (int i = 0; i < n; i++) {
        if (act && !act[i]) continue;
        double mag = sqrt(ex[i] * ex[i] + ey[i] * ey[i] + ez[i] * ez[i]);
        if (mag < 1e-30) continue;
        double dx = tx[i] - ex[i], dy = ty[i] - ey[i], dz = tz[i] - ez[i];
        double e = sqrt(dx * dx + dy * dy + dz * dz) / mag;
        s += e; s2 += e * e;
        if (e > mx) mx = e;
        errs[cnt++] = e;
    }
    out->count = cnt;
    out->mean = cnt ? s / cnt : 0.0;
    out->rms = cnt ? sqrt(s2 / cnt) : 0.0;
    out->max = mx;
    if (cnt) {
        qsort(errs, (size_t)cnt, sizeof(double), cmp_double);
        out->median = (cnt & 1) ? errs[cnt / 2]
                                : 0.5 * (errs[cnt / 2 - 1] + errs[cnt / 2]);
    } else {
        out->median = 0.0;
    }
    free(errs);
}

/* ======================================================================== */
/* Collisions                                                               */
/* ======================================================================== */

/* CollisionDetector.TryAddEvent (Collisions/CollisionDetector.cs:70-135). */
static int try_event(int i, int j, const double* px, const double* py,
                     const double* pz, const double* vx, const double* vy,
                     const double* vz, const double* m, const double* r,
                     cmo_event_t* ev) {
    double dx = px[i] - px[j];
    double dy = py[i] - py[j];
    double dz = pz[i] - pz[j];
    double dist2 = dx * dx + dy * dy + dz * dz;
    double sumr = r[i] + r[j];
    double sumr2 = sumr * sumr;
    if (dist2 >= sumr2) return 0;
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
    ev->a = a; ev->b = b; ev->overlap = overlap; ev->dist = dist;
    ev->nx = nx; ev->ny = ny; ev->nz = nz; ev->rel_normal_speed = rns;
    return 1;
}

static void merge_sort_events(cmo_event_t* e, cmo_event_t* tmp, int lo, int hi) {
    if (hi - lo < 2) return;
    int mid = (lo + hi) / 2;
    merge_sort_events(e, tmp, lo, mid);
    merge_sort_events(e, tmp, mid, hi);
    int i = lo, j = mid, k = lo;
    while (i < mid && j < hi) {
        /* descending overlap; equal overlaps in (i, j) order of the pair */
        int before = e[j].overlap > e[i].overlap;
        if (e[j].overlap == e[i].overlap) {
            int ji = e[j].a < e[j].b ? e[j].a : e[j].b, jj = e[j].a < e[j].b ? e[j].b : e[j].a;
            int ii = e[i].a < e[i].b ? e[i].a : e[i].b, ij = e[i].a < e[i].b ? e[i].b : e[i].a;
            before = ji < ii || (ji == ii && jj < ij);
        }
        if (before) tmp[k++] = e[j++];
        else tmp[k++] = e[i++];
    }
    while (i < mid) tmp[k++] = e[i++];
    while (j < hi) tmp[k++] = e[j++];
    memcpy(e + lo, tmp + lo, (size_t)(hi - lo) * sizeof(cmo_event_t));
}

/* CollisionDetector.Detect (Collisions/CollisionDetector.cs:35-68).  The
 * spatial-hash broad phase (SpatialHashBroadPhase.cs:15-94) uses cells of
 * 2*max radius and scans 27 neighbours, so it yields every pair with
 * d < r_i + r_j: the candidate set after the narrow phase equals the
 * brute-force i<j sweep used here.  Events are in (i, j) order before the
 * depth sort; List.Sort leaves exact-tie order unspecified, this oracle (and
 * the GPU) keep ties in (i, j) order.  Returns the event count (may exceed
 * cap; only cap events are stored). */
int cmo_detect(int n, const double* px, const double* py, const double* pz,
               const double* vx, const double* vy, const double* vz,
               const double* m, const double* rad, const uint8_t* act,
               const uint8_t* col, cmo_event_t* out, int cap) {
    /* every pair i < j is tested (brute force); threads collect the events
     * of interleaved rows, and the sort below fixes the order, so the result
     * does not depend on the thread count */
    int ne = 0, ecap = 256;
    cmo_event_t* ev = (cmo_event_t*)malloc((size_t)ecap * sizeof(cmo_event_t));
#pragma omp parallel
    {
        int lne = 0, lcap = 64;
        cmo_event_t* loc = (cmo_event_t*)malloc((size_t)lcap * sizeof(cmo_event_t));
#pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < n; i++) {
            if (!act[i] || !col[i]) continue;
            for (int j = i + 1; j < n; j++) {
                if (!act[j] || !col[j]) continue;
                cmo_event_t e;
                if (try_event(i, j, px, py, pz, vx, vy, vz, m, rad, &e)) {
                    if (lne == lcap) {
                        lcap *= 2;
                        loc = (cmo_event_t*)realloc(loc, (size_t)lcap * sizeof(cmo_event_t));
                    }
                    loc[lne++] = e;
                }
            }
        }
#pragma omp critical
        {
            if (ne + lne > ecap) {
                while (ne + lne > ecap) ecap *= 2;
                ev = (cmo_event_t*)realloc(ev, (size_t)ecap * sizeof(cmo_event_t));
            }
            if (lne > 0) memcpy(ev + ne, loc, (size_t)lne * sizeof(cmo_event_t));
            ne += lne;
        }
        free(loc);
    }
    cmo_event_t* tmp = (cmo_event_t*)malloc((size_t)(ne + 1) * sizeof(cmo_event_t));
    merge_sort_events(ev, tmp, 0, ne);
    free(tmp);
    int k = ne < cap ? ne : cap;
    if (k > 0) memcpy(out, ev, (size_t)k * sizeof(cmo_event_t));
    free(ev);
    return ne;
}

/* Cube root rounded to the nearest double, by a fixed IEEE sequence:
 * exponent split x = t * 2^(3q) with t in [0.5, 4), a cubic seed refined by
 * four Newton steps y -= (y^3 - t) / (3 y^2), rescaled by 2^q, then the
 * neighbour among {y-ulp, y, y+ulp} whose cube is closest to x, compared
 * with exact (FMA-split) residuals.  Stands in for Math.Cbrt
 * (Types/DensityModel.cs:60-66); agrees with libm cbrt to <= 1 ulp. */
static double cube_residual(double c, double x) {
    /* c^3 - x evaluated as an unevaluated sum of exact products */
    double c2 = c * c;
    double e2 = fma(c, c, -c2);
    double p = c2 * c;
    double ep = fma(c2, c, -p);
    double qv = e2 * c;
    double eq = fma(e2, c, -qv);
    return ((p - x) + ep) + (qv + eq);
}

double cmo_cbrt(double x) {
    if (x == 0.0 || !isfinite(x)) return x;
    int neg = x < 0.0;
    double a = neg ? -x : x;
    int e;
    double f = frexp(a, &e);           /* a = f * 2^e, f in [0.5, 1) */
    int q = e >= 0 ? e / 3 : -((-e + 2) / 3);
    int r = e - 3 * q;                 /* 0, 1 or 2 */
    double t = ldexp(f, r);            /* t in [0.5, 4) */
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
    if (rh < rb) { best = hi; }
    return neg ? -best : best;
}

/* DensityModel.ComputeBodyRadius (Types/DensityModel.cs:54-100): BlackHole
 * (type 7) uses the Schwarzschild radius in AU, others cbrt(3(m/rho)/(4 pi)). */
double cmo_body_radius(double mass, double density, int32_t type) {
    if (type == 7) {
        if (mass <= 0.0) return 0.0;
        double kg = mass * 1.989e30;
        double rm = 2.0 * 6.67430e-11 * kg / (2.998e8 * 2.998e8);
        return rm / 1.496e11;
    }
    if (mass <= 0.0 || density <= 0.0) return 0.0;
    double volume = mass / density;
    return cmo_cbrt(3.0 * volume / (4.0 * 3.14159265358979323846));
}

/* CollisionResolver.Resolve in MergeOnly mode (Collisions/CollisionResolver.cs:58-130)
 * with ResolveMerge -> MergePolicy.Merge (Collisions/MergePolicy.cs:31-80).
 * Telemetry (burst energy, CollisionEnergyModel) does not touch the state in
 * this mode and is not restated.  Returns the number of merges performed. */
int cmo_resolve_merge(int ne, const cmo_event_t* ev, int n, double* px,
                      double* py, double* pz, double* vx, double* vy,
                      double* vz, double* ax, double* ay, double* az,
                      double* oax, double* oay, double* oaz, double* m,
                      double* rad, double* dens, uint8_t* act, uint8_t* col,
                      int32_t* type) {
    (void)n;
    int merged = 0;
    for (int k = 0; k < ne; k++) {
        int a = ev[k].a, b = ev[k].b;
        if (!act[a] || !act[b]) continue;
        double ma = m[a], mb = m[b];
        if (ma <= 0.0 || mb <= 0.0) continue;
        double mt = ma + mb;
        if (mt <= 0.0) { merged++; continue; }
        double inv = 1.0 / mt;
        vx[a] = (ma * vx[a] + mb * vx[b]) * inv;
        vy[a] = (ma * vy[a] + mb * vy[b]) * inv;
        vz[a] = (ma * vz[a] + mb * vz[b]) * inv;
        px[a] = (ma * px[a] + mb * px[b]) * inv;
        py[a] = (ma * py[a] + mb * py[b]) * inv;
        pz[a] = (ma * pz[a] + mb * pz[b]) * inv;
        m[a] = mt;
        dens[a] = (ma * dens[a] + mb * dens[b]) * inv;
        if (type[b] == 7) type[a] = 7;
        rad[a] = cmo_body_radius(mt, dens[a], type[a]);
        act[b] = 0; col[b] = 0;
        ax[b] = 0.0; ay[b] = 0.0; az[b] = 0.0;
        oax[b] = 0.0; oay[b] = 0.0; oaz[b] = 0.0;
        merged++;
    }
    return merged;
}

/* Order-preserving removal of inactive bodies, ParticleSystem::compact
 * (engine/src/physics/particle_system.cpp:156-184), applied to every array
 * passed in.  old_index[k] = pre-compaction index of survivor k. */
int cmo_compact(int n, double* const* f64_arrays, int n_f64,
                uint8_t* const* u8_arrays, int n_u8, int32_t* const* i32_arrays,
                int n_i32, const uint8_t* act, int32_t* old_index) {
    uint8_t* keep = (uint8_t*)malloc((size_t)n + 1);
    memcpy(keep, act, (size_t)n);
    int w = 0;
    for (int r = 0; r < n; r++) {
        if (!keep[r]) continue;
        if (w != r) {
            for (int a = 0; a < n_f64; a++) f64_arrays[a][w] = f64_arrays[a][r];
            for (int a = 0; a < n_u8; a++) u8_arrays[a][w] = u8_arrays[a][r];
            for (int a = 0; a < n_i32; a++) i32_arrays[a][w] = i32_arrays[a][r];
        }
        if (old_index) old_index[w] = r;
        w++;
    }
    free(keep);
    return w;
}
