// engine.hip -- the C ABI of libcmhip.so: handle lifecycle, BodySoA upload /
// download, backend selection, the device-resident Verlet loop and the
// kick-drift / kick-rotate kernels.
//
// Compiled with -ffp-contract=off: the Verlet updates reproduce the
// reference's C# arithmetic bit for bit (SoAVerletIntegrator.cs:74-115).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "cm_internal.h"

namespace cm {

int set_error(Engine& e, int code, const std::string& msg) {
    e.err = msg;
    return code;
}

int finalize_stats(Engine& e) {
    CM_HIP_TRY(e, hipStreamSynchronize(e.stream));
    cm_force_stats s{};
    auto ms = [&](int a, int b) {
        float t = 0.0f;
        if (hipEventElapsedTime(&t, e.ev[a], e.ev[b]) != hipSuccess) t = -1.0f;
        return t;
    };
    if (e.count > 0 && e.cfg.kind == CM_BACKEND_BARNES_HUT) {
        const TreeHeader& h = *e.hdr_host;
        s.cell_interactions = (int64_t)h.n_cell_int;
        s.body_interactions = (int64_t)h.n_body_int;
        s.far_cells = (int64_t)h.n_far_int;
        s.n_active = h.n_active;
        s.n_nodes = h.n_nodes;
        s.n_groups = h.n_groups;
        // any evaluation of the run (not only the last, whose header this
        // is) that outgrew the node pool fails the call
        s.flags = (h.overflow || h.overflow_run) ? 1 : 0;
        if (e.timing) {
            s.ms_tree = ms(0, 3);
            s.ms_sort = ms(1, 2);
            s.ms_force = ms(3, 5);
            s.ms_total = ms(0, 5);
        }
    } else if (e.count > 0) {
        s.n_active = -1;
        s.body_interactions = -1;
        if (e.timing) {
            s.ms_force = ms(0, 5);
            s.ms_total = s.ms_force;
        }
    }
    for (int w = 0; w < e.wn; w++) {
        float a = 0.0f, b = 0.0f;
        if (hipEventElapsedTime(&a, e.wev[3 * w + 1], e.wev[3 * w + 2]) == hipSuccess &&
            hipEventElapsedTime(&b, e.wev[3 * w], e.wev[3 * w + 2]) == hipSuccess) {
            e.acc_ms_force += a;
            e.acc_ms_total += b;
            e.acc_timed++;
        }
    }
    e.wn = 0;
    e.acc_evals += e.evals_since;
    e.evals_since = 0;
    s.acc_interactions = (int64_t)e.hdr_host->acc_int;
    s.acc_far_cells = (int64_t)e.hdr_host->acc_far;
    s.acc_ms_force = e.acc_ms_force;
    s.acc_ms_total = e.acc_ms_total;
    s.acc_evaluations = e.acc_evals;
    s.acc_timed_evaluations = e.acc_timed;
    e.stats = s;
    return CM_OK;
}

__global__ void k_reset_acc(TreeHeader* hdr) {
    hdr->acc_int = 0ull;
    hdr->acc_far = 0ull;
}

__global__ void k_begin_run(TreeHeader* hdr) { hdr->overflow_run = 0; }

int begin_run(Engine& e) {
    e.eval_seq = 0;   // the first force evaluation of every call is timed
    if (e.cfg.kind != CM_BACKEND_BARNES_HUT || e.count == 0) return CM_OK;
    hipLaunchKernelGGL(k_begin_run, dim3(1), dim3(1), 0, e.stream, e.hdr);
    CM_HIP_TRY(e, hipGetLastError());
    return CM_OK;
}

// ---- Verlet kernels --------------------------------------------------------
// SoAVerletIntegrator.Step phases 1+2 (Integrators/SoAVerletIntegrator.cs:74-90):
// v += (0.5 dt) * a_old; x += dt * v, active bodies only.
__global__ void __launch_bounds__(256)
k_kick_drift(int n, double dt, const uint8_t* __restrict__ act,
             double* __restrict__ px, double* __restrict__ py,
             double* __restrict__ pz, double* __restrict__ vx,
             double* __restrict__ vy, double* __restrict__ vz,
             const double* __restrict__ oax, const double* __restrict__ oay,
             const double* __restrict__ oaz) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !act[i]) return;
    double half = 0.5 * dt;
    double x = vx[i] + half * oax[i];
    double y = vy[i] + half * oay[i];
    double z = vz[i] + half * oaz[i];
    vx[i] = x; vy[i] = y; vz[i] = z;
    px[i] = px[i] + dt * x;
    py[i] = py[i] + dt * y;
    pz[i] = pz[i] + dt * z;
}

// Phases 4+5 (:99-115): v += (0.5 dt) * a for active bodies; a_old = a for all.
__global__ void __launch_bounds__(256)
k_kick_rotate(int n, double dt, const uint8_t* __restrict__ act,
              double* __restrict__ vx, double* __restrict__ vy,
              double* __restrict__ vz, const double* __restrict__ ax,
              const double* __restrict__ ay, const double* __restrict__ az,
              double* __restrict__ oax, double* __restrict__ oay,
              double* __restrict__ oaz) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a0 = ax[i], a1 = ay[i], a2 = az[i];
    if (act[i]) {
        double half = 0.5 * dt;
        vx[i] = vx[i] + half * a0;
        vy[i] = vy[i] + half * a1;
        vz[i] = vz[i] + half * a2;
    }
    oax[i] = a0; oay[i] = a1; oaz[i] = a2;
}

// Phases 4+5 of one step fused with phases 1+2 of the next: the same
// operations in the same order on each body, in one pass over the arrays
// (v += (0.5 dt0) a; a_old = a; v += (0.5 dt1) a_old; x += dt1 v).
__global__ void __launch_bounds__(256)
k_kick_rotate_drift(int n, double dt0, double dt1, const uint8_t* __restrict__ act,
                    double* __restrict__ px, double* __restrict__ py,
                    double* __restrict__ pz, double* __restrict__ vx,
                    double* __restrict__ vy, double* __restrict__ vz,
                    const double* __restrict__ ax, const double* __restrict__ ay,
                    const double* __restrict__ az, double* __restrict__ oax,
                    double* __restrict__ oay, double* __restrict__ oaz) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a0 = ax[i], a1 = ay[i], a2 = az[i];
    oax[i] = a0; oay[i] = a1; oaz[i] = a2;
    if (!act[i]) return;
    const double h0 = 0.5 * dt0, h1 = 0.5 * dt1;
    const double x = (vx[i] + h0 * a0) + h1 * a0;
    const double y = (vy[i] + h0 * a1) + h1 * a1;
    const double z = (vz[i] + h0 * a2) + h1 * a2;
    vx[i] = x; vy[i] = y; vz[i] = z;
    px[i] = px[i] + dt1 * x;
    py[i] = py[i] + dt1 * y;
    pz[i] = pz[i] + dt1 * z;
}

int launch_kick_rotate_drift(Engine& e, double dt0, double dt1) {
    if (e.count == 0) return CM_OK;
    int nb = (e.count + 255) / 256;
    hipLaunchKernelGGL(k_kick_rotate_drift, dim3(nb), dim3(256), 0, e.stream, e.count,
                       dt0, dt1, e.act, e.f[PX], e.f[PY], e.f[PZ], e.f[VX], e.f[VY],
                       e.f[VZ], e.f[AX], e.f[AY], e.f[AZ], e.f[OAX], e.f[OAY], e.f[OAZ]);
    CM_HIP_TRY(e, hipGetLastError());
    return CM_OK;
}

int launch_kick_drift(Engine& e, double dt) {
    if (e.count == 0) return CM_OK;
    int nb = (e.count + 255) / 256;
    hipLaunchKernelGGL(k_kick_drift, dim3(nb), dim3(256), 0, e.stream, e.count,
                       dt, e.act, e.f[PX], e.f[PY], e.f[PZ], e.f[VX], e.f[VY],
                       e.f[VZ], e.f[OAX], e.f[OAY], e.f[OAZ]);
    CM_HIP_TRY(e, hipGetLastError());
    return CM_OK;
}

int launch_kick_rotate(Engine& e, double dt) {
    if (e.count == 0) return CM_OK;
    int nb = (e.count + 255) / 256;
    hipLaunchKernelGGL(k_kick_rotate, dim3(nb), dim3(256), 0, e.stream, e.count,
                       dt, e.act, e.f[VX], e.f[VY], e.f[VZ], e.f[AX], e.f[AY],
                       e.f[AZ], e.f[OAX], e.f[OAY], e.f[OAZ]);
    CM_HIP_TRY(e, hipGetLastError());
    return CM_OK;
}

// The configured backend's Newtonian forces, then the 1PN decorator when it
// is on (PostNewtonianBackend.ComputeForces, Solvers/PostNewtonianBackend.cs:33-43).
static int compute_forces(Engine& e, double softening) {
    int rc = e.cfg.kind == CM_BACKEND_BARNES_HUT ? launch_barnes_hut(e, softening)
                                                 : launch_direct(e, softening);
    if (rc != CM_OK) return rc;
    return launch_pn(e, softening);
}

// A run of SoAVerletIntegrator.Step calls (Integrators/SoAVerletIntegrator.cs:
// 48-115) with step sizes dts[0..n-1]: the closing kick of each step and the
// opening kick-drift of the next share one pass over the bodies.
static int verlet_run(Engine& e, double softening, const double* dts, int n) {
    if (n <= 0) return CM_OK;
    int rc = launch_kick_drift(e, dts[0]);
    for (int k = 0; k < n && rc == CM_OK; k++) {
        e.stats_eval = k + 1 == n;
        rc = compute_forces(e, softening);
        e.stats_eval = true;
        if (rc != CM_OK) break;
        rc = k + 1 < n ? launch_kick_rotate_drift(e, dts[k], dts[k + 1])
                       : launch_kick_rotate(e, dts[k]);
    }
    return rc;
}

static int verlet_one(Engine& e, double softening, double dt) {
    return verlet_run(e, softening, &dt, 1);
}

static double* const* field_ptrs(const cm_soa_ptrs* p, double* out[NF64]) {
    double* const src[NF64] = {p->pos_x, p->pos_y, p->pos_z, p->vel_x,
                               p->vel_y, p->vel_z, p->acc_x, p->acc_y,
                               p->acc_z, p->old_acc_x, p->old_acc_y,
                               p->old_acc_z, p->mass, p->radius, p->density};
    for (int k = 0; k < NF64; k++) out[k] = src[k];
    return out;
}

static void free_all(Engine& e) {
    for (int k = 0; k < NF64; k++) if (e.f[k]) (void)hipFree(e.f[k]);
    void* ptrs[] = {e.act, e.col, e.type, e.keys_a, e.keys_b, e.vals_a,
                    e.vals_b, e.hist, e.sorted, e.bbox_part, e.bbox_cnt, e.hdr,
                    e.nd_info, e.nd_cm, e.nd_sum, e.nd_parent, e.tree_dl, e.tree_tcount, e.pm4, e.nd_quad,
                    e.red, e.sample,
                    e.groups, e.gq, e.bcost, e.gcost, e.gwork, e.nd_box, e.gmark, e.gtile};
    for (void* p : ptrs) if (p) (void)hipFree(p);
    if (e.hdr_host) (void)hipHostFree(e.hdr_host);
    collide_free(e);
    mlp_free(e);
    pn_free(e);
    for (auto& ev : e.ev) if (ev) (void)hipEventDestroy(ev);
    if (e.wev) {
        for (int k = 0; k < 3 * kWindow; k++) if (e.wev[k]) (void)hipEventDestroy(e.wev[k]);
        delete[] e.wev;
    }
    if (e.stream) (void)hipStreamDestroy(e.stream);
}

}  // namespace cm

using cm::Engine;

static Engine* as_engine(void* h) { return static_cast<Engine*>(h); }

extern "C" {

int32_t cm_default_config(cm_backend_config* cfg) {
    if (!cfg) return CM_ERR_INVALID_ARGUMENT;
    cfg->kind = CM_BACKEND_BARNES_HUT;
    cfg->precision = CM_PREC_FP64;
    cfg->shell_theorem = 0;
    cfg->morton_levels = 10;
    cfg->theta = 0.5;   // BarnesHutBackend.Theta default (BarnesHutBackend.cs:75)
    cfg->leaf_capacity = 16;
    cfg->group_size = 64;
    cfg->learned_correction = 0;
    cfg->far_field_ratio = 0.0;
    return CM_OK;
}

int32_t cm_create(int32_t capacity, int32_t device, void** out_engine) {
    if (!out_engine || capacity <= 0) return CM_ERR_INVALID_ARGUMENT;
    *out_engine = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return CM_ERR_DEVICE_NOT_FOUND;
    if (device < 0 || device >= ndev) return CM_ERR_DEVICE_NOT_FOUND;
    Engine* e = new (std::nothrow) Engine();
    if (!e) return CM_ERR_OUT_OF_MEMORY;
    e->device = device;
    e->capacity = capacity;
    cm_default_config(&e->cfg);
    cm_default_pn_config(&e->pn);
    cm_default_softening_config(&e->soft);
    cm_default_collision_config(&e->coll_cfg);
    auto fail = [&](int code) {
        cm::free_all(*e);
        delete e;
        return code;
    };
    if (hipSetDevice(device) != hipSuccess) return fail(CM_ERR_HIP);
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(CM_ERR_HIP);
    size_t n = (size_t)capacity;
    bool ok = true;
    auto alloc = [&](void** p, size_t bytes) {
        if (!ok) return;
        if (hipMalloc(p, std::max<size_t>(bytes, 256)) != hipSuccess) ok = false;
    };
    for (int k = 0; k < cm::NF64; k++) alloc((void**)&e->f[k], n * 8);
    alloc((void**)&e->act, n);
    alloc((void**)&e->col, n);
    alloc((void**)&e->type, n * 4);
    alloc((void**)&e->keys_a, n * 8);
    alloc((void**)&e->keys_b, n * 8);
    alloc((void**)&e->vals_a, n * 4);
    alloc((void**)&e->vals_b, n * 4);
    size_t tiles = (n + cm::kSortTile - 1) / cm::kSortTile;
    alloc((void**)&e->hist, 256 * (tiles + 1) * 4);
    alloc((void**)&e->sorted, 5 * n * 8);
    alloc((void**)&e->pm4, n * sizeof(double4));
    alloc((void**)&e->bbox_part, 6 * 1024 * 8);
    alloc((void**)&e->bbox_cnt, 1024 * 4);
    alloc((void**)&e->hdr, sizeof(cm::TreeHeader));
    e->node_cap = (int)std::min<size_t>(3 * n + 1024, (size_t)INT32_MAX / 2);
    alloc((void**)&e->nd_info, (size_t)e->node_cap * sizeof(int4));
    alloc((void**)&e->nd_cm, (size_t)e->node_cap * sizeof(double4));
    alloc((void**)&e->nd_sum, (size_t)e->node_cap * 3 * 8);
    alloc((void**)&e->nd_box, (size_t)e->node_cap * 6 * 8);
    alloc((void**)&e->nd_parent, (size_t)e->node_cap * 4);
    alloc((void**)&e->nd_quad, (size_t)e->node_cap * 6 * 4);
    alloc((void**)&e->tree_dl, n * sizeof(uchar2));
    alloc((void**)&e->tree_tcount, (size_t)(cm::kMaxLevels + 1) * cm::tree_tiles((int)n) * 4);
    alloc((void**)&e->groups, n * sizeof(int2));
    // queue entries: every group once, plus the parts of split groups
    e->gq_cap = (int)std::min<size_t>(2 * n, (size_t)INT32_MAX / 2);
    alloc((void**)&e->gq, (size_t)e->gq_cap * sizeof(int4));
    alloc((void**)&e->bcost, n * 4);
    alloc((void**)&e->gcost, n * 4);
    alloc((void**)&e->gwork, ((n + 1023) / 1024 + 1) * 8);
    alloc((void**)&e->gmark, n * 4);
    alloc((void**)&e->gtile, ((n + 1023) / 1024 + 1) * 4);
    alloc((void**)&e->red, 64 * 1024 * 8);
    alloc((void**)&e->sample, 8);
    if (ok && hipHostMalloc((void**)&e->hdr_host, sizeof(cm::TreeHeader)) != hipSuccess)
        ok = false;
    if (!ok) return fail(CM_ERR_OUT_OF_MEMORY);
    std::memset(e->hdr_host, 0, sizeof(cm::TreeHeader));
    for (auto& ev : e->ev)
        if (hipEventCreate(&ev) != hipSuccess) return fail(CM_ERR_HIP);
    e->wev = new (std::nothrow) hipEvent_t[3 * cm::kWindow]();
    if (!e->wev) return fail(CM_ERR_OUT_OF_MEMORY);
    for (int k = 0; k < 3 * cm::kWindow; k++)
        if (hipEventCreate(&e->wev[k]) != hipSuccess) return fail(CM_ERR_HIP);
    if (hipMemsetAsync(e->hdr, 0, sizeof(cm::TreeHeader), e->stream) != hipSuccess ||
        hipMemsetAsync(e->bcost, 0, n * 4, e->stream) != hipSuccess ||
        hipStreamSynchronize(e->stream) != hipSuccess)
        return fail(CM_ERR_HIP);
    // device address of cm_energy_noalloc's static scratch on this GPU
    if (int rc = cm::noalloc_scratch_init(*e); rc != CM_OK) return fail(rc);
    *out_engine = e;
    return CM_OK;
}

int32_t cm_destroy(void* h) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->stream);
    cm::free_all(*e);
    delete e;
    return CM_OK;
}

const char* cm_last_error(void* h) {
    Engine* e = as_engine(h);
    if (!e) return "null engine handle";
    return e->err.c_str();
}

int32_t cm_configure(void* h, const cm_backend_config* cfg) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!cfg) return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "null config");
    if (cfg->kind < 0 || cfg->kind > 2)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "unknown backend kind");
    if (cfg->precision < 0 || cfg->precision > 2)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "unknown precision");
    if (cfg->morton_levels < 1 || cfg->morton_levels > cm::kMaxLevels)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT,
                             "morton_levels must be in 1..21");
    if (cfg->leaf_capacity < 1 || cfg->leaf_capacity > 64)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT,
                             "leaf_capacity must be in 1..64");
    if (cfg->group_size < 1 || cfg->group_size > 64)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT,
                             "group_size must be in 1..64");
    if (!(cfg->theta >= 0.0) || !(cfg->theta < 1e6))
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "theta must be finite, >= 0");
    if (cfg->learned_correction != 0 && cfg->learned_correction != 1)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "learned_correction must be 0 or 1");
    if (!(cfg->far_field_ratio >= 0.0) || cfg->far_field_ratio > 0.5)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "far_field_ratio must be in [0, 0.5]");
    e->cfg = *cfg;
    return CM_OK;
}

int32_t cm_upload_soa(void* h, const cm_soa_ptrs* src, int32_t count) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!src || count < 0)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "null source or negative count");
    if (count > e->capacity)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT,
                             "body count exceeds capacity");   // engine.cpp:170-173
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    double* p[cm::NF64];
    cm::field_ptrs(src, p);
    size_t b8 = (size_t)count * 8;
    for (int k = 0; k < cm::NF64; k++) {
        const double* from = p[k];
        if (!from && k >= cm::OAX && k <= cm::OAZ) from = p[k - 3];  // CopyFrom seeds OldAcc = Acc
        if (from)
            CM_HIP_TRY(*e, hipMemcpyAsync(e->f[k], from, b8, hipMemcpyHostToDevice, e->stream));
        else
            CM_HIP_TRY(*e, hipMemsetAsync(e->f[k], 0, b8, e->stream));
    }
    if (src->is_active)
        CM_HIP_TRY(*e, hipMemcpyAsync(e->act, src->is_active, count, hipMemcpyHostToDevice, e->stream));
    else
        CM_HIP_TRY(*e, hipMemsetAsync(e->act, 0, count, e->stream));
    if (src->is_collidable)
        CM_HIP_TRY(*e, hipMemcpyAsync(e->col, src->is_collidable, count, hipMemcpyHostToDevice, e->stream));
    else
        CM_HIP_TRY(*e, hipMemsetAsync(e->col, 0, count, e->stream));
    if (src->body_type)
        CM_HIP_TRY(*e, hipMemcpyAsync(e->type, src->body_type, (size_t)count * 4, hipMemcpyHostToDevice, e->stream));
    else
        CM_HIP_TRY(*e, hipMemsetAsync(e->type, 0, (size_t)count * 4, e->stream));
    CM_HIP_TRY(*e, hipStreamSynchronize(e->stream));
    e->count = count;
    e->sim_time = 0.0;
    e->initial_energy = 0.0;
    e->initial_energy_set = false;
    return CM_OK;
}

int32_t cm_upload_force_inputs(void* h, const cm_soa_ptrs* src, int32_t count) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!src || count < 0)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "null source or negative count");
    if (count > e->capacity)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "body count exceeds capacity");
    const bool need_vel = e->pn.enabled != 0;
    const bool need_type = e->soft.mode == CM_SOFTENING_PER_TYPE;
    const bool need_rad = e->cfg.shell_theorem != 0;
    if (!src->pos_x || !src->pos_y || !src->pos_z || !src->mass || !src->is_active ||
        (need_rad && !src->radius) ||
        (need_vel && (!src->vel_x || !src->vel_y || !src->vel_z)) ||
        (need_type && !src->body_type))
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT,
                             "missing an array the force evaluation reads");
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    const size_t b8 = (size_t)count * 8;
    const double* from[cm::NF64] = {src->pos_x, src->pos_y, src->pos_z, src->vel_x,
                                    src->vel_y, src->vel_z, nullptr, nullptr, nullptr,
                                    nullptr, nullptr, nullptr, src->mass, src->radius,
                                    nullptr};
    for (int k = 0; k < cm::NF64; k++) {
        if (!from[k]) continue;
        if (k >= cm::VX && k <= cm::VZ && !need_vel) continue;
        CM_HIP_TRY(*e, hipMemcpyAsync(e->f[k], from[k], b8, hipMemcpyHostToDevice, e->stream));
    }
    CM_HIP_TRY(*e, hipMemcpyAsync(e->act, src->is_active, count, hipMemcpyHostToDevice, e->stream));
    if (need_type)
        CM_HIP_TRY(*e, hipMemcpyAsync(e->type, src->body_type, (size_t)count * 4,
                                      hipMemcpyHostToDevice, e->stream));
    CM_HIP_TRY(*e, hipStreamSynchronize(e->stream));
    if (count != e->count) {
        e->sim_time = 0.0;
        e->initial_energy_set = false;
    }
    e->count = count;
    return CM_OK;
}

int32_t cm_download_soa(void* h, const cm_soa_ptrs* dst, int32_t count) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!dst || count < 0 || count > e->count)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "bad destination or count");
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    double* p[cm::NF64];
    cm::field_ptrs(dst, p);
    size_t b8 = (size_t)count * 8;
    for (int k = 0; k < cm::NF64; k++)
        if (p[k])
            CM_HIP_TRY(*e, hipMemcpyAsync(p[k], e->f[k], b8, hipMemcpyDeviceToHost, e->stream));
    if (dst->is_active)
        CM_HIP_TRY(*e, hipMemcpyAsync(dst->is_active, e->act, count, hipMemcpyDeviceToHost, e->stream));
    if (dst->is_collidable)
        CM_HIP_TRY(*e, hipMemcpyAsync(dst->is_collidable, e->col, count, hipMemcpyDeviceToHost, e->stream));
    if (dst->body_type)
        CM_HIP_TRY(*e, hipMemcpyAsync(dst->body_type, e->type, (size_t)count * 4, hipMemcpyDeviceToHost, e->stream));
    CM_HIP_TRY(*e, hipStreamSynchronize(e->stream));
    return CM_OK;
}

int32_t cm_get_count(void* h, int32_t* count) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!count) return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "null count");
    *count = e->count;
    return CM_OK;
}

int32_t cm_compute_forces(void* h, double softening) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!(softening >= 0.0))
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "softening must be >= 0");
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    int rc = cm::begin_run(*e);
    if (rc == CM_OK) rc = cm::compute_forces(*e, softening);
    if (rc != CM_OK) return rc;
    rc = cm::finalize_stats(*e);
    if (rc != CM_OK) return rc;
    if (e->stats.flags & 1)
        return cm::set_error(*e, CM_ERR_POOL_EXHAUSTED, "octree node pool exhausted");
    return CM_OK;
}

int32_t cm_verlet_step(void* h, double softening, double dt, int32_t n_steps) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!(softening >= 0.0) || n_steps < 0)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "bad softening or step count");
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    constexpr int kChunk = 256;   // steps per host loop iteration
    double dts[kChunk];
    for (int k = 0; k < kChunk; k++) dts[k] = dt;
    if (int rc0 = cm::begin_run(*e); rc0 != CM_OK) return rc0;
    for (int s = 0; s < n_steps; s += kChunk) {
        int rc = cm::verlet_run(*e, softening, dts, std::min(kChunk, n_steps - s));
        if (rc != CM_OK) return rc;
    }
    int rc = cm::finalize_stats(*e);
    if (rc != CM_OK) return rc;
    if (e->stats.flags & 1)
        return cm::set_error(*e, CM_ERR_POOL_EXHAUSTED, "octree node pool exhausted");
    return CM_OK;
}

int32_t cm_yoshida4_step(void* h, double softening, double dt, int32_t n_steps) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!(softening >= 0.0) || n_steps < 0)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "bad softening or step count");
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    // YoshidaCoefficients (engine/include/celestial/sim/yoshida.hpp:18-28)
    const double cbrt2 = 1.2599210498948732;
    const double w1 = 1.0 / (2.0 - cbrt2);
    const double w0 = -cbrt2 / (2.0 - cbrt2);
    const double sub[3] = {w1 * dt, w0 * dt, w1 * dt};
    constexpr int kChunk = 255;   // sub-steps per host loop iteration (85 steps)
    double dts[kChunk];
    for (int k = 0; k < kChunk; k++) dts[k] = sub[k % 3];
    if (int rc0 = cm::begin_run(*e); rc0 != CM_OK) return rc0;
    for (int s = 0; s < n_steps; s += kChunk / 3) {
        int rc = cm::verlet_run(*e, softening, dts, 3 * std::min(kChunk / 3, n_steps - s));
        if (rc != CM_OK) return rc;
    }
    int rc = cm::finalize_stats(*e);
    if (rc != CM_OK) return rc;
    if (e->stats.flags & 1)
        return cm::set_error(*e, CM_ERR_POOL_EXHAUSTED, "octree node pool exhausted");
    return CM_OK;
}

int32_t cm_default_softening_config(cm_softening_config* cfg) {
    if (!cfg) return CM_ERR_INVALID_ARGUMENT;
    // SofteningConfig defaults (engine/include/celestial/sim/simulation_config.hpp:20-33)
    cfg->mode = CM_SOFTENING_GLOBAL;
    for (int k = 0; k < 16; k++) cfg->type_softening[k] = 0.0;
    cfg->adaptive_scale = 0.01;
    cfg->adaptive_min = 1e-8;
    return CM_OK;
}

int32_t cm_set_softening(void* h, const cm_softening_config* cfg) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!cfg) return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "null softening config");
    if (cfg->mode < CM_SOFTENING_GLOBAL || cfg->mode > CM_SOFTENING_ADAPTIVE)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "unknown softening mode");
    for (int k = 0; k < 16; k++)
        if (!(cfg->type_softening[k] >= 0.0))
            return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "type softening must be >= 0");
    if (!(cfg->adaptive_scale >= 0.0) || !(cfg->adaptive_min >= 0.0))
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT,
                             "adaptive scale and minimum must be >= 0");
    e->soft = *cfg;
    return CM_OK;
}

int32_t cm_energy(void* h, double softening, int32_t with_potential,
                  cm_energy_report* out) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!out) return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "null output");
    if (with_potential < 0 || with_potential > 2)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "with_potential must be 0, 1 or 2");
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    return cm::launch_energy(*e, softening, with_potential, out);
}

int32_t cm_energy_noalloc(void* h, double softening, int32_t with_potential,
                          cm_energy_report* out) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!out) return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "null output");
    if (with_potential < 0 || with_potential > 2)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "with_potential must be 0, 1 or 2");
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    return cm::launch_energy_noalloc(*e, softening, with_potential, out);
}

int32_t cm_potential_rows(void* h, double softening, int32_t method, double* rows) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (method != 1 && method != 2)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "method must be 1 or 2");
    if (!rows && e->count > 0) return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "null rows");
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    return cm::potential_rows(*e, softening, method, rows);
}

int32_t cm_max_acceleration(void* h, double* out) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!out) return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "null output");
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    return cm::max_acceleration(*e, out);
}

int32_t cm_direct_sample(void* h, double softening, const int32_t* targets,
                         int32_t n_targets, double* ax, double* ay, double* az) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (n_targets < 0 || (n_targets > 0 && (!targets || !ax || !ay || !az)))
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "bad sample arguments");
    for (int t = 0; t < n_targets; t++)
        if (targets[t] < 0 || targets[t] >= e->count)
            return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "target index out of range");
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    return cm::launch_direct_sample(*e, softening, targets, n_targets, ax, ay, az);
}

int32_t cm_compare_forces(void* h, double softening, double theta, const int32_t* targets,
                          int32_t n_targets, cm_force_comparison* out) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!out || n_targets < 0 || (n_targets > 0 && !targets))
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "bad comparison arguments");
    if (!(softening >= 0.0) || !(theta >= 0.0) || !(theta < 1e6))
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "bad softening or theta");
    for (int t = 0; t < n_targets; t++)
        if (targets[t] < 0 || targets[t] >= e->count)
            return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "target index out of range");
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    return cm::compare_forces(*e, softening, theta, targets, n_targets, out);
}

int32_t cm_get_force_stats(void* h, cm_force_stats* out) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!out) return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "null output");
    *out = e->stats;
    return CM_OK;
}

int32_t cm_reset_force_stats(void* h) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    int rc = cm::finalize_stats(*e);
    if (rc != CM_OK) return rc;
    hipLaunchKernelGGL(cm::k_reset_acc, dim3(1), dim3(1), 0, e->stream, e->hdr);
    CM_HIP_TRY(*e, hipGetLastError());
    CM_HIP_TRY(*e, hipStreamSynchronize(e->stream));
    e->hdr_host->acc_int = 0ull;
    e->hdr_host->acc_far = 0ull;
    e->acc_ms_force = 0.0;
    e->acc_ms_total = 0.0;
    e->acc_evals = 0;
    e->acc_timed = 0;
    e->stats.acc_interactions = 0;
    e->stats.acc_far_cells = 0;
    e->stats.acc_ms_force = 0.0;
    e->stats.acc_ms_total = 0.0;
    e->stats.acc_evaluations = 0;
    e->stats.acc_timed_evaluations = 0;
    return CM_OK;
}

int32_t cm_detect_collisions(void* h, cm_collision_event* out, int32_t cap,
                             int32_t* n_events) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!n_events || cap < 0)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "bad event arguments");
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    return cm::collide_detect(*e, out, cap, n_events);
}

int32_t cm_resolve_merges(void* h, int32_t* n_merged) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    cm_collision_config merge_only;
    cm_default_collision_config(&merge_only);
    return cm::collide_resolve(*e, merge_only, n_merged);
}

int32_t cm_nbody_step(void* h, double softening, double dt, int32_t n_steps,
                      int32_t collisions, int32_t* n_resolved) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!(softening >= 0.0) || n_steps < 0 || (collisions != 0 && collisions != 1))
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "bad softening, step count or flag");
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    int32_t total = 0;
    if (int rc0 = cm::begin_run(*e); rc0 != CM_OK) return rc0;
    for (int s = 0; s < n_steps; s++) {
        int rc = cm::verlet_run(*e, softening, &dt, 1);
        if (rc == CM_OK && collisions) {
            int32_t ne = 0, k = 0;
            rc = cm::collide_detect(*e, nullptr, 0, &ne);
            // NBodySolver resolves once per step even with no contacts
            if (rc == CM_OK) rc = cm::collide_resolve(*e, e->coll_cfg, &k);
            total += k;
        }
        if (rc != CM_OK) return rc;
    }
    if (n_resolved) *n_resolved = total;
    int rc = cm::finalize_stats(*e);
    if (rc != CM_OK) return rc;
    if (e->stats.flags & 1)
        return cm::set_error(*e, CM_ERR_POOL_EXHAUSTED, "octree node pool exhausted");
    return CM_OK;
}

int32_t cm_nbody_step_diag(void* h, double softening, double dt, int32_t n_steps,
                           int32_t collisions, int32_t with_potential,
                           cm_simulation_state* states) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!(softening >= 0.0) || n_steps < 0 || (collisions != 0 && collisions != 1) ||
        with_potential < 0 || with_potential > 2 || (n_steps > 0 && !states))
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT,
                             "bad softening, step count, flag, potential mode or output");
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    if (int rc0 = cm::begin_run(*e); rc0 != CM_OK) return rc0;
    int rc = cm::nbody_step_diag(*e, softening, dt, n_steps, collisions, with_potential,
                                 cm::verlet_one, states);
    if (rc != CM_OK) return rc;
    rc = cm::finalize_stats(*e);
    if (rc != CM_OK) return rc;
    if (e->stats.flags & 1)
        return cm::set_error(*e, CM_ERR_POOL_EXHAUSTED, "octree node pool exhausted");
    return CM_OK;
}

int32_t cm_default_collision_config(cm_collision_config* cfg) {
    if (!cfg) return CM_ERR_INVALID_ARGUMENT;
    // CollisionResolver's field defaults (Collisions/CollisionResolver.cs:28-32)
    cfg->mode = CM_COLLISION_MERGE_ONLY;
    cfg->default_restitution = 0.15;
    cfg->fragmentation_specific_energy_threshold = 0.6;
    cfg->fragmentation_mass_loss_cap = 0.3;
    cfg->capture_velocity_factor = 0.9;
    cfg->promote_compact_remnants = 0;
    cfg->max_merges_per_frame = 0;
    cfg->max_merges_per_body = 0;
    return CM_OK;
}

static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

int32_t cm_configure_collisions(void* h, const cm_collision_config* cfg) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!cfg) return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "null collision config");
    if (cfg->mode < CM_COLLISION_MERGE_ONLY || cfg->mode > CM_COLLISION_REALISTIC)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "unknown collision mode");
    if (!(cfg->default_restitution == cfg->default_restitution) ||
        !(cfg->fragmentation_mass_loss_cap == cfg->fragmentation_mass_loss_cap))
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "NaN collision parameter");
    cm_collision_config c = *cfg;
    // CollisionResolver.Configure (Collisions/CollisionResolver.cs:38-49)
    c.default_restitution = clampd(c.default_restitution, 0.0, 1.0);
    c.fragmentation_specific_energy_threshold =
        std::max(1e-8, c.fragmentation_specific_energy_threshold);
    c.fragmentation_mass_loss_cap = clampd(c.fragmentation_mass_loss_cap, 0.0, 0.9);
    c.capture_velocity_factor = std::max(0.0, c.capture_velocity_factor);
    if (c.fragmentation_mass_loss_cap < 0.05)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT,
                             "fragmentation_mass_loss_cap below 0.05 (the reference's "
                             "Math.Clamp throws on fragmentation-class events)");
    if (c.promote_compact_remnants != 0 && c.promote_compact_remnants != 1)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "promote_compact_remnants must be 0 or 1");
    if ((c.max_merges_per_frame > 0 || c.max_merges_per_body > 0) &&
        c.mode != CM_COLLISION_MERGE_ONLY)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT,
                             "merge caps apply to the MergeOnly mode only");
    e->coll_cfg = c;
    return CM_OK;
}

int32_t cm_resolve_collisions(void* h, int32_t* n_resolved) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    return cm::collide_resolve(*e, e->coll_cfg, n_resolved);
}

int32_t cm_compact(void* h, int32_t* new_count, int32_t* old_index) {
    Engine* e = as_engine(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    return cm::collide_compact(*e, new_count, old_index);
}

}  // extern "C"
