// cm_internal.h -- device state and kernel entry points of libcmhip.so.
//
// Everything here is MI355X/gfx950 code: 64-lane wavefronts (__ballot returns
// a 64-bit mask), HIP streams of default priority, hipMalloc'd SoA arrays.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/cmhip.h"

namespace cm {

constexpr int kWave = 64;
constexpr int kMaxLevels = 21;          // 63-bit keys
constexpr int kStackCap = 1536;         // per-wave LDS traversal stack (ints)
constexpr int kSortTile = 2048;         // keys per radix-sort tile
constexpr int kSortThreads = 256;
constexpr int kTreeTile = 1024;         // bodies per tree-build tile
constexpr int kGroupTile = 1024;        // bodies per tile of the target-group ordering
constexpr int kMaxDevices = 64;         // per-device tables of static device memory
// largest body count of cm_energy_noalloc with a potential (its per-body rows
// live in a static device array of this many doubles, 32 MiB per GPU)
constexpr int kNoAllocMaxBodies = 1 << 22;
// FP32 pair loops may leave a body's own (dx = 0) term in the sum when the
// softening makes it exactly zero: r^2 = eps^2, rsq(eps^2)^3 <= 1e18 for
// eps^2 >= 1e-12, so m rsq^3 stays finite for any mass below 3e20 and the
// term is +-0.  Below this softening the loops skip the body itself by index.
constexpr double kSelfTermMinEps2 = 1e-12;

// rows of per-tile node counts the tree build keeps per level
inline int tree_tiles(int capacity) { return (capacity + kTreeTile - 1) / kTreeTile + 1; }

// SoA field order of BodySoA (SoA/BodySoA.cs:42-66).
enum Field : int {
    PX, PY, PZ, VX, VY, VZ, AX, AY, AZ, OAX, OAY, OAZ, MASS, RAD, DENS, NF64
};

// Device-resident per-evaluation header written and read by the kernels so
// that the host never needs the body count or level sizes of the tree.
struct TreeHeader {
    double box[4];             // root cube: centre x, y, z and half size
    int32_t n_active;
    int32_t n_nodes;
    int32_t overflow;
    int32_t n_levels_used;
    int32_t level_start[kMaxLevels + 3];
    unsigned long long n_cell_int;
    unsigned long long n_body_int;
    unsigned long long acc_int;   // cell + body terms since the last stats reset
    unsigned long long n_far_int; // (group, cell) far-field expansion terms
    unsigned long long acc_far;   // the same since the last stats reset
    unsigned long long pot_int[5];   // terms counted by potential evaluations (unused)
    int32_t n_groups;             // target groups formed for the traversal
    uint32_t work_next[8];        // per-XCD queues: next group of each eighth
    int32_t overflow_run;         // set by any evaluation of a run whose tree outgrew
                                  // the node pool; cleared by the host per API call
    // per-XCD queues of the traversal (k_group_order): entries
    // [qbase[q], qend[q]) of the queue array; a cache line of their own, away
    // from the atomically advanced work_next
    alignas(128) int32_t qbase[8];
    int32_t qend[8];
};

// Workgroups of one launch are dispatched round-robin over the 8 XCDs of an
// MI355X (each with its own L2): blockIdx.x % kXcds is the block's XCD.
constexpr int kXcds = 8;

constexpr int kWindow = 1024;     // force evaluations timed per host sync
// the first force evaluation of each API call and every kTimingStride-th
// one after it are timed with the three window events (each event record
// holds the stream for a few us)
constexpr int kTimingStride = 4;

struct Engine {
    int device = 0;
    int capacity = 0;
    int count = 0;
    hipStream_t stream = nullptr;
    cm_backend_config cfg{};
    std::string err;

    // BodySoA arrays (device)
    double* f[NF64] = {};
    uint8_t* act = nullptr;
    uint8_t* col = nullptr;
    int32_t* type = nullptr;

    // Barnes-Hut scratch
    uint64_t* keys_a = nullptr;
    uint64_t* keys_b = nullptr;
    int32_t* vals_a = nullptr;
    int32_t* vals_b = nullptr;
    uint32_t* hist = nullptr;          // 256 x sort tiles
    double* sorted = nullptr;          // 5 x capacity: x, y, z, m, radius
    double4* pm4 = nullptr;            // per body (x, y, z, m), written with the keys
    double* bbox_part = nullptr;       // per-block min/max partials
    int32_t* bbox_cnt = nullptr;
    TreeHeader* hdr = nullptr;
    TreeHeader* hdr_host = nullptr;   // pinned copy, refreshed after each evaluation
    int node_cap = 0;
    int4* nd_info = nullptr;           // first, nchild | level << 8, start, count
    double4* nd_cm = nullptr;          // com x, y, z, mass
    double* nd_sum = nullptr;          // 3 x node_cap: sum m*x, m*y, m*z
    double* nd_box = nullptr;          // 6 x node_cap: body bounding box lo, hi
    int32_t* nd_parent = nullptr;      // parent node of every node but the root
    uchar2* tree_dl = nullptr;         // per sorted body: shared digits with the
                                       // previous key, deepest level it opens
    uint32_t* tree_tcount = nullptr;   // (kMaxLevels + 1) x tree_tiles: nodes per tile
    int2* groups = nullptr;            // target groups: sorted body range [x, y), in body order
    int4* gq = nullptr;                // traversal queue entries (k_group_order): the group's
                                       // sorted range [x, y) and the targets [z, w) it computes
    int gq_cap = 0;                    // entries gq holds
    int32_t* gcost = nullptr;          // per group in body order: work in the last traversal
    long long* gwork = nullptr;        // per tile of 1024 bodies: its groups' work (>= 1 each)
    int32_t* bcost = nullptr;          // per body (caller's index): work units of its group
                                       // in the last traversal (orders the next one)
    int32_t* gmark = nullptr;          // per sorted body: end of the group starting there, or 0
    int32_t* gtile = nullptr;          // group counts / offsets per tile of 1024 bodies
    float* nd_quad = nullptr;          // 6 x node_cap second moments (learned correction)

    // learned far-field correction (k_mlp.hip)
    void* mlp = nullptr;

    // collisions / compaction scratch
    void* coll = nullptr;

    // 1PN decorator (k_pn.hip) and direct-sum softening modes (k_direct.hip)
    cm_pn_config pn{};
    void* pn_buf = nullptr;
    cm_softening_config soft{};
    cm_collision_config coll_cfg{};

    // diagnostics scratch
    double* red = nullptr;             // reduction partials
    // NBodySolver's diagnostics state (_currentTime, _initialEnergy), reset
    // when new bodies are uploaded
    double sim_time = 0.0;
    double initial_energy = 0.0;
    bool initial_energy_set = false;
    double* noalloc_scratch = nullptr; // device address of the static scratch of
                                       // cm_energy_noalloc on this engine's GPU
    double* sample = nullptr;          // direct-sample partials

    // timing
    bool timing = true;
    // the force statistics describe the last evaluation of a run: only that
    // one records the per-phase timing events and copies the tree header
    // back to the host (each event record costs the stream a few us)
    bool stats_eval = true;
    bool detail_timing() const { return timing && stats_eval; }
    hipEvent_t ev[6] = {};
    cm_force_stats stats{};
    // per-evaluation (start, force-kernel start, end) events of the timed
    // evaluations since the last host sync, folded into the accumulators by
    // finalize_stats; evals_since counts every evaluation
    hipEvent_t* wev = nullptr;        // 3 x kWindow
    int wn = 0;
    int evals_since = 0;
    long long eval_seq = 0;           // force evaluations since the call began
    double acc_ms_force = 0.0, acc_ms_total = 0.0;
    int acc_evals = 0, acc_timed = 0;

    // marks phase `which` (0 start, 1 force kernel, 2 end) of one force
    // evaluation: the events are recorded for every kTimingStride-th one
    void mark(int which) {
        const bool timed = eval_seq % kTimingStride == 0 && wn < kWindow;
        if (timed) (void)hipEventRecord(wev[3 * wn + which], stream);
        if (which == 2) {
            if (timed) wn++;
            evals_since++;
            eval_seq++;
        }
    }
};

// ---- kernels launchers (one .hip file each) --------------------------------
int launch_kick_drift(Engine& e, double dt);
int launch_kick_rotate(Engine& e, double dt);
int launch_direct(Engine& e, double softening);
int launch_barnes_hut(Engine& e, double softening);
int launch_direct_sample(Engine& e, double softening, const int32_t* targets,
                         int32_t nt, double* ax, double* ay, double* az);
int launch_energy(Engine& e, double softening, int with_pe, cm_energy_report* out);
int launch_energy_noalloc(Engine& e, double softening, int with_pe, cm_energy_report* out);
int noalloc_scratch_init(Engine& e);
int nbody_step_diag(Engine& e, double softening, double dt, int n_steps, int collisions,
                    int with_pe, int (*verlet_one)(Engine&, double, double),
                    cm_simulation_state* states);
int compare_forces(Engine& e, double softening, double theta, const int32_t* targets,
                   int32_t nt, cm_force_comparison* out);
int launch_bh_potential(Engine& e, double softening, double* rows);
int potential_rows(Engine& e, double softening, int method, double* rows);
int max_acceleration(Engine& e, double* out);
int collide_detect(Engine& e, cm_collision_event* out, int32_t cap, int32_t* ne);
int collide_resolve(Engine& e, const cm_collision_config& cfg, int32_t* merged);
int collide_compact(Engine& e, int32_t* new_count, int32_t* old_index);
void collide_free(Engine& e);
// Sorted Morton keys are 32-bit words when 3 * levels + 1 bits fit (levels
// <= 10), else 64-bit.
inline bool tree_keys32(const Engine& e) { return e.cfg.morton_levels <= 10; }
int build_tree(Engine& e, const void** keys_out, const int32_t** order_out);
// order: sorted position -> caller's index, from build_tree
int form_groups(Engine& e, const int32_t* order);
void mlp_free(Engine& e);
int launch_pn(Engine& e, double softening);
void pn_free(Engine& e);
bool mlp_ready(const Engine& e);
const uint16_t* mlp_weights(const Engine& e);   // bf16 [out][in] per layer
const float* mlp_params(const Engine& e);       // FP32 master (biases)

void radix_sort_pairs(hipStream_t st, int n, int first_pass, int passes,
                      uint64_t** k, int32_t** v, uint64_t** kt, int32_t** vt,
                      uint32_t* hist);
void radix_sort_pairs32(hipStream_t st, int n, int first_pass, int passes,
                        uint32_t** k, int32_t** v, uint32_t** kt, int32_t** vt,
                        uint32_t* hist);

int set_error(Engine& e, int code, const std::string& msg);
int finalize_stats(Engine& e);   // syncs the stream, fills e.stats
int begin_run(Engine& e);        // clears the run's sticky overflow flag

}  // namespace cm

#define CM_HIP_TRY(e, call)                                                    \
    do {                                                                       \
        hipError_t _st = (call);                                               \
        if (_st != hipSuccess)                                                 \
            return ::cm::set_error((e), CM_ERR_HIP,                            \
                                   std::string(#call) + ": " +                 \
                                       hipGetErrorString(_st));                \
    } while (0)
