/*
 * cm_oracle.h -- CPU oracle for the celestial-mechanics Barnes-Hut hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  It is used by tests/, by __graft_entry__.smoke() to check
 * its result, and by bench.py's cpu_baseline leg (timed as the CPU restatement
 * of the reference's FP64 path).
 *
 * Every function restates either
 *   (a) the reference's C# FP64 physics path (file:line cited per function), or
 *   (b) a definition this build adds on top of it (30-bit/63-bit Morton keys,
 *       stable key sort, the level-order bucket octree, the group-MAC
 *       traversal, its group far-field Taylor expansion, Kahan reductions,
 *       the correctly-rounded cube root),
 * in plain C with IEEE double arithmetic and no FMA contraction
 * (-ffp-contract=off), so that results are bit-reproducible.
 */
#ifndef CM_ORACLE_H
#define CM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- gravity ----------------------------------------------------------- */
double cmo_gravity_coeff(double dist_sq, double src_mass, double src_radius,
                         double eps2, int shell);
void cmo_direct_n3(int n, const double* px, const double* py, const double* pz,
                   const double* m, const uint8_t* act, double eps,
                   double* ax, double* ay, double* az);
void cmo_direct_gk(int n, const double* px, const double* py, const double* pz,
                   const double* m, const double* rad, const uint8_t* act,
                   double eps, int shell, double* ax, double* ay, double* az);
void cmo_direct_targets(int n, const double* px, const double* py,
                        const double* pz, const double* m, const uint8_t* act,
                        double eps, int nt, const int32_t* targets,
                        double* ax, double* ay, double* az, int nthreads);

/* ---- reference Barnes-Hut (C# insertion octree) ------------------------ */
typedef struct {
    int64_t node_count;
    int64_t interactions;   /* body-node + body-body terms evaluated */
    double build_ms;
    double traverse_ms;
} cmo_bh_stats;
int cmo_bh_csharp(int n, const double* px, const double* py, const double* pz,
                  const double* m, const double* rad, const uint8_t* act,
                  double eps, double theta, int shell, double min_node_size,
                  int nthreads, double* ax, double* ay, double* az,
                  cmo_bh_stats* stats);

/* ---- Verlet ------------------------------------------------------------ */
void cmo_verlet_kick_drift(int n, const uint8_t* act, double dt,
                           double* px, double* py, double* pz,
                           double* vx, double* vy, double* vz,
                           const double* oax, const double* oay, const double* oaz);
void cmo_verlet_kick_rotate(int n, const uint8_t* act, double dt,
                            double* vx, double* vy, double* vz,
                            const double* ax, const double* ay, const double* az,
                            double* oax, double* oay, double* oaz);

/* ---- build-defined octree ---------------------------------------------- */
void cmo_root_box(int n, const double* px, const double* py, const double* pz,
                  const uint8_t* act, double* box4 /* cx,cy,cz,half */);
uint64_t cmo_morton_key(double x, double y, double z, const double* box4,
                        int levels);
void cmo_morton_keys(int n, const double* px, const double* py,
                     const double* pz, const uint8_t* act, const double* box4,
                     int levels, uint64_t* keys);
void cmo_stable_sort_keys(int n, const uint64_t* keys, uint64_t* keys_sorted,
                          int32_t* order);
int cmo_tree_build(int n, const double* px, const double* py, const double* pz,
                   const double* m, const uint8_t* act, int levels, int leaf_cap,
                   int node_cap, double* box4, int32_t* n_active,
                   uint64_t* keys_sorted, int32_t* order,
                   int32_t* nd_level, int32_t* nd_start, int32_t* nd_count,
                   int32_t* nd_first, int32_t* nd_nchild, double* nd_mass,
                   double* nd_com /* 3 per node */, int32_t* level_start);
int cmo_bh_tree(int n, const double* px, const double* py, const double* pz,
                const double* m, const double* rad, const uint8_t* act,
                double eps, double theta, int shell, int levels, int leaf_cap,
                int group_size, int split_depth, int lifo /* stack capacity; 0 = FIFO */, int nthreads,
                double far_ratio /* 0 = no far-field expansion */,
                double* ax, double* ay, double* az,
                int32_t* n_cells, int32_t* n_parts, int32_t* max_queue,
                int64_t* far_total /* (group, cell) expansion terms; may be NULL */);
/* Far-field expansion of one cell (this build's definition, no reference
 * counterpart): adds to t[19] the Taylor coefficients about a group centre
 * of the softened monopole field of mass M at offset r = (rx, ry, rz) from
 * it: T0 (x, y, z), T1 (xx, yy, zz, xy, xz, yz), T2 (xxx, yyy, zzz, xxy,
 * xxz, yyx, yyz, zzx, zzy, xyz).  cmo_far_eval evaluates
 * T0 + T1 d + (1/2) T2 d d at offset d. */
void cmo_far_coeffs(double M, double rx, double ry, double rz, double eps2, double* t19);
void cmo_far_eval(const double* t19, double dx, double dy, double dz, double* out3);
/* Barnes-Hut potential shares rows[i] = 1/2 m_i phi_i: through the group
 * traversal (this build's definition), and as the engine's
 * compute_potential_bh_kernel over the C# insertion octree. */
int cmo_bh_potential(int n, const double* px, const double* py, const double* pz,
                     const double* m, const uint8_t* act, double eps, double theta,
                     int levels, int leaf_cap, int group_size, int split_depth, int lifo,
                     int nthreads, double* rows, int32_t* n_cells, int32_t* n_parts);
int cmo_bh_potential_reference(int n, const double* px, const double* py, const double* pz,
                               const double* m, const uint8_t* act, double eps, double theta,
                               int nthreads, double* rows);

/* ---- diagnostics -------------------------------------------------------- */
typedef struct {
    double ke, pe, px, py, pz, lx, ly, lz, l_mag;
} cmo_energy_t;
void cmo_energy_naive(int n, const double* px, const double* py,
                      const double* pz, const double* vx, const double* vy,
                      const double* vz, const double* m, const uint8_t* act,
                      double eps, int with_pe, cmo_energy_t* out);
void cmo_energy_exact(int n, const double* px, const double* py,
                      const double* pz, const double* vx, const double* vy,
                      const double* vz, const double* m, const uint8_t* act,
                      double eps, int with_pe, int nthreads, cmo_energy_t* out);
void cmo_center_of_mass(int n, const double* px, const double* py, const double* pz,
                        const double* m, const uint8_t* act, double* out4);
double cmo_max_acceleration(int n, const double* ax, const double* ay, const double* az,
                            const uint8_t* act);
typedef struct {
    double mean, max, rms, median;
    int32_t count;
} cmo_ferr_t;
void cmo_force_error(int n, const uint8_t* act, const double* ex,
                     const double* ey, const double* ez, const double* tx,
                     const double* ty, const double* tz, cmo_ferr_t* out);

/* ---- collisions -------------------------------------------------------- */
typedef struct {
    int32_t a, b;
    double overlap, dist, nx, ny, nz, rel_normal_speed;
} cmo_event_t;
int cmo_detect(int n, const double* px, const double* py, const double* pz,
               const double* vx, const double* vy, const double* vz,
               const double* m, const double* rad, const uint8_t* act,
               const uint8_t* col, cmo_event_t* out, int cap);
double cmo_cbrt(double x);
double cmo_body_radius(double mass, double density, int32_t type);
int cmo_resolve_merge(int ne, const cmo_event_t* ev, int n, double* px,
                      double* py, double* pz, double* vx, double* vy,
                      double* vz, double* ax, double* ay, double* az,
                      double* oax, double* oay, double* oaz, double* m,
                      double* rad, double* dens, uint8_t* act, uint8_t* col,
                      int32_t* type);
int cmo_compact(int n, double* const* f64_arrays, int n_f64,
                uint8_t* const* u8_arrays, int n_u8, int32_t* const* i32_arrays,
                int n_i32, const uint8_t* act, int32_t* old_index);

/* ---- widening rows (cm_oracle_ext.c) ----------------------------------- */
double cmo_c_sim(void);
int cmo_pn_apply(int n, const double* px, const double* py, const double* pz,
                 const double* vx, const double* vy, const double* vz,
                 const double* m, const uint8_t* act, double eps,
                 double max_vel_frac, double warn_factor, double* ax, double* ay,
                 double* az, int32_t* warn_i, int32_t* warn_j, double* warn_dist,
                 double* warn_rs, int warn_cap);
void cmo_direct_softened(int n, const double* px, const double* py,
                         const double* pz, const double* m, const uint8_t* act,
                         const int32_t* type, int mode, const double* table,
                         double scale, double mn, double global_eps, double* ax,
                         double* ay, double* az);
void cmo_yoshida_coefficients(double* w3);
void cmo_collision_energy(double m1, double m2, double eff_radius_au, double rel_speed,
                          double max_loss, int32_t* is_merge, int32_t* is_frag,
                          double* ejecta_kg, double* ratio_out);
int cmo_resolve_events(int ne, const cmo_event_t* ev, int n, double* px, double* py,
                       double* pz, double* vx, double* vy, double* vz, double* ax,
                       double* ay, double* az, double* oax, double* oay, double* oaz,
                       double* m, double* rad, double* dens, uint8_t* act, uint8_t* col,
                       int32_t* type, int mode, double default_restitution,
                       double mass_loss_cap, int promote_remnants, int max_per_frame,
                       int max_per_body, int32_t* merge_count_scratch);

/* .NET System.Random(seed) draws and PlummerSphere.Generate (cm_oracle_ext.c) */
void cmo_dotnet_random(int32_t seed, int count, int32_t* out);
void cmo_plummer_dotnet(int n, double total_mass, double a, int32_t seed, double* px,
                        double* py, double* pz, double* vx, double* vy, double* vz,
                        double* m);

#ifdef __cplusplus
}
#endif
#endif
