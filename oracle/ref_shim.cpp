// ref_shim.cpp -- C entry points over the reference engine's own CPU code
// (engine/src/physics/*.cpp of mfscpayload-690/celestial-mechanics-desktop,
// compiled from where it lies by ref.mk).  TEST INFRASTRUCTURE ONLY.
#include <celestial/physics/barnes_hut_solver.hpp>
#include <celestial/physics/particle_system.hpp>

#include <cstdint>
#include <cstring>

using celestial::physics::BarnesHutSolver;
using celestial::physics::ParticleSystem;

static void load(ParticleSystem& ps, int n, const double* px, const double* py,
                 const double* pz, const double* m, const uint8_t* act) {
    ps.allocate(n > 0 ? n : 1);
    ps.set_count(n);
    for (int i = 0; i < n; i++) {
        ps.pos_x[i] = px[i]; ps.pos_y[i] = py[i]; ps.pos_z[i] = pz[i];
        ps.mass[i] = m[i]; ps.radius[i] = 0.0; ps.is_active[i] = act[i];
    }
}

extern "C" int ref_bh_forces(int n, const double* px, const double* py, const double* pz,
                             const double* m, const uint8_t* act, double eps, double theta,
                             double* ax, double* ay, double* az) {
    ParticleSystem ps;
    load(ps, n, px, py, pz, m, act);
    BarnesHutSolver bh;
    bh.theta = theta;
    bh.use_parallel = false;
    bh.compute_forces(ps, eps);
    for (int i = 0; i < n; i++) { ax[i] = ps.acc_x[i]; ay[i] = ps.acc_y[i]; az[i] = ps.acc_z[i]; }
    ps.free();
    return 0;
}

extern "C" int ref_compact(int n, double* px, double* mass, double* radius, uint8_t* act) {
    ParticleSystem ps;
    ps.allocate(n > 0 ? n : 1);
    ps.set_count(n);
    for (int i = 0; i < n; i++) {
        ps.pos_x[i] = px[i]; ps.mass[i] = mass[i]; ps.radius[i] = radius[i];
        ps.is_active[i] = act[i];
    }
    int k = ps.compact();
    for (int i = 0; i < k; i++) { px[i] = ps.pos_x[i]; mass[i] = ps.mass[i]; radius[i] = ps.radius[i]; }
    ps.free();
    return k;
}

#include <celestial/physics/post_newtonian.hpp>

// PostNewtonianCorrection::apply_corrections (engine/src/physics/post_newtonian.cpp:7-23)
// on a ParticleSystem whose accelerations start at (ax, ay, az).
extern "C" int ref_pn_apply(int n, const double* px, const double* py, const double* pz,
                            const double* vx, const double* vy, const double* vz,
                            const double* m, const uint8_t* act, double eps,
                            double* ax, double* ay, double* az) {
    ParticleSystem ps;
    load(ps, n, px, py, pz, m, act);
    for (int i = 0; i < n; i++) {
        ps.vel_x[i] = vx[i]; ps.vel_y[i] = vy[i]; ps.vel_z[i] = vz[i];
        ps.acc_x[i] = ax[i]; ps.acc_y[i] = ay[i]; ps.acc_z[i] = az[i];
    }
    celestial::physics::PostNewtonianCorrection pn;
    pn.softening_epsilon = eps;
    pn.apply_corrections(ps);
    for (int i = 0; i < n; i++) { ax[i] = ps.acc_x[i]; ay[i] = ps.acc_y[i]; az[i] = ps.acc_z[i]; }
    ps.free();
    return 0;
}
