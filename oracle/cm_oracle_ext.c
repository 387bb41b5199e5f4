/*
 * cm_oracle_ext.c -- CPU oracle for the rows that widen the hot path: the 1PN
 * decorator, the engine's softening modes and the Yoshida-4 composition.
 *
 * TEST INFRASTRUCTURE ONLY (see cm_oracle.h).  Plain C, IEEE double, no FMA
 * contraction (-ffp-contract=off): every expression keeps the reference's
 * operation order, so the GPU kernels can be compared bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "cm_oracle.h"

/* PhysicalConstants (src/CelestialMechanics.Math/PhysicalConstants.cs:12-50),
 * equal to the engine's core constants (engine/include/celestial/core/
 * constants.hpp:23-46).  C# folds these constant expressions in double. */
static const double kC = 2.998e8;
static const double kTimeUnit = 5.0226e6;
static const double kAU = 1.496e11;

double cmo_c_sim(void) { return kC * kTimeUnit / kAU; }

/* PostNewtonian1Correction.ComputeCorrection
 * (src/CelestialMechanics.Physics/Extensions/PostNewtonian1Correction.cs:66-165)
 * for target i over every j != i with m_j > 0 (inactive sources included:
 * the reference loops over all `count` bodies), and ApplyCorrections (:171-188):
 * acc_i += (dax, day, daz) for active i only.  The velocity cut is
 * max_vel_frac * c (the C# class reads the static 0.3 c; the engine port,
 * engine/src/physics/post_newtonian.cpp:36-38, reads its field, default 0.3).
 * Pairs inside warn_factor Schwarzschild radii are skipped and reported in
 * (i, j) order, as OnSchwarzschildProximityWarning would be invoked. */
int cmo_pn_apply(int n, const double* px, const double* py, const double* pz,
                 const double* vx, const double* vy, const double* vz,
                 const double* m, const uint8_t* act, double eps,
                 double max_vel_frac, double warn_factor, double* ax, double* ay,
                 double* az, int32_t* warn_i, int32_t* warn_j, double* warn_dist,
                 double* warn_rs, int warn_cap) {
    const double c_sim = kC * kTimeUnit / kAU;
    const double c_sim2 = c_sim * c_sim;
    const double G = 1.0;
    const double inv_c2 = 1.0 / c_sim2;
    const double max_vel = max_vel_frac * c_sim;
    const double max_vel2 = max_vel * max_vel;
    const double rs_factor = 2.0 * G / c_sim2;
    const double eps2 = eps * eps;
    int nw = 0;
    for (int i = 0; i < n; i++) {
        if (!act[i]) continue;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        const double xi = px[i], yi = py[i], zi = pz[i];
        const double vxi = vx[i], vyi = vy[i], vzi = vz[i];
        const double mi = m[i];
        const double vi2 = vxi * vxi + vyi * vyi + vzi * vzi;
        for (int j = 0; j < n; j++) {
            if (j == i) continue;
            const double mj = m[j];
            if (mj <= 0.0) continue;
            const double dx = px[j] - xi, dy = py[j] - yi, dz = pz[j] - zi;
            const double dist2 = dx * dx + dy * dy + dz * dz + eps2;
            const double dist = sqrt(dist2);
            const double inv = 1.0 / dist;
            const double dvx = vxi - vx[j], dvy = vyi - vy[j], dvz = vzi - vz[j];
            const double vrel2 = dvx * dvx + dvy * dvy + dvz * dvz;
            if (vrel2 > max_vel2) continue;
            const double rs = rs_factor * (mi + mj);
            if (dist < warn_factor * rs) {
                if (nw < warn_cap) {
                    warn_i[nw] = i; warn_j[nw] = j;
                    warn_dist[nw] = dist; warn_rs[nw] = rs;
                }
                nw++;
                continue;
            }
            const double nx = dx * inv, ny = dy * inv, nz = dz * inv;
            const double vxj = vx[j], vyj = vy[j], vzj = vz[j];
            const double vj2 = vxj * vxj + vyj * vyj + vzj * vzj;
            const double vivj = vxi * vxj + vyi * vyj + vzi * vzj;
            const double nvi = nx * vxi + ny * vyi + nz * vzi;
            const double nvj = nx * vxj + ny * vyj + nz * vzj;
            const double gmi_r = G * mi * inv;
            const double gmj_r = G * mj * inv;
            const double pre = gmj_r * inv * inv_c2;
            const double t1 = -vi2 - 2.0 * vj2 + 4.0 * vivj + 1.5 * nvj * nvj
                              + 5.0 * gmi_r + 4.0 * gmj_r;
            const double t2 = 4.0 * nvi - 3.0 * nvj;
            sx += pre * (t1 * nx + t2 * dvx);
            sy += pre * (t1 * ny + t2 * dvy);
            sz += pre * (t1 * nz + t2 * dvz);
        }
        ax[i] += sx;
        ay[i] += sy;
        az[i] += sz;
    }
    return nw;
}

/* Engine::compute_cpu_brute_forces (engine/src/sim/engine.cpp:581-660): zero
 * every acceleration, then for each active i the sweep over active j != i in
 * ascending j with the pairwise softening of the configured mode
 *   mode 1 (PerBodyType, :587-616): eps = (t[type_i] + t[type_j]) * 0.5,
 *   mode 2 (Adaptive,    :617-650): eps_k = max(scale * cbrt(m_k), min),
 *                                   eps = (eps_i + eps_j) * 0.5,
 *   mode 0 (Global,      :651-660): eps = global_eps,
 * and the engine's operation order: inv3 = inv*inv*inv, factor = m_j * inv3.
 * Types index the table modulo 16 as the GPU kernel does
 * (engine/src/cuda/gravity_kernel.cu:131, 168).  The cube root is the
 * correctly rounded one (cmo_cbrt); glibc's cbrt, which the engine calls,
 * is within one ulp of it. */
static double soft_eps(int mode, const double* table, double scale, double mn,
                       double global_eps, double mass, int32_t type) {
    if (mode == 1) return table[type & 15];
    if (mode == 2) {
        double e = scale * cmo_cbrt(mass);
        return e < mn ? mn : e;   /* std::max(e, mn) */
    }
    return global_eps;
}

void cmo_direct_softened(int n, const double* px, const double* py,
                         const double* pz, const double* m, const uint8_t* act,
                         const int32_t* type, int mode, const double* table,
                         double scale, double mn, double global_eps, double* ax,
                         double* ay, double* az) {
    for (int i = 0; i < n; i++) { ax[i] = 0.0; ay[i] = 0.0; az[i] = 0.0; }
    for (int i = 0; i < n; i++) {
        if (!act[i]) continue;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        const double xi = px[i], yi = py[i], zi = pz[i];
        const double ei = soft_eps(mode, table, scale, mn, global_eps, m[i], type[i]);
        for (int j = 0; j < n; j++) {
            if (i == j || !act[j]) continue;
            double eps2;
            if (mode == 1 || mode == 2) {
                const double ej = soft_eps(mode, table, scale, mn, global_eps, m[j], type[j]);
                const double e = (ei + ej) * 0.5;
                eps2 = e * e;
            } else {
                eps2 = global_eps * global_eps;
            }
            const double dx = xi - px[j], dy = yi - py[j], dz = zi - pz[j];
            const double d2 = dx * dx + dy * dy + dz * dz + eps2;
            const double inv = 1.0 / sqrt(d2);
            const double inv3 = inv * inv * inv;
            const double f = m[j] * inv3;
            sx -= f * dx; sy -= f * dy; sz -= f * dz;
        }
        ax[i] = sx; ay[i] = sy; az[i] = sz;
    }
}

/* YoshidaCoefficients (engine/include/celestial/sim/yoshida.hpp:18-28):
 * W1 = 1/(2 - cbrt 2), W0 = -cbrt 2/(2 - cbrt 2) with the header's literal
 * for cbrt 2; sub-step s of a step of size dt uses SUB_DT[s] * dt. */
void cmo_yoshida_coefficients(double* w3) {
    const double cbrt2 = 1.2599210498948732;
    const double w1 = 1.0 / (2.0 - cbrt2);
    const double w0 = -cbrt2 / (2.0 - cbrt2);
    w3[0] = w1; w3[1] = w0; w3[2] = w1;
}

/* ---- collision outcomes ------------------------------------------------- */
/* .NET Math.Max / Math.Clamp for doubles (Clamp's min <= max is the caller's
 * guarantee here; the reference throws otherwise). */
static double cs_max(double a, double b) { return (a > b || isnan(a)) ? a : b; }
static double cs_clamp(double v, double lo, double hi) {
    if (v < lo) return lo;
    if (v > hi) return hi;
    return v;
}

static const double kSolarMass = 1.989e30;
static const double kGSi = 6.67430e-11;

/* CollisionEnergyModel.Evaluate (src/CelestialMechanics.Physics/Astrophysics/
 * CollisionEnergyModel.cs:20-70): the fields that change the state. */
void cmo_collision_energy(double m1, double m2, double eff_radius_au, double rel_speed,
                          double max_loss, int32_t* is_merge, int32_t* is_frag,
                          double* ejecta_kg, double* ratio_out) {
    const double m1kg = cs_max(m1, 1e-15) * kSolarMass;
    const double m2kg = cs_max(m2, 1e-15) * kSolarMass;
    const double mt = m1kg + m2kg;
    const double rm = cs_max(eff_radius_au, 1e-12) * kAU;
    const double v = (cs_max(rel_speed, 0.0) * kAU) / kTimeUnit;
    const double mu = (m1kg * m2kg) / cs_max(mt, 1e-15);
    const double ec = 0.5 * mu * v * v;
    const double eb = (3.0 / 5.0) * kGSi * mt * mt / cs_max(rm, 1.0);
    const double ratio = eb > 0.0 ? ec / eb : 0.0;
    const int merge = ec < 0.5 * eb;
    const int frag = !merge && ec < eb;
    const int cat = !merge && !frag;
    double f = 0.0;
    if (frag) {
        const double nrm = cs_clamp((ratio - 0.5) / 0.5, 0.0, 1.0);
        f = cs_clamp(0.08 + nrm * 0.32, 0.05, max_loss);
    } else if (cat) {
        const double nrm = cs_clamp(ratio - 1.0, 0.0, 12.0) / 12.0;
        const double ccap = cs_max(max_loss, 0.55);
        f = cs_clamp(0.5 + nrm * 0.4, 0.45, ccap);
    }
    *is_merge = merge;
    *is_frag = frag;
    *ejecta_kg = mt * f;
    if (ratio_out) *ratio_out = ratio;
}

/* CollisionResolver.GetRestitutionForType (Collisions/CollisionResolver.cs:456-470),
 * BodyType order Star, Planet, GasGiant, RockyPlanet, Moon, Asteroid,
 * NeutronStar, BlackHole, Comet, Custom (Types/BodyType.cs). */
static double restitution_for(int32_t t, double dflt) {
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

static int is_compact(int32_t t) { return t == 6 || t == 7; }

/* DensityModel.GetDefaultDensity (Types/DensityModel.cs:76-87). */
static double default_density(int32_t t) {
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

typedef struct {
    double *px, *py, *pz, *vx, *vy, *vz, *ax, *ay, *az, *oax, *oay, *oaz, *m, *rad, *dens;
    uint8_t *act, *col;
    int32_t* type;
} cmo_soa_t;

/* ResolveImpulse (CollisionResolver.cs:378-405). */
static void impulse(cmo_soa_t* s, int a, int b, double nx, double ny, double nz,
                    double rns, double e) {
    if (rns >= 0.0) return;
    const double ma = s->m[a], mb = s->m[b];
    if (ma <= 0.0 || mb <= 0.0) return;
    const double ia = 1.0 / ma, ib = 1.0 / mb;
    const double j = -(1.0 + e) * rns / (ia + ib);
    s->vx[a] -= j * ia * nx; s->vy[a] -= j * ia * ny; s->vz[a] -= j * ia * nz;
    s->vx[b] += j * ib * nx; s->vy[b] += j * ib * ny; s->vz[b] += j * ib * nz;
}

/* PositionalCorrection (CollisionResolver.cs:407-437). */
static void positional(cmo_soa_t* s, int a, int b, double depth, double nx, double ny,
                       double nz) {
    if (depth <= 0.0) return;
    const double slop = 1e-4, percent = 0.6;
    const double ma = s->m[a], mb = s->m[b];
    if (ma <= 0.0 || mb <= 0.0) return;
    const double ia = 1.0 / ma, ib = 1.0 / mb;
    const double c = percent * cs_max(depth - slop, 0.0) / (ia + ib);
    s->px[a] -= c * ia * nx; s->py[a] -= c * ia * ny; s->pz[a] -= c * ia * nz;
    s->px[b] += c * ib * nx; s->py[b] += c * ib * ny; s->pz[b] += c * ib * nz;
}

/* ConservePairMomentum (CollisionResolver.cs:527-546). */
static void conserve_pair(cmo_soa_t* s, int a, int b, double dx, double dy, double dz) {
    const double ma = cs_max(s->m[a], 1e-15), mb = cs_max(s->m[b], 1e-15);
    const double cx = ma * s->vx[a] + mb * s->vx[b];
    const double cy = ma * s->vy[a] + mb * s->vy[b];
    const double cz = ma * s->vz[a] + mb * s->vz[b];
    const double ex = dx - cx, ey = dy - cy, ez = dz - cz;
    const double it = 1.0 / (ma + mb);
    s->vx[a] += ex * it; s->vy[a] += ey * it; s->vz[a] += ez * it;
    s->vx[b] += ex * it; s->vy[b] += ey * it; s->vz[b] += ez * it;
}

static void deactivate(cmo_soa_t* s, int b) {
    s->act[b] = 0; s->col[b] = 0;
    s->ax[b] = 0.0; s->ay[b] = 0.0; s->az[b] = 0.0;
    s->oax[b] = 0.0; s->oay[b] = 0.0; s->oaz[b] = 0.0;
}

/* MergePolicy.Merge (Collisions/MergePolicy.cs:31-80). */
static void merge_policy(cmo_soa_t* s, int a, int b) {
    const double ma = s->m[a], mb = s->m[b];
    const double mt = ma + mb;
    if (mt <= 0.0) return;
    const double inv = 1.0 / mt;
    s->vx[a] = (ma * s->vx[a] + mb * s->vx[b]) * inv;
    s->vy[a] = (ma * s->vy[a] + mb * s->vy[b]) * inv;
    s->vz[a] = (ma * s->vz[a] + mb * s->vz[b]) * inv;
    s->px[a] = (ma * s->px[a] + mb * s->px[b]) * inv;
    s->py[a] = (ma * s->py[a] + mb * s->py[b]) * inv;
    s->pz[a] = (ma * s->pz[a] + mb * s->pz[b]) * inv;
    s->m[a] = mt;
    s->dens[a] = (ma * s->dens[a] + mb * s->dens[b]) * inv;
    if (s->type[b] == 7) s->type[a] = 7;
    s->rad[a] = cmo_body_radius(mt, s->dens[a], s->type[a]);
    deactivate(s, b);
}

/* PromoteCompactRemnant (CollisionResolver.cs:577-607). */
static void promote(cmo_soa_t* s, int a) {
    const int32_t cur = s->type[a];
    if (!(cur == 0 || cur == 6 || cur == 7)) return;
    const double mass = s->m[a];
    int32_t tgt = cur;
    if (mass >= 3.0) tgt = 7;
    else if (mass >= 1.6) tgt = 6;
    if (tgt == cur) return;
    const double d = default_density(tgt);
    s->type[a] = tgt;
    s->dens[a] = d;
    s->rad[a] = cmo_body_radius(mass, d, tgt);
}

/* CollisionResolver.Resolve (Collisions/CollisionResolver.cs:58-247) over an
 * ordered event list, every state-changing branch: MergeOnly (0),
 * BounceOnly (1) and Realistic (2) modes; outcomes Merge / Accretion
 * (MergePolicy.Merge, optional PromoteCompactRemnant), Fragmentation
 * (impulse at 0.35 x restitution, positional correction, mass loss, pair
 * momentum restored), CatastrophicDisruption (:285-333) and Bounce (impulse
 * + positional correction).  Telemetry (bursts, accretion disks) does not
 * touch the body state and is not restated.  The optional caps are the
 * engine's merge safeguards (engine/src/physics/collision_resolver.cpp:46-76,
 * simulation_config.hpp:52-60), MergeOnly only; <= 0 means no cap.  Returns
 * the number of events resolved (LastCollisionCount). */
int cmo_resolve_events(int ne, const cmo_event_t* ev, int n, double* px, double* py,
                       double* pz, double* vx, double* vy, double* vz, double* ax,
                       double* ay, double* az, double* oax, double* oay, double* oaz,
                       double* m, double* rad, double* dens, uint8_t* act, uint8_t* col,
                       int32_t* type, int mode, double default_restitution,
                       double mass_loss_cap, int promote_remnants, int max_per_frame,
                       int max_per_body, int32_t* merge_count_scratch) {
    cmo_soa_t S = {px, py, pz, vx, vy, vz, ax, ay, az, oax, oay, oaz, m, rad, dens,
                   act, col, type};
    cmo_soa_t* s = &S;
    int resolved = 0, merges = 0;
    if (merge_count_scratch) memset(merge_count_scratch, 0, sizeof(int32_t) * (size_t)n);
    for (int k = 0; k < ne; k++) {
        int a = ev[k].a, b = ev[k].b;
        if (!act[a] || !act[b]) continue;
        double ma = m[a], mb = m[b];
        if (ma <= 0.0 || mb <= 0.0) continue;
        double nx = px[b] - px[a], ny = py[b] - py[a], nz = pz[b] - pz[a];
        const double dist = sqrt(nx * nx + ny * ny + nz * nz);
        if (dist > 1e-12) {
            const double inv = 1.0 / dist;
            nx *= inv; ny *= inv; nz *= inv;
        } else {
            nx = 1.0; ny = 0.0; nz = 0.0;
        }
        const double rvx = vx[b] - vx[a], rvy = vy[b] - vy[a], rvz = vz[b] - vz[a];
        const double rns = rvx * nx + rvy * ny + rvz * nz;
        const double rs = sqrt(rvx * rvx + rvy * rvy + rvz * rvz);
        int32_t is_merge, is_frag;
        double ej_kg;
        cmo_collision_energy(ma, mb, 0.5 * cs_max(rad[a] + rad[b], 1e-12), rs, mass_loss_cap,
                             &is_merge, &is_frag, &ej_kg, NULL);
        /* SelectOutcome (:249-267): 0 Bounce, 1 Merge, 2 Fragmentation,
         * 3 Accretion, 4 CatastrophicDisruption */
        int outcome;
        if (is_compact(type[a]) || is_compact(type[b])) outcome = 3;
        else if (is_merge) outcome = 1;
        else if (is_frag) outcome = 2;
        else outcome = 4;
        if (mode == 1) outcome = 0;
        else if (mode == 0) outcome = 1;
        if (outcome == 3 && !is_compact(type[a]) && is_compact(type[b])) {
            int t = a; a = b; b = t;   /* EnsureAccretorSurvivor (:565-575) */
        }
        ma = m[a];
        mb = m[b];
        if (outcome == 1 || outcome == 3) {
            if (max_per_frame > 0 && merges >= max_per_frame) continue;
            if (max_per_body > 0 && merge_count_scratch &&
                (merge_count_scratch[a] >= max_per_body || merge_count_scratch[b] >= max_per_body))
                continue;
            merge_policy(s, a, b);
            if (promote_remnants) promote(s, a);
            if (merge_count_scratch) merge_count_scratch[a]++;
            merges++;
        } else if (outcome == 2) {
            const double e = cs_clamp(0.5 * (restitution_for(type[a], default_restitution) +
                                              restitution_for(type[b], default_restitution)),
                                      0.0, 1.0);
            impulse(s, a, b, nx, ny, nz, rns, e * 0.35);
            positional(s, a, b, ev[k].overlap, nx, ny, nz);
            const double ej = ej_kg / kSolarMass;
            const double ef = cs_clamp(ej / cs_max(ma + mb, 1e-15), 0.02, mass_loss_cap);
            const double keep = 1.0 - ef;
            const double pbx = ma * vx[a] + mb * vx[b];
            const double pby = ma * vy[a] + mb * vy[b];
            const double pbz = ma * vz[a] + mb * vz[b];
            m[a] = cs_max(1e-12, ma * keep);
            m[b] = cs_max(1e-12, mb * keep);
            rad[a] = cmo_body_radius(m[a], dens[a], type[a]);
            rad[b] = cmo_body_radius(m[b], dens[b], type[b]);
            conserve_pair(s, a, b, pbx, pby, pbz);
        } else if (outcome == 4) {
            const double mt = ma + mb;
            if (mt > 0.0) {
                const double ej = ej_kg / kSolarMass;
                const double sm = cs_max(1e-12, mt - ej);
                const double qx = ma * vx[a] + mb * vx[b];
                const double qy = ma * vy[a] + mb * vy[b];
                const double qz = ma * vz[a] + mb * vz[b];
                px[a] = (ma * px[a] + mb * px[b]) / mt;
                py[a] = (ma * py[a] + mb * py[b]) / mt;
                pz[a] = (ma * pz[a] + mb * pz[b]) / mt;
                vx[a] = qx / sm; vy[a] = qy / sm; vz[a] = qz / sm;
                m[a] = sm;
                const int32_t ta = type[a], tb = type[b];
                const int32_t st = (is_compact(ta) || is_compact(tb))
                                       ? (is_compact(ta) ? ta : tb)
                                       : (ma >= mb ? ta : tb);
                type[a] = st;
                rad[a] = cmo_body_radius(m[a], dens[a], st);
                deactivate(s, b);
            }
        } else {
            const double e = cs_clamp(0.5 * (restitution_for(type[a], default_restitution) +
                                              restitution_for(type[b], default_restitution)),
                                      0.0, 1.0);
            impulse(s, a, b, nx, ny, nz, rns, e);
            positional(s, a, b, ev[k].overlap, nx, ny, nz);
        }
        resolved++;
    }
    return resolved;
}

/* ---- .NET System.Random and PlummerSphere.Generate ---------------------- */
/* System.Random(seed) as .NET implements it for an explicit seed (Knuth's
 * subtractive generator, "Seminumerical Algorithms" 3.2.2, in the form of
 * Numerical Recipes' ran3): 55 words modulo MBIG = 2^31 - 1, seeded from
 * MSEED - |seed| (MSEED = 161803398) by the 21-step index walk, four mixing
 * rounds, then draws a[i] - a[i + 31 mod 55] with the two indices starting at
 * 0 and 21.  NextDouble = draw * (1 / MBIG).  The reference's
 * PlummerSphere.Generate (ScenarioGenerators/PlummerSphere.cs:44) draws from
 * it; .NET itself is absent here, so known answers of Random(0) pin it. */
typedef struct { int32_t a[56]; int p, q; } cmo_netrand;

static void netrand_seed(cmo_netrand* r, int32_t seed) {
    const int32_t mbig = 2147483647;
    int32_t mj = 161803398 - (seed == INT32_MIN ? mbig : abs(seed));
    int32_t mk = 1;
    memset(r->a, 0, sizeof(r->a));
    r->a[55] = mj;
    for (int i = 1; i < 55; i++) {
        int ii = (21 * i) % 55;
        r->a[ii] = mk;
        mk = mj - mk;
        if (mk < 0) mk += mbig;
        mj = r->a[ii];
    }
    for (int k = 1; k < 5; k++)
        for (int i = 1; i < 56; i++) {
            r->a[i] -= r->a[1 + (i + 30) % 55];
            if (r->a[i] < 0) r->a[i] += mbig;
        }
    r->p = 0;
    r->q = 21;
}

static int32_t netrand_next(cmo_netrand* r) {
    if (++r->p >= 56) r->p = 1;
    if (++r->q >= 56) r->q = 1;
    int32_t v = r->a[r->p] - r->a[r->q];
    if (v == 2147483647) v--;
    if (v < 0) v += 2147483647;
    r->a[r->p] = v;
    return v;
}

static double netrand_double(cmo_netrand* r) { return netrand_next(r) * (1.0 / 2147483647.0); }

/* The first `count` raw draws (InternalSample) of System.Random(seed). */
void cmo_dotnet_random(int32_t seed, int count, int32_t* out) {
    cmo_netrand r;
    netrand_seed(&r, seed);
    for (int i = 0; i < count; i++) out[i] = netrand_next(&r);
}

/* PlummerSphere.Generate(count, totalMass, scaleLength, seed)
 * (ScenarioGenerators/PlummerSphere.cs:41-116), G = 1. */
void cmo_plummer_dotnet(int n, double total_mass, double a, int32_t seed, double* px,
                        double* py, double* pz, double* vx, double* vy, double* vz,
                        double* m) {
    cmo_netrand r;
    netrand_seed(&r, seed);
    const double pi = 3.14159265358979323846;
    double bm = total_mass / n;
    for (int i = 0; i < n; i++) {
        double u = netrand_double(&r);
        if (u < 1e-10) u = 1e-10;                 /* Math.Max(u, 1e-10) */
        if (u > 1.0 - 1e-10) u = 1.0 - 1e-10;     /* Math.Min(u, 1 - 1e-10) */
        double rad = a / sqrt(pow(u, -2.0 / 3.0) - 1.0);
        double ct = 2.0 * netrand_double(&r) - 1.0;
        double st = sqrt(1.0 - ct * ct);
        double phi = 2.0 * pi * netrand_double(&r);
        px[i] = rad * st * cos(phi);
        py[i] = rad * st * sin(phi);
        pz[i] = rad * ct;
        double vesc = sqrt(2.0 * 1.0 * total_mass / sqrt(rad * rad + a * a));
        double speed;
        for (;;) {
            double q = netrand_double(&r);
            double g = q * q * pow(1.0 - q * q, 3.5);
            if (netrand_double(&r) * 0.1 < g) {
                speed = q * vesc;
                break;
            }
        }
        double vct = 2.0 * netrand_double(&r) - 1.0;
        double vst = sqrt(1.0 - vct * vct);
        double vphi = 2.0 * pi * netrand_double(&r);
        vx[i] = speed * vst * cos(vphi);
        vy[i] = speed * vst * sin(vphi);
        vz[i] = speed * vct;
        m[i] = bm;
    }
}
