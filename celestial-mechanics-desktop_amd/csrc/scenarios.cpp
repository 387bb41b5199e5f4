// scenarios.cpp -- deterministic synthetic initial conditions (host code).
//
// Random numbers come from one SplitMix64 stream per call (the mixing
// constants of the reference's DeterministicMode::deterministic_hash,
// engine/src/sim/deterministic.cpp:21-28); NextDouble = top 53 bits / 2^53.
// cm_generate_plummer_dotnet instead draws from the stream of .NET's seeded
// System.Random (Knuth's subtractive generator as .NET implements it for an
// explicit seed), so that it reproduces PlummerSphere.Generate(count, M, a,
// seed) body for body.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/cmhip.h"

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

// System.Random(seed) of .NET (the seeded, framework-compatible algorithm):
// a 55-word lagged subtractive generator modulo 2^31 - 1, seeded from
// 161803398 - |seed| with the 21-step index walk and four mixing rounds;
// NextDouble = Sample() = draw * (1 / (2^31 - 1)).
struct DotNetRandom {
    int32_t a[56];
    int inext = 0, inextp = 21;
    explicit DotNetRandom(int32_t seed) {
        const int32_t big = 2147483647;
        int32_t mj = 161803398 - (seed == INT32_MIN ? big : (seed < 0 ? -seed : seed));
        int32_t mk = 1;
        a[0] = 0;
        a[55] = mj;
        for (int i = 1; i < 55; i++) {
            int ii = (21 * i) % 55;
            a[ii] = mk;
            mk = mj - mk;
            if (mk < 0) mk += big;
            mj = a[ii];
        }
        for (int round = 1; round < 5; round++)
            for (int i = 1; i < 56; i++) {
                a[i] -= a[1 + (i + 30) % 55];
                if (a[i] < 0) a[i] += big;
            }
    }
    int32_t next() {
        int p = inext + 1 >= 56 ? 1 : inext + 1;
        int q = inextp + 1 >= 56 ? 1 : inextp + 1;
        int32_t v = a[p] - a[q];
        if (v == 2147483647) v--;
        if (v < 0) v += 2147483647;
        a[p] = v;
        inext = p;
        inextp = q;
        return v;
    }
    double uniform() { return next() * (1.0 / 2147483647.0); }
};

constexpr double kPi = 3.14159265358979323846;

// PlummerSphere.Generate's per-body loop (ScenarioGenerators/PlummerSphere.cs:
// 52-106), drawing NextDouble from `rng` in the reference's order.
template <class Rng>
void plummer_bodies(Rng& rng, int32_t n, double total_mass, double a, double cut,
                    const cm_soa_ptrs* out) {
    const double body_mass = total_mass / n;
    for (int32_t i = 0; i < n; i++) {
        double u = rng.uniform();
        if (cut < 1.0) u *= cut;   // truncation
        u = std::max(u, 1e-10);
        u = std::min(u, 1.0 - 1e-10);
        double r = a / std::sqrt(std::pow(u, -2.0 / 3.0) - 1.0);
        double ct = 2.0 * rng.uniform() - 1.0;
        double stt = std::sqrt(1.0 - ct * ct);
        double phi = 2.0 * kPi * rng.uniform();
        out->pos_x[i] = r * stt * std::cos(phi);
        out->pos_y[i] = r * stt * std::sin(phi);
        out->pos_z[i] = r * ct;
        double vesc = std::sqrt(2.0 * total_mass / std::sqrt(r * r + a * a));
        double speed;
        for (;;) {
            double q = rng.uniform();
            double g = q * q * std::pow(1.0 - q * q, 3.5);
            if (rng.uniform() * 0.1 < g) { speed = q * vesc; break; }
        }
        double vct = 2.0 * rng.uniform() - 1.0;
        double vst = std::sqrt(1.0 - vct * vct);
        double vphi = 2.0 * kPi * rng.uniform();
        out->vel_x[i] = speed * vst * std::cos(vphi);
        out->vel_y[i] = speed * vst * std::sin(vphi);
        out->vel_z[i] = speed * vct;
        out->mass[i] = body_mass;
        if (out->radius) out->radius[i] = 0.02;
    }
}

bool plummer_args_ok(int32_t n, double total_mass, double scale_length, const cm_soa_ptrs* out) {
    return n >= 0 && out && out->pos_x && out->pos_y && out->pos_z && out->vel_x &&
           out->vel_y && out->vel_z && out->mass && total_mass > 0.0 && scale_length > 0.0;
}

}  // namespace

extern "C" int32_t cm_generate_plummer(int32_t n, double total_mass,
                                       double scale_length, uint64_t seed,
                                       double mass_fraction_cut,
                                       int32_t com_frame,
                                       const cm_soa_ptrs* out) {
    if (!plummer_args_ok(n, total_mass, scale_length, out) ||
        !(mass_fraction_cut > 0.0) || mass_fraction_cut > 1.0)
        return CM_ERR_INVALID_ARGUMENT;
    if (n == 0) return CM_OK;
    SplitMix64 rng(seed);
    plummer_bodies(rng, n, total_mass, scale_length, mass_fraction_cut, out);
    if (com_frame) {
        long double m = 0, cx = 0, cy = 0, cz = 0, ux = 0, uy = 0, uz = 0;
        for (int32_t i = 0; i < n; i++) {
            long double mi = out->mass[i];
            m += mi;
            cx += mi * out->pos_x[i]; cy += mi * out->pos_y[i]; cz += mi * out->pos_z[i];
            ux += mi * out->vel_x[i]; uy += mi * out->vel_y[i]; uz += mi * out->vel_z[i];
        }
        double px = (double)(cx / m), py = (double)(cy / m), pz = (double)(cz / m);
        double vx = (double)(ux / m), vy = (double)(uy / m), vz = (double)(uz / m);
        for (int32_t i = 0; i < n; i++) {
            out->pos_x[i] -= px; out->pos_y[i] -= py; out->pos_z[i] -= pz;
            out->vel_x[i] -= vx; out->vel_y[i] -= vy; out->vel_z[i] -= vz;
        }
    }
    return CM_OK;
}

extern "C" int32_t cm_generate_plummer_dotnet(int32_t n, double total_mass,
                                              double scale_length, int32_t seed,
                                              const cm_soa_ptrs* out) {
    if (!plummer_args_ok(n, total_mass, scale_length, out) || seed < 0)
        return CM_ERR_INVALID_ARGUMENT;
    if (n == 0) return CM_OK;
    DotNetRandom rng(seed);
    plummer_bodies(rng, n, total_mass, scale_length, 1.0, out);
    return CM_OK;
}

namespace {

// Kepler elements -> Cartesian state about a central mass mu (G = 1).
void kepler_to_cartesian(double mu, double a, double e, double inc, double node,
                         double peri, double mean_anom, double* r, double* v) {
    double E = mean_anom;
    for (int it = 0; it < 30; it++) {
        double f = E - e * std::sin(E) - mean_anom;
        double d = 1.0 - e * std::cos(E);
        double step = f / d;
        E -= step;
        if (std::fabs(step) < 1e-15) break;
    }
    double cE = std::cos(E), sE = std::sin(E);
    double b = std::sqrt(1.0 - e * e);
    double xo = a * (cE - e), yo = a * b * sE;
    double nmean = std::sqrt(mu / (a * a * a));
    double denom = 1.0 - e * cE;
    double vxo = -a * nmean * sE / denom, vyo = a * nmean * b * cE / denom;
    double cO = std::cos(node), sO = std::sin(node);
    double cw = std::cos(peri), sw = std::sin(peri);
    double ci = std::cos(inc), si = std::sin(inc);
    double r11 = cO * cw - sO * sw * ci, r12 = -cO * sw - sO * cw * ci;
    double r21 = sO * cw + cO * sw * ci, r22 = -sO * sw + cO * cw * ci;
    double r31 = sw * si, r32 = cw * si;
    r[0] = r11 * xo + r12 * yo; r[1] = r21 * xo + r22 * yo; r[2] = r31 * xo + r32 * yo;
    v[0] = r11 * vxo + r12 * vyo; v[1] = r21 * vxo + r22 * vyo; v[2] = r31 * vxo + r32 * vyo;
}

double rayleigh(SplitMix64& rng, double sigma) {
    return sigma * std::sqrt(-2.0 * std::log(1.0 - rng.uniform()));
}

void com_shift(int32_t n, const cm_soa_ptrs* out) {
    long double m = 0, c[3] = {0, 0, 0}, u[3] = {0, 0, 0};
    for (int32_t i = 0; i < n; i++) {
        long double mi = out->mass[i];
        m += mi;
        c[0] += mi * out->pos_x[i]; c[1] += mi * out->pos_y[i]; c[2] += mi * out->pos_z[i];
        u[0] += mi * out->vel_x[i]; u[1] += mi * out->vel_y[i]; u[2] += mi * out->vel_z[i];
    }
    double p[3], w[3];
    for (int k = 0; k < 3; k++) { p[k] = (double)(c[k] / m); w[k] = (double)(u[k] / m); }
    for (int32_t i = 0; i < n; i++) {
        out->pos_x[i] -= p[0]; out->pos_y[i] -= p[1]; out->pos_z[i] -= p[2];
        out->vel_x[i] -= w[0]; out->vel_y[i] -= w[1]; out->vel_z[i] -= w[2];
    }
}

}  // namespace

// Cold thin disk around a central star (the collision workload): body 0 is a
// 1 Msun star of the given radius (density consistent with that radius);
// bodies 1..n-1 have semi-major axes log-uniform in [a_min, a_max],
// Rayleigh eccentricity (sigma_e) and inclination (sigma_i), uniform angles,
// masses log-uniform in [m_min, m_max] and radii from the given density
// (Msun / AU^3) by DensityModel.ComputeRadius (Types/DensityModel.cs:60-66).
extern "C" int32_t cm_generate_disk(int32_t n, uint64_t seed, double star_radius,
                                    double a_min, double a_max, double sigma_e,
                                    double sigma_i, double m_min, double m_max,
                                    double density, const cm_soa_ptrs* out) {
    if (n < 1 || !out || !out->pos_x || !out->vel_x || !out->mass || !out->radius ||
        !out->density || !out->is_active || !out->is_collidable || !out->body_type)
        return CM_ERR_INVALID_ARGUMENT;
    if (!(a_min > 0.0) || !(a_max >= a_min) || !(m_min > 0.0) || !(m_max >= m_min) ||
        !(density > 0.0) || !(star_radius > 0.0))
        return CM_ERR_INVALID_ARGUMENT;
    SplitMix64 rng(seed);
    const double mstar = 1.0;
    out->pos_x[0] = out->pos_y[0] = out->pos_z[0] = 0.0;
    out->vel_x[0] = out->vel_y[0] = out->vel_z[0] = 0.0;
    out->mass[0] = mstar;
    out->radius[0] = star_radius;
    out->density[0] = 3.0 * mstar / (4.0 * kPi * star_radius * star_radius * star_radius);
    out->body_type[0] = 0;   // Star
    out->is_active[0] = 1;
    out->is_collidable[0] = 1;
    const double la = std::log(a_min), lb = std::log(a_max);
    const double lm0 = std::log(m_min), lm1 = std::log(m_max);
    for (int32_t i = 1; i < n; i++) {
        double a = std::exp(la + (lb - la) * rng.uniform());
        double e = std::min(rayleigh(rng, sigma_e), 0.9);
        double inc = rayleigh(rng, sigma_i);
        double node = 2.0 * kPi * rng.uniform();
        double peri = 2.0 * kPi * rng.uniform();
        double mean = 2.0 * kPi * rng.uniform();
        double m = std::exp(lm0 + (lm1 - lm0) * rng.uniform());
        double r[3], v[3];
        kepler_to_cartesian(mstar + m, a, e, inc, node, peri, mean, r, v);
        out->pos_x[i] = r[0]; out->pos_y[i] = r[1]; out->pos_z[i] = r[2];
        out->vel_x[i] = v[0]; out->vel_y[i] = v[1]; out->vel_z[i] = v[2];
        out->mass[i] = m;
        out->density[i] = density;
        out->radius[i] = std::cbrt(3.0 * (m / density) / (4.0 * kPi));
        out->body_type[i] = 5;   // Asteroid
        out->is_active[i] = 1;
        out->is_collidable[i] = 1;
    }
    com_shift(n, out);
    return CM_OK;
}

extern "C" int32_t cm_generate_ensemble(int32_t n_systems, uint64_t seed, int32_t n_min,
                                        int32_t n_max, double m_min, double m_max,
                                        double a_min, double a_max, double sigma_e,
                                        int32_t* offsets, const cm_soa_ptrs* out) {
    if (n_systems < 1 || n_min < 1 || n_max < n_min || !offsets || !(m_min > 0.0) ||
        !(m_max >= m_min) || !(a_min > 0.0) || !(a_max >= a_min))
        return CM_ERR_INVALID_ARGUMENT;
    // sizes from their own stream so that the offsets can be had alone
    SplitMix64 sizes(seed ^ 0x5EEDull);
    offsets[0] = 0;
    for (int32_t s = 0; s < n_systems; s++) {
        int32_t span = n_max - n_min + 1;
        int32_t nb = n_min + (int32_t)(sizes.next() % (uint64_t)span);
        offsets[s + 1] = offsets[s] + nb;
    }
    if (!out) return CM_OK;
    if (!out->pos_x || !out->pos_y || !out->pos_z || !out->vel_x || !out->vel_y ||
        !out->vel_z || !out->mass)
        return CM_ERR_INVALID_ARGUMENT;
    SplitMix64 rng(seed);
    const double la = std::log(a_min), lb = std::log(a_max);
    const double lm0 = std::log(m_min), lm1 = std::log(m_max);
    for (int32_t s = 0; s < n_systems; s++) {
        int32_t b0 = offsets[s], b1 = offsets[s + 1];
        out->pos_x[b0] = out->pos_y[b0] = out->pos_z[b0] = 0.0;
        out->vel_x[b0] = out->vel_y[b0] = out->vel_z[b0] = 0.0;
        out->mass[b0] = 1.0;
        for (int32_t i = b0 + 1; i < b1; i++) {
            double a = std::exp(la + (lb - la) * rng.uniform());
            double e = std::min(rayleigh(rng, sigma_e), 0.9);
            double inc = std::acos(2.0 * rng.uniform() - 1.0);
            double node = 2.0 * kPi * rng.uniform();
            double peri = 2.0 * kPi * rng.uniform();
            double mean = 2.0 * kPi * rng.uniform();
            double m = std::exp(lm0 + (lm1 - lm0) * rng.uniform());
            double r[3], v[3];
            kepler_to_cartesian(1.0 + m, a, e, inc, node, peri, mean, r, v);
            out->pos_x[i] = r[0]; out->pos_y[i] = r[1]; out->pos_z[i] = r[2];
            out->vel_x[i] = v[0]; out->vel_y[i] = v[1]; out->vel_z[i] = v[2];
            out->mass[i] = m;
        }
        cm_soa_ptrs sub = *out;
        sub.pos_x += b0; sub.pos_y += b0; sub.pos_z += b0;
        sub.vel_x += b0; sub.vel_y += b0; sub.vel_z += b0;
        sub.mass += b0;
        com_shift(b1 - b0, &sub);
        for (int32_t i = b0; i < b1; i++) {
            if (out->radius) out->radius[i] = 0.0;
            if (out->is_active) out->is_active[i] = 1;
        }
    }
    return CM_OK;
}
