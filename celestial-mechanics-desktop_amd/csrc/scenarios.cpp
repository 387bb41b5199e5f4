// scenarios.cpp -- deterministic synthetic initial conditions (host code).
//
// Random numbers come from one SplitMix64 stream per call (the mixing
// constants of the reference's DeterministicMode::deterministic_hash,
// engine/src/sim/deterministic.cpp:21-28); NextDouble = top 53 bits / 2^53.
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

constexpr double kPi = 3.14159265358979323846;

}  // namespace

extern "C" int32_t cm_generate_plummer(int32_t n, double total_mass,
                                       double scale_length, uint64_t seed,
                                       double mass_fraction_cut,
                                       int32_t com_frame,
                                       const cm_soa_ptrs* out) {
    if (n < 0 || !out || !out->pos_x || !out->pos_y || !out->pos_z ||
        !out->vel_x || !out->vel_y || !out->vel_z || !out->mass)
        return CM_ERR_INVALID_ARGUMENT;
    if (!(total_mass > 0.0) || !(scale_length > 0.0) ||
        !(mass_fraction_cut > 0.0) || mass_fraction_cut > 1.0)
        return CM_ERR_INVALID_ARGUMENT;
    if (n == 0) return CM_OK;
    SplitMix64 rng(seed);
    const double body_mass = total_mass / n;
    const double a = scale_length;
    // PlummerSphere.Generate (ScenarioGenerators/PlummerSphere.cs:41-116)
    for (int32_t i = 0; i < n; i++) {
        double u = rng.uniform();
        if (mass_fraction_cut < 1.0) u *= mass_fraction_cut;   // truncation
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
