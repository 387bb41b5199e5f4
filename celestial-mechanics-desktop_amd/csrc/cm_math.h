// cm_math.h -- FP64 device math shared by the kernels: a correctly rounded
// cube root and DensityModel's body radius.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cm {

// Correctly rounded cube root by a fixed IEEE sequence (see DensityModel use).
__device__ inline double cube_residual(double c, double x) {
    double c2 = c * c;
    double e2 = fma(c, c, -c2);
    double p = c2 * c;
    double ep = fma(c2, c, -p);
    double qv = e2 * c;
    double eq = fma(e2, c, -qv);
    return ((p - x) + ep) + (qv + eq);
}

__device__ inline double cbrt_rn(double x) {
    if (x == 0.0 || !isfinite(x)) return x;
    bool neg = x < 0.0;
    double a = neg ? -x : x;
    int ex;
    double f = frexp(a, &ex);
    int q = ex >= 0 ? ex / 3 : -((-ex + 2) / 3);
    int r = ex - 3 * q;
    double t = ldexp(f, r);
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
    if (rh < rb) best = hi;
    return neg ? -best : best;
}

// DensityModel.ComputeBodyRadius (Types/DensityModel.cs:54-100).
__device__ inline double body_radius(double mass, double density, int32_t type) {
    if (type == 7) {   // BodyType.BlackHole: Schwarzschild radius in AU
        if (mass <= 0.0) return 0.0;
        double kg = mass * 1.989e30;
        double rm = 2.0 * 6.67430e-11 * kg / (2.998e8 * 2.998e8);
        return rm / 1.496e11;
    }
    if (mass <= 0.0 || density <= 0.0) return 0.0;
    double volume = mass / density;
    return cbrt_rn(3.0 * volume / (4.0 * 3.14159265358979323846));
}

}  // namespace cm
