// bh_device.h -- device code shared by the Barnes-Hut traversal (k_traverse.hip)
// and the learned far-field correction (k_mlp.hip): the cell acceptance test
// and the MLP's layout, bf16 rounding, MFMA tile step and input features.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "cm_internal.h"

namespace cm {

// ---- cell acceptance (BarnesHutBackend.cs:392-459 as a group test) --------
// kind: 0 skip (zero mass, BarnesHutBackend.cs:402), 1 accept the monopole,
// 2 open a leaf (its bodies), 3 open an internal node (its children).  The
// reference's (2h)^2 < theta^2 d^2 (BarnesHutBackend.cs:431-433) with d the
// distance from the node's centre of mass to the target group's box; leaves
// are approximated like internal nodes unless the shell theorem is on.
__device__ __forceinline__ int mac_kind(double4 c, int4 nd, double lo0, double lo1,
                                        double lo2, double hi0, double hi1, double hi2,
                                        double h, double theta2, bool shell) {
    if (c.w == 0.0) return 0;
    const int level = nd.y >> 8;
    const bool leaf = nd.x < 0;
    // distance to the box per axis: (lo - x) if x < lo, (x - hi) if x > hi,
    // else 0; written with max, which gives the same doubles (at most one of
    // the two terms is positive, and adding +0 changes nothing)
    double ddx = fmax(lo0 - c.x, 0.0) + fmax(c.x - hi0, 0.0);
    double ddy = fmax(lo1 - c.y, 0.0) + fmax(c.y - hi1, 0.0);
    double ddz = fmax(lo2 - c.z, 0.0) + fmax(c.z - hi2, 0.0);
    double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
    double hl = ldexp(h, -level);
    double size2 = 4.0 * hl * hl;
    if (!(shell && leaf) && d2 > 0.0 && size2 < theta2 * d2) return 1;
    return leaf ? 2 : 3;
}

// ---- MLP layout -------------------------------------------------------------
// 16 inputs -> 64 -> 64 -> 64 -> 16 outputs (8 used), ReLU on the hidden
// layers.  FP32 master parameters, row-major W[out][in] then b[out] per layer:
constexpr int kMlpIn = 16, kMlpHid = 64, kMlpOut = 16, kMlpUsed = 8;
constexpr int kPW1 = 0;
constexpr int kPB1 = kPW1 + kMlpHid * kMlpIn;
constexpr int kPW2 = kPB1 + kMlpHid;
constexpr int kPB2 = kPW2 + kMlpHid * kMlpHid;
constexpr int kPW3 = kPB2 + kMlpHid;
constexpr int kPB3 = kPW3 + kMlpHid * kMlpHid;
constexpr int kPW4 = kPB3 + kMlpHid;
constexpr int kPB4 = kPW4 + kMlpOut * kMlpHid;
constexpr int kMlpParams = kPB4 + kMlpOut;   // 10448
// bf16 copies of the weights, row-major [out][in] (wb) and [in][out] (wtb)
constexpr int kBW1 = 0;
constexpr int kBW2 = kBW1 + kMlpHid * kMlpIn;
constexpr int kBW3 = kBW2 + kMlpHid * kMlpHid;
constexpr int kBW4 = kBW3 + kMlpHid * kMlpHid;
constexpr int kMlpWeights = kBW4 + kMlpOut * kMlpHid;   // 10240

// float -> bf16, round to nearest even (inputs are finite)
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) {
    return __uint_as_float((uint32_t)b << 16);
}

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// One 16x16x16 bf16 MFMA step, fp32 accumulate.  Lane l supplies
// A[l % 16][4 (l / 16) + r] and B[4 (l / 16) + r][l % 16], r = 0..3, and
// receives C[4 (l / 16) + r][l % 16].
__device__ __forceinline__ v4f mfma16(v4s a, v4s b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// C (M x N) = A (M x K) * B (K x N) for one wavefront, all of M, N, K
// multiples of 16; A(m, k) = A[m * am + k * ak], B(k, n) = B[k * bk + n * bn]
// (bf16 in LDS or global memory); store(m, n, value) receives every element.
// Must be called by all 64 lanes.
template <class Store>
__device__ __forceinline__ void wave_gemm(int M, int N, int K, const uint16_t* A, int am,
                                          int ak, const uint16_t* B, int bk, int bn,
                                          Store store) {
    const int lane = threadIdx.x & (kWave - 1), lr = lane & 15, lg = lane >> 4;
    for (int mt = 0; mt < M; mt += 16) {
        for (int nt = 0; nt < N; nt += 16) {
            v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int kt = 0; kt < K; kt += 16) {
                v4s a, b;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    a[r] = (short)A[(mt + lr) * am + (kt + 4 * lg + r) * ak];
                    b[r] = (short)B[(kt + 4 * lg + r) * bk + (nt + lr) * bn];
                }
                acc = mfma16(a, b, acc);
            }
#pragma unroll
            for (int r = 0; r < 4; r++) store(mt + 4 * lg + r, nt + lr, acc[r]);
        }
    }
}

// ---- MLP inputs and output scales for one (group, cell) pair ---------------
// r: cell centre of mass minus the group's centre (float, as the split
// precision traversal forms it), M: cell mass, hl: cell half size, Q: second
// moments of the cell's bodies about its centre of mass (xx, yy, zz, xy, xz,
// yz).  Everything is expressed in the cell's frame F, the orthogonal map
// that takes u = r / |r| to +z (F = diag(1, 1, -s) (I - w w^T / (1 + |u_z|)),
// w = u + s z, s = sign(u_z)): in that frame the quadrupole part of the
// residual is linear in the rotated moments, which a small network fits far
// better than the bilinear form in the original axes.
// Inputs: hl / |r|, F Q F^T / (M hl^2) (xx, yy, zz, xy, xz, yz), zero padded.
// Outputs 0..2: F res, res = exact minus monopole acceleration at the group
// centre, in units of M hl^2 / |r|^4; outputs 3..7: F J F^T with
// J_ij = d res_i / d x_j (xx, yy, xy, xz, yz; zz = -xx - yy), in units of
// M hl^2 / |r|^5.  F is returned row-major in f[9].
__device__ __forceinline__ void mlp_frame(float ux, float uy, float uz, float f[9]) {
    const float s = uz >= 0.0f ? 1.0f : -1.0f;
    const float wz = uz + s;
    const float c = 1.0f / (1.0f + s * uz);
    // H = I - w w^T c, then the third row times -s
    f[0] = 1.0f - ux * ux * c; f[1] = -ux * uy * c;        f[2] = -ux * wz * c;
    f[3] = -uy * ux * c;        f[4] = 1.0f - uy * uy * c; f[5] = -uy * wz * c;
    f[6] = s * wz * ux * c;     f[7] = s * wz * uy * c;    f[8] = -s * (1.0f - wz * wz * c);
}

__device__ __forceinline__ void mlp_inputs(float rx, float ry, float rz, float M, float hl,
                                           const float* Q, float x[kMlpIn], float& sres,
                                           float& sgrad, float f[9]) {
    float d2 = rx * rx + ry * ry + rz * rz;
    float inv = __builtin_amdgcn_rsqf(d2);
    mlp_frame(rx * inv, ry * inv, rz * inv, f);
    float qn = 1.0f / (M * hl * hl);
    // q' = F q F^T with q symmetric
    const float q[9] = {Q[0], Q[3], Q[4], Q[3], Q[1], Q[5], Q[4], Q[5], Q[2]};
    float t[9];   // F q
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            t[3 * i + j] = f[3 * i] * q[j] + f[3 * i + 1] * q[3 + j] + f[3 * i + 2] * q[6 + j];
    auto qp = [&](int i, int j) {
        return (t[3 * i] * f[3 * j] + t[3 * i + 1] * f[3 * j + 1] + t[3 * i + 2] * f[3 * j + 2]) * qn;
    };
    x[0] = hl * inv;
    x[1] = qp(0, 0); x[2] = qp(1, 1); x[3] = qp(2, 2);
    x[4] = qp(0, 1); x[5] = qp(0, 2); x[6] = qp(1, 2);
#pragma unroll
    for (int k = 7; k < kMlpIn; k++) x[k] = 0.0f;
    sres = M * hl * hl * inv * inv * inv * inv;
    sgrad = sres * inv;
}

}  // namespace cm
