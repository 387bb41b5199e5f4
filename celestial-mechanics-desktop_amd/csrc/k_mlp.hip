// k_mlp.hip -- the learned far-field correction: a small MLP that predicts,
// for one (target group, accepted cell) pair of the Barnes-Hut traversal, the
// monopole truncation residual of that cell at the group's centre and its
// gradient (layout and units in bh_device.h).  The traversal adds
// sum over cells (res + grad . (x - centre)) to every body of the group.
//
// Training data come from the engine's own tree: k_mlp_collect walks sampled
// target groups with the traversal's acceptance test and records every
// accepted cell of two or more bodies; k_mlp_targets sums the cell's bodies in
// FP64 (softened exactly like the traversal) to get the exact field and its
// gradient at the group centre and subtracts the monopole's.  Training is
// Adam on the mean squared error of the 8 used outputs, with the forward and
// backward passes on bf16 MFMA tiles (fp32 accumulation) in k_mlp_train and
// the update in k_mlp_adam.  Rounding points (reproduced by the CPU oracle):
// inputs, weights, hidden activations and the backward deltas dZ are rounded
// to bf16; biases, pre-activations, outputs, losses and gradients stay fp32;
// bias gradients sum the bf16 deltas.
//
// No reference counterpart: the reference's Barnes-Hut is monopole-only
// (BarnesHutBackend.cs:392-459).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "bh_device.h"
#include "cm_internal.h"

namespace cm {

struct Mlp {
    float* params = nullptr;      // FP32 master, kMlpParams
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    uint16_t* wb = nullptr;       // bf16 weights [out][in]
    uint16_t* wtb = nullptr;      // bf16 weights [in][out]
    float* grad_part = nullptr;   // per-wavefront gradients, part_cap x kMlpParams
    float* grad_sum = nullptr;    // kMlpParams
    float* loss_part = nullptr;   // part_cap
    int part_cap = 0;
    int2* pairs = nullptr;        // (group, node)
    double4* pair_c = nullptr;    // group centre
    float* px = nullptr;          // inputs, pair_cap x kMlpIn
    float* pt = nullptr;          // targets, pair_cap x kMlpUsed
    int32_t* idx = nullptr;       // batch order, pair_cap
    int32_t* gsel = nullptr;      // sampled groups
    int32_t* counter = nullptr;   // pairs found
    double* loss_acc = nullptr;   // summed batch losses of the running epoch
    int pair_cap = 0;
    int gsel_cap = 0;
    int n_pairs = 0;
    int n_sel = 0;                // groups in the last sample
    int step = 0;
    bool ready = false;           // parameters set
};

static Mlp* mlp_of(Engine& e) { return static_cast<Mlp*>(e.mlp); }

// ------------------------------------------------------------------ kernels --
__global__ void __launch_bounds__(256)
k_mlp_refresh(const float* __restrict__ params, uint16_t* __restrict__ wb,
              uint16_t* __restrict__ wtb) {
    int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= kMlpParams) return;
    int boff, in, q;
    if (p < kPB1) { q = p - kPW1; boff = kBW1; in = kMlpIn; }
    else if (p >= kPW2 && p < kPB2) { q = p - kPW2; boff = kBW2; in = kMlpHid; }
    else if (p >= kPW3 && p < kPB3) { q = p - kPW3; boff = kBW3; in = kMlpHid; }
    else if (p >= kPW4 && p < kPB4) { q = p - kPW4; boff = kBW4; in = kMlpHid; }
    else return;   // a bias
    int out = (boff == kBW4) ? kMlpOut : kMlpHid;
    int o = q / in, i = q % in;
    uint16_t b = f2bf(params[p]);
    wb[boff + q] = b;
    wtb[boff + i * out + o] = b;
}

// Forward pass of 64 samples (one wavefront) into LDS; returns nothing, the
// last layer's pre-activations are handed to `out(o, s, y)`.
template <class Out>
__device__ __forceinline__ void mlp_forward_wave(const uint16_t* wb, const float* params,
                                                 const uint16_t* sX, uint16_t* sA1,
                                                 uint16_t* sA2, uint16_t* sA3, Out out) {
    // layer 1: Z^T[o][s] = W1[o][i] X[s][i]; B(k = i, n = s) = sX[s * 16 + i]
    wave_gemm(kMlpHid, 64, kMlpIn, wb + kBW1, kMlpIn, 1, sX, 1, kMlpIn,
              [&](int o, int s, float z) {
                  float a = z + params[kPB1 + o];
                  sA1[o * 64 + s] = f2bf(a > 0.0f ? a : 0.0f);
              });
    __syncthreads();
    wave_gemm(kMlpHid, 64, kMlpHid, wb + kBW2, kMlpHid, 1, sA1, 64, 1,
              [&](int o, int s, float z) {
                  float a = z + params[kPB2 + o];
                  sA2[o * 64 + s] = f2bf(a > 0.0f ? a : 0.0f);
              });
    __syncthreads();
    wave_gemm(kMlpHid, 64, kMlpHid, wb + kBW3, kMlpHid, 1, sA2, 64, 1,
              [&](int o, int s, float z) {
                  float a = z + params[kPB3 + o];
                  sA3[o * 64 + s] = f2bf(a > 0.0f ? a : 0.0f);
              });
    __syncthreads();
    wave_gemm(kMlpOut, 64, kMlpHid, wb + kBW4, kMlpHid, 1, sA3, 64, 1,
              [&](int o, int s, float z) { out(o, s, z + params[kPB4 + o]); });
}

// Outputs for n samples given as fp32 inputs [n][16]; y is [n][16].
__global__ void __launch_bounds__(64)
k_mlp_forward(const uint16_t* __restrict__ wb, const float* __restrict__ params,
              const float* __restrict__ x, int n, float* __restrict__ y) {
    __shared__ uint16_t sX[64 * kMlpIn];
    __shared__ uint16_t sA[3][64 * 64];
    const int lane = threadIdx.x;
    const int base = blockIdx.x * 64;
    const int s = base + lane;
    for (int i = 0; i < kMlpIn; i++)
        sX[lane * kMlpIn + i] = s < n ? f2bf(x[(size_t)s * kMlpIn + i]) : (uint16_t)0;
    __syncthreads();
    mlp_forward_wave(wb, params, sX, sA[0], sA[1], sA[2], [&](int o, int sl, float v) {
        if (base + sl < n) y[(size_t)(base + sl) * kMlpOut + o] = v;
    });
}

// Forward + backward of up to 64 samples per wavefront (samples
// idx[blockIdx.x * 64 + lane]); writes this wavefront's gradient of
// L = inv_norm / 2 * sum (y - t)^2 over the used outputs (inv_norm = 2 /
// (8 * batch) gives the batch mean) and its sum of squared errors.
__global__ void __launch_bounds__(64)
k_mlp_train(const uint16_t* __restrict__ wb, const uint16_t* __restrict__ wtb,
            const float* __restrict__ params, const float* __restrict__ x,
            const float* __restrict__ t, const int32_t* __restrict__ idx, int n,
            float inv_norm, float* __restrict__ grad_part, float* __restrict__ loss_part) {
    __shared__ uint16_t sX[64 * kMlpIn];        // [s][i]
    __shared__ uint16_t sA[3][64 * 64];         // activations [feature][s]
    __shared__ uint16_t sD[2][64 * 64];         // deltas dZ [feature][s]
    __shared__ float sT[64 * kMlpUsed];
    const int lane = threadIdx.x;
    const int base = blockIdx.x * 64;
    const bool valid = base + lane < n;
    const int src = valid ? idx[base + lane] : 0;
    for (int i = 0; i < kMlpIn; i++)
        sX[lane * kMlpIn + i] = valid ? f2bf(x[(size_t)src * kMlpIn + i]) : (uint16_t)0;
    for (int o = 0; o < kMlpUsed; o++)
        sT[lane * kMlpUsed + o] = valid ? t[(size_t)src * kMlpUsed + o] : 0.0f;
    __syncthreads();
    float lsum = 0.0f;
    mlp_forward_wave(wb, params, sX, sA[0], sA[1], sA[2], [&](int o, int s, float y) {
        float g = 0.0f;
        if (o < kMlpUsed && base + s < n) {
            float d = y - sT[s * kMlpUsed + o];
            lsum += d * d;
            g = d * inv_norm;
        }
        sD[0][o * 64 + s] = f2bf(g);
    });
    __syncthreads();
    float* gp = grad_part + (size_t)blockIdx.x * kMlpParams;
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, kWave);
    if (lane == 0) loss_part[blockIdx.x] = lsum;

    // layer 4: dW4[o][i] = sum_s dY[o][s] A3[i][s]; db4[o] = sum_s dY[o][s]
    if (lane < kMlpOut) {
        float acc = 0.0f;
        for (int s = 0; s < 64; s++) acc += bf2f(sD[0][lane * 64 + s]);
        gp[kPB4 + lane] = acc;
    }
    wave_gemm(kMlpOut, kMlpHid, 64, sD[0], 64, 1, sA[2], 1, 64,
              [&](int o, int i, float v) { gp[kPW4 + o * kMlpHid + i] = v; });
    // dA3^T[i][s] = sum_o W4[o][i] dY[o][s], masked by the ReLU
    wave_gemm(kMlpHid, 64, kMlpOut, wtb + kBW4, kMlpOut, 1, sD[0], 64, 1,
              [&](int i, int s, float v) {
                  sD[1][i * 64 + s] = f2bf(bf2f(sA[2][i * 64 + s]) > 0.0f ? v : 0.0f);
              });
    __syncthreads();
    // layer 3
    {
        float acc = 0.0f;
        for (int s = 0; s < 64; s++) acc += bf2f(sD[1][lane * 64 + s]);
        gp[kPB3 + lane] = acc;
    }
    wave_gemm(kMlpHid, kMlpHid, 64, sD[1], 64, 1, sA[1], 1, 64,
              [&](int o, int i, float v) { gp[kPW3 + o * kMlpHid + i] = v; });
    wave_gemm(kMlpHid, 64, kMlpHid, wtb + kBW3, kMlpHid, 1, sD[1], 64, 1,
              [&](int i, int s, float v) {
                  sD[0][i * 64 + s] = f2bf(bf2f(sA[1][i * 64 + s]) > 0.0f ? v : 0.0f);
              });
    __syncthreads();
    // layer 2
    {
        float acc = 0.0f;
        for (int s = 0; s < 64; s++) acc += bf2f(sD[0][lane * 64 + s]);
        gp[kPB2 + lane] = acc;
    }
    wave_gemm(kMlpHid, kMlpHid, 64, sD[0], 64, 1, sA[0], 1, 64,
              [&](int o, int i, float v) { gp[kPW2 + o * kMlpHid + i] = v; });
    wave_gemm(kMlpHid, 64, kMlpHid, wtb + kBW2, kMlpHid, 1, sD[0], 64, 1,
              [&](int i, int s, float v) {
                  sD[1][i * 64 + s] = f2bf(bf2f(sA[0][i * 64 + s]) > 0.0f ? v : 0.0f);
              });
    __syncthreads();
    // layer 1: dW1[o][i] = sum_s dZ1[o][s] X[s][i]
    {
        float acc = 0.0f;
        for (int s = 0; s < 64; s++) acc += bf2f(sD[1][lane * 64 + s]);
        gp[kPB1 + lane] = acc;
    }
    wave_gemm(kMlpHid, kMlpIn, 64, sD[1], 64, 1, sX, kMlpIn, 1,
              [&](int o, int i, float v) { gp[kPW1 + o * kMlpIn + i] = v; });
}

// Sums the per-wavefront gradients in order; with update != 0 applies one
// Adam step (beta1 0.9, beta2 0.999, eps 1e-8) and refreshes the bf16 copies.
__global__ void __launch_bounds__(256)
k_mlp_adam(const float* __restrict__ grad_part, int nparts, float* __restrict__ params,
           float* __restrict__ m, float* __restrict__ v, float lr, float bc1, float bc2,
           int update, float* __restrict__ grad_sum, const float* __restrict__ loss_part,
           float loss_scale, double* __restrict__ loss_acc) {
    int p = blockIdx.x * 256 + threadIdx.x;
    if (p == 0 && loss_acc) {
        // running sum of batch losses (mean squared error of each batch)
        double l = 0.0;
        for (int w = 0; w < nparts; w++) l += (double)loss_part[w];
        *loss_acc += l * (double)loss_scale;
    }
    if (p >= kMlpParams) return;
    float g = 0.0f;
    for (int w = 0; w < nparts; w++) g += grad_part[(size_t)w * kMlpParams + p];
    grad_sum[p] = g;
    if (!update) return;
    float mm = 0.9f * m[p] + 0.1f * g;
    float vv = 0.999f * v[p] + 0.001f * g * g;
    m[p] = mm;
    v[p] = vv;
    params[p] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + 1e-8f);
}

// Target groups of the sample: accepted cells of >= 2 bodies, recorded with
// the group centre (the traversal's acceptance test and centre).
__global__ void __launch_bounds__(kWave)
k_mlp_collect(const TreeHeader* __restrict__ hdr, const int4* __restrict__ info,
              const double4* __restrict__ cm, const int2* __restrict__ groups,
              const int32_t* __restrict__ gsel, int nsel, const double* __restrict__ sorted,
              int cap, int levels, double theta2, int shell, int2* __restrict__ pairs,
              double4* __restrict__ pair_c, int pair_cap, int32_t* __restrict__ counter) {
    __shared__ int s_stack[kStackCap];
    __shared__ int s_base;
    const int lane = threadIdx.x;
    const double h = hdr->box[3];
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const double big = 1.7976931348623157e308;
    for (int sel = blockIdx.x; sel < nsel; sel += gridDim.x) {
        const int g = gsel[sel];
        const int2 gr = groups[g];
        const int my = gr.x + lane;
        const bool member = my < gr.y;
        double xi = member ? sorted[my] : 0.0;
        double yi = member ? sorted[cap + my] : 0.0;
        double zi = member ? sorted[2 * (size_t)cap + my] : 0.0;
        double lo0 = member ? xi : big, lo1 = member ? yi : big, lo2 = member ? zi : big;
        double hi0 = member ? xi : -big, hi1 = member ? yi : -big, hi2 = member ? zi : -big;
        for (int o = 32; o > 0; o >>= 1) {
            double t;
            t = __shfl_xor(lo0, o, kWave); lo0 = t < lo0 ? t : lo0;
            t = __shfl_xor(lo1, o, kWave); lo1 = t < lo1 ? t : lo1;
            t = __shfl_xor(lo2, o, kWave); lo2 = t < lo2 ? t : lo2;
            t = __shfl_xor(hi0, o, kWave); hi0 = t > hi0 ? t : hi0;
            t = __shfl_xor(hi1, o, kWave); hi1 = t > hi1 ? t : hi1;
            t = __shfl_xor(hi2, o, kWave); hi2 = t > hi2 ? t : hi2;
        }
        const double4 centre = make_double4((lo0 + hi0) * 0.5, (lo1 + hi1) * 0.5,
                                            (lo2 + hi2) * 0.5, 0.0);
        // a forced single pop can still push 8 children per level below the
        // limit, so one path's worth of slots stays free
        const int reserve = 8 * (levels + 2);
        int S = 1;
        if (lane == 0) s_stack[0] = 0;
        __syncthreads();
        while (S > 0) {
            int lim = (kStackCap - reserve - S) / 7;
            if (lim < 1) lim = 1;
            int take = S < kWave ? S : kWave;
            if (take > lim) take = lim;
            const int b = S - take;
            const int node = lane < take ? s_stack[b + lane] : -1;
            S = b;
            __syncthreads();
            int kind = 0;
            int4 nd = make_int4(-1, 0, 0, 0);
            if (node >= 0) {
                nd = info[node];
                kind = mac_kind(cm[node], nd, lo0, lo1, lo2, hi0, hi1, hi2, h, theta2,
                                shell != 0);
            }
            const bool rec = kind == 1 && nd.w >= 2;
            const uint64_t rm = __ballot(rec);
            if (lane == 0 && rm) s_base = atomicAdd(counter, __popcll(rm));
            __syncthreads();
            if (rec) {
                int p = s_base + __popcll(rm & lt);
                if (p < pair_cap) {
                    pairs[p] = make_int2(sel, node);
                    pair_c[p] = centre;
                }
            }
            int nop = kind == 3 ? (nd.y & 255) : 0;
            int incl = nop;
            for (int o = 1; o < kWave; o <<= 1) {
                int tt = __shfl_up(incl, o, kWave);
                if (lane >= o) incl += tt;
            }
            int tot = __shfl(incl, kWave - 1, kWave);
            if (kind == 3)
                for (int c = 0; c < nop; c++) s_stack[S + incl - nop + c] = nd.x + c;
            S += tot;
            __syncthreads();
        }
    }
}

// Inputs and FP64 targets of every collected pair (one wavefront per pair).
__global__ void __launch_bounds__(kWave)
k_mlp_targets(const TreeHeader* __restrict__ hdr, const int4* __restrict__ info,
              const double4* __restrict__ cm, const float* __restrict__ quad,
              const int2* __restrict__ pairs, const double4* __restrict__ pair_c, int np,
              const double* __restrict__ sorted, int cap, double eps2,
              float* __restrict__ px, float* __restrict__ pt) {
    const int lane = threadIdx.x;
    const double h = hdr->box[3];
    for (int p = blockIdx.x; p < np; p += gridDim.x) {
        const int node = pairs[p].y;
        const double4 g = pair_c[p];
        const int4 nd = info[node];
        // exact field and gradient at g from the cell's bodies
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        double gxx = 0.0, gyy = 0.0, gzz = 0.0, gxy = 0.0, gxz = 0.0, gyz = 0.0;
        for (int j = nd.z + lane; j < nd.z + nd.w; j += kWave) {
            double rx = sorted[j] - g.x, ry = sorted[cap + j] - g.y;
            double rz = sorted[2 * (size_t)cap + j] - g.z, m = sorted[3 * (size_t)cap + j];
            double D = rx * rx + ry * ry + rz * rz + eps2;
            double i1 = 1.0 / sqrt(D);
            double i3 = i1 * i1 * i1, i5 = i3 * i1 * i1;
            a0 += m * rx * i3; a1 += m * ry * i3; a2 += m * rz * i3;
            gxx += m * (3.0 * rx * rx * i5 - i3);
            gyy += m * (3.0 * ry * ry * i5 - i3);
            gzz += m * (3.0 * rz * rz * i5 - i3);
            gxy += m * 3.0 * rx * ry * i5;
            gxz += m * 3.0 * rx * rz * i5;
            gyz += m * 3.0 * ry * rz * i5;
        }
        for (int o = 32; o > 0; o >>= 1) {
            a0 += __shfl_xor(a0, o, kWave); a1 += __shfl_xor(a1, o, kWave);
            a2 += __shfl_xor(a2, o, kWave);
            gxx += __shfl_xor(gxx, o, kWave); gyy += __shfl_xor(gyy, o, kWave);
            gzz += __shfl_xor(gzz, o, kWave); gxy += __shfl_xor(gxy, o, kWave);
            gxz += __shfl_xor(gxz, o, kWave); gyz += __shfl_xor(gyz, o, kWave);
        }
        if (lane != 0) continue;
        const double4 c = cm[node];
        double rx = c.x - g.x, ry = c.y - g.y, rz = c.z - g.z;
        double D = rx * rx + ry * ry + rz * rz + eps2;
        double i1 = 1.0 / sqrt(D);
        double i3 = i1 * i1 * i1, i5 = i3 * i1 * i1;
        double M = c.w;
        a0 -= M * rx * i3; a1 -= M * ry * i3; a2 -= M * rz * i3;
        gxx -= M * (3.0 * rx * rx * i5 - i3);
        gyy -= M * (3.0 * ry * ry * i5 - i3);
        gzz -= M * (3.0 * rz * rz * i5 - i3);
        gxy -= M * 3.0 * rx * ry * i5;
        gxz -= M * 3.0 * rx * rz * i5;
        gyz -= M * 3.0 * ry * rz * i5;
        const double hl = ldexp(h, -(nd.y >> 8));
        float x[kMlpIn], sres, sgrad, f[9];
        mlp_inputs((float)rx, (float)ry, (float)rz, (float)M, (float)hl,
                   quad + 6 * (size_t)node, x, sres, sgrad, f);
        for (int i = 0; i < kMlpIn; i++) px[(size_t)p * kMlpIn + i] = x[i];
        // targets in the cell's frame: F res and F J F^T
        const double R[3] = {a0, a1, a2};
        const double J[9] = {gxx, gxy, gxz, gxy, gyy, gyz, gxz, gyz, gzz};
        double FR[3], FJ[9];
        for (int i = 0; i < 3; i++) {
            FR[i] = f[3 * i] * R[0] + f[3 * i + 1] * R[1] + f[3 * i + 2] * R[2];
            for (int j = 0; j < 3; j++)
                FJ[3 * i + j] = f[3 * i] * J[j] + f[3 * i + 1] * J[3 + j] + f[3 * i + 2] * J[6 + j];
        }
        auto jp = [&](int i, int j) {
            return FJ[3 * i] * f[3 * j] + FJ[3 * i + 1] * f[3 * j + 1] + FJ[3 * i + 2] * f[3 * j + 2];
        };
        float* tp = pt + (size_t)p * kMlpUsed;
        tp[0] = (float)(FR[0] / sres); tp[1] = (float)(FR[1] / sres); tp[2] = (float)(FR[2] / sres);
        tp[3] = (float)(jp(0, 0) / sgrad); tp[4] = (float)(jp(1, 1) / sgrad);
        tp[5] = (float)(jp(0, 1) / sgrad); tp[6] = (float)(jp(0, 2) / sgrad);
        tp[7] = (float)(jp(1, 2) / sgrad);
    }
}

// ------------------------------------------------------------------- host ---
static int ensure_mlp(Engine& e) {
    if (e.mlp) return CM_OK;
    Mlp* m = new (std::nothrow) Mlp();
    if (!m) return set_error(e, CM_ERR_OUT_OF_MEMORY, "mlp state");
    e.mlp = m;
    bool ok = true;
    auto alloc = [&](void** p, size_t bytes) {
        if (ok && hipMalloc(p, bytes) != hipSuccess) ok = false;
    };
    alloc((void**)&m->params, kMlpParams * 4);
    alloc((void**)&m->adam_m, kMlpParams * 4);
    alloc((void**)&m->adam_v, kMlpParams * 4);
    alloc((void**)&m->wb, kMlpWeights * 2);
    alloc((void**)&m->wtb, kMlpWeights * 2);
    alloc((void**)&m->grad_sum, kMlpParams * 4);
    alloc((void**)&m->counter, 4);
    alloc((void**)&m->loss_acc, sizeof(double));
    if (!ok) return set_error(e, CM_ERR_OUT_OF_MEMORY, "mlp buffers");
    return CM_OK;
}

static int ensure_parts(Engine& e, int nparts) {
    Mlp* m = mlp_of(e);
    if (nparts <= m->part_cap) return CM_OK;
    if (m->grad_part) (void)hipFree(m->grad_part);
    if (m->loss_part) (void)hipFree(m->loss_part);
    m->grad_part = nullptr;
    m->loss_part = nullptr;
    m->part_cap = 0;
    if (hipMalloc((void**)&m->grad_part, (size_t)nparts * kMlpParams * 4) != hipSuccess ||
        hipMalloc((void**)&m->loss_part, (size_t)nparts * 4) != hipSuccess)
        return set_error(e, CM_ERR_OUT_OF_MEMORY, "mlp gradient partials");
    m->part_cap = nparts;
    return CM_OK;
}

static int ensure_pairs(Engine& e, int cap, int nsel) {
    Mlp* m = mlp_of(e);
    if (cap > m->pair_cap) {
        void* ptrs[] = {m->pairs, m->pair_c, m->px, m->pt, m->idx};
        for (void* p : ptrs) if (p) (void)hipFree(p);
        m->pairs = nullptr; m->pair_c = nullptr; m->px = nullptr; m->pt = nullptr;
        m->idx = nullptr; m->pair_cap = 0;
        bool ok = hipMalloc((void**)&m->pairs, (size_t)cap * sizeof(int2)) == hipSuccess &&
                  hipMalloc((void**)&m->pair_c, (size_t)cap * sizeof(double4)) == hipSuccess &&
                  hipMalloc((void**)&m->px, (size_t)cap * kMlpIn * 4) == hipSuccess &&
                  hipMalloc((void**)&m->pt, (size_t)cap * kMlpUsed * 4) == hipSuccess &&
                  hipMalloc((void**)&m->idx, (size_t)cap * 4) == hipSuccess;
        if (!ok) return set_error(e, CM_ERR_OUT_OF_MEMORY, "mlp pair buffers");
        m->pair_cap = cap;
    }
    if (nsel > m->gsel_cap) {
        if (m->gsel) (void)hipFree(m->gsel);
        m->gsel = nullptr;
        m->gsel_cap = 0;
        if (hipMalloc((void**)&m->gsel, (size_t)nsel * 4) != hipSuccess)
            return set_error(e, CM_ERR_OUT_OF_MEMORY, "mlp group sample");
        m->gsel_cap = nsel;
    }
    return CM_OK;
}

void mlp_free(Engine& e) {
    Mlp* m = mlp_of(e);
    if (!m) return;
    void* ptrs[] = {m->params, m->adam_m, m->adam_v, m->wb, m->wtb, m->grad_part,
                    m->grad_sum, m->loss_part, m->pairs, m->pair_c, m->px, m->pt,
                    m->idx, m->gsel, m->counter, m->loss_acc};
    for (void* p : ptrs) if (p) (void)hipFree(p);
    delete m;
    e.mlp = nullptr;
}

bool mlp_ready(const Engine& e) {
    return e.mlp && static_cast<const Mlp*>(e.mlp)->ready;
}

const uint16_t* mlp_weights(const Engine& e) { return static_cast<const Mlp*>(e.mlp)->wb; }
const float* mlp_params(const Engine& e) { return static_cast<const Mlp*>(e.mlp)->params; }

static uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double uniform01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }

static int upload_params(Engine& e, const float* host) {
    Mlp* m = mlp_of(e);
    CM_HIP_TRY(e, hipMemcpyAsync(m->params, host, kMlpParams * 4, hipMemcpyHostToDevice,
                                 e.stream));
    CM_HIP_TRY(e, hipMemsetAsync(m->adam_m, 0, kMlpParams * 4, e.stream));
    CM_HIP_TRY(e, hipMemsetAsync(m->adam_v, 0, kMlpParams * 4, e.stream));
    hipLaunchKernelGGL(k_mlp_refresh, dim3((kMlpParams + 255) / 256), dim3(256), 0,
                       e.stream, m->params, m->wb, m->wtb);
    CM_HIP_TRY(e, hipGetLastError());
    CM_HIP_TRY(e, hipStreamSynchronize(e.stream));
    m->step = 0;
    m->ready = true;
    return CM_OK;
}

// one gradient (and optionally Adam step) over samples idx[0 .. n); the batch
// loss is returned in *loss (a host sync) and/or added to the device sum
// *loss_acc (no sync)
static int train_batch(Engine& e, const float* x, const float* t, const int32_t* idx,
                       int n, float lr, int update, float* loss, double* loss_acc = nullptr) {
    Mlp* m = mlp_of(e);
    int nparts = (n + 63) / 64;
    int rc = ensure_parts(e, nparts);
    if (rc != CM_OK) return rc;
    float inv_norm = 2.0f / (float)(kMlpUsed * n);
    hipLaunchKernelGGL(k_mlp_train, dim3(nparts), dim3(64), 0, e.stream, m->wb, m->wtb,
                       m->params, x, t, idx, n, inv_norm, m->grad_part, m->loss_part);
    float bc1 = 1.0f, bc2 = 1.0f;
    if (update) {
        m->step++;
        bc1 = 1.0f - std::pow(0.9f, (float)m->step);
        bc2 = 1.0f - std::pow(0.999f, (float)m->step);
    }
    hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, e.stream,
                       m->grad_part, nparts, m->params, m->adam_m, m->adam_v, lr, bc1, bc2,
                       update, m->grad_sum, m->loss_part, 1.0f / (float)(kMlpUsed * n),
                       loss_acc);
    if (update)
        hipLaunchKernelGGL(k_mlp_refresh, dim3((kMlpParams + 255) / 256), dim3(256), 0,
                           e.stream, m->params, m->wb, m->wtb);
    CM_HIP_TRY(e, hipGetLastError());
    if (loss) {
        std::vector<float> lp(nparts);
        CM_HIP_TRY(e, hipMemcpyAsync(lp.data(), m->loss_part, (size_t)nparts * 4,
                                     hipMemcpyDeviceToHost, e.stream));
        CM_HIP_TRY(e, hipStreamSynchronize(e.stream));
        double s = 0.0;
        for (float v : lp) s += v;
        *loss = (float)(s / (double)(kMlpUsed * n));
    }
    return CM_OK;
}

}  // namespace cm

using cm::Engine;

extern "C" {

int32_t cm_mlp_init(void* h, uint64_t seed) {
    Engine* e = static_cast<Engine*>(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    int rc = cm::ensure_mlp(*e);
    if (rc != CM_OK) return rc;
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    // He-uniform weights, zero biases
    std::vector<float> p(cm::kMlpParams, 0.0f);
    uint64_t s = seed;
    auto fill = [&](int off, int out, int in) {
        double bound = std::sqrt(6.0 / in);
        for (int k = 0; k < out * in; k++)
            p[off + k] = (float)((2.0 * cm::uniform01(s) - 1.0) * bound);
    };
    fill(cm::kPW1, cm::kMlpHid, cm::kMlpIn);
    fill(cm::kPW2, cm::kMlpHid, cm::kMlpHid);
    fill(cm::kPW3, cm::kMlpHid, cm::kMlpHid);
    fill(cm::kPW4, cm::kMlpOut, cm::kMlpHid);
    for (int k = 0; k < cm::kMlpOut * cm::kMlpHid; k++) p[cm::kPW4 + k] *= 0.1f;
    return cm::upload_params(*e, p.data());
}

int32_t cm_mlp_set_params(void* h, const float* params, int32_t n) {
    Engine* e = static_cast<Engine*>(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!params || n != cm::kMlpParams)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "params must hold 10448 floats");
    int rc = cm::ensure_mlp(*e);
    if (rc != CM_OK) return rc;
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    return cm::upload_params(*e, params);
}

int32_t cm_mlp_get_params(void* h, float* params, int32_t n) {
    Engine* e = static_cast<Engine*>(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!params || n != cm::kMlpParams)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "params must hold 10448 floats");
    if (!cm::mlp_ready(*e)) return cm::set_error(*e, CM_ERR_NOT_INITIALIZED, "no network");
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    CM_HIP_TRY(*e, hipMemcpyAsync(params, cm::mlp_of(*e)->params, cm::kMlpParams * 4,
                                  hipMemcpyDeviceToHost, e->stream));
    CM_HIP_TRY(*e, hipStreamSynchronize(e->stream));
    return CM_OK;
}

int32_t cm_mlp_forward(void* h, const float* x, int32_t n, float* y) {
    Engine* e = static_cast<Engine*>(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!x || !y || n <= 0) return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "bad arrays");
    if (!cm::mlp_ready(*e)) return cm::set_error(*e, CM_ERR_NOT_INITIALIZED, "no network");
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    cm::Mlp* m = cm::mlp_of(*e);
    float *dx = nullptr, *dy = nullptr;
    if (hipMalloc((void**)&dx, (size_t)n * cm::kMlpIn * 4) != hipSuccess ||
        hipMalloc((void**)&dy, (size_t)n * cm::kMlpOut * 4) != hipSuccess) {
        if (dx) (void)hipFree(dx);
        return cm::set_error(*e, CM_ERR_OUT_OF_MEMORY, "forward buffers");
    }
    int rc = CM_OK;
    if (hipMemcpyAsync(dx, x, (size_t)n * cm::kMlpIn * 4, hipMemcpyHostToDevice,
                       e->stream) != hipSuccess)
        rc = cm::set_error(*e, CM_ERR_HIP, "copy in");
    if (rc == CM_OK) {
        hipLaunchKernelGGL(cm::k_mlp_forward, dim3((n + 63) / 64), dim3(64), 0, e->stream,
                           m->wb, m->params, dx, n, dy);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(y, dy, (size_t)n * cm::kMlpOut * 4, hipMemcpyDeviceToHost,
                           e->stream) != hipSuccess ||
            hipStreamSynchronize(e->stream) != hipSuccess)
            rc = cm::set_error(*e, CM_ERR_HIP, "mlp forward");
    }
    (void)hipFree(dx);
    (void)hipFree(dy);
    return rc;
}

int32_t cm_mlp_gradient(void* h, const float* x, const float* t, int32_t n, float* grad,
                        float* loss) {
    Engine* e = static_cast<Engine*>(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!x || !t || !grad || n <= 0)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "bad arrays");
    if (!cm::mlp_ready(*e)) return cm::set_error(*e, CM_ERR_NOT_INITIALIZED, "no network");
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    cm::Mlp* m = cm::mlp_of(*e);
    float *dx = nullptr, *dt = nullptr;
    int32_t* di = nullptr;
    bool ok = hipMalloc((void**)&dx, (size_t)n * cm::kMlpIn * 4) == hipSuccess &&
              hipMalloc((void**)&dt, (size_t)n * cm::kMlpUsed * 4) == hipSuccess &&
              hipMalloc((void**)&di, (size_t)n * 4) == hipSuccess;
    int rc = ok ? CM_OK : cm::set_error(*e, CM_ERR_OUT_OF_MEMORY, "gradient buffers");
    if (rc == CM_OK) {
        std::vector<int32_t> id(n);
        for (int k = 0; k < n; k++) id[k] = k;
        if (hipMemcpyAsync(dx, x, (size_t)n * cm::kMlpIn * 4, hipMemcpyHostToDevice, e->stream) != hipSuccess ||
            hipMemcpyAsync(dt, t, (size_t)n * cm::kMlpUsed * 4, hipMemcpyHostToDevice, e->stream) != hipSuccess ||
            hipMemcpyAsync(di, id.data(), (size_t)n * 4, hipMemcpyHostToDevice, e->stream) != hipSuccess)
            rc = cm::set_error(*e, CM_ERR_HIP, "copy in");
    }
    if (rc == CM_OK) rc = cm::train_batch(*e, dx, dt, di, n, 0.0f, 0, loss);
    if (rc == CM_OK) {
        if (hipMemcpyAsync(grad, m->grad_sum, cm::kMlpParams * 4, hipMemcpyDeviceToHost,
                           e->stream) != hipSuccess ||
            hipStreamSynchronize(e->stream) != hipSuccess)
            rc = cm::set_error(*e, CM_ERR_HIP, "copy out");
    }
    if (dx) (void)hipFree(dx);
    if (dt) (void)hipFree(dt);
    if (di) (void)hipFree(di);
    return rc;
}

int32_t cm_mlp_collect(void* h, double softening, int32_t max_groups, uint64_t seed,
                       int32_t max_pairs, int32_t* n_pairs) {
    Engine* e = static_cast<Engine*>(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!n_pairs || max_groups <= 0 || max_pairs <= 0)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "bad sample sizes");
    if (e->cfg.kind != CM_BACKEND_BARNES_HUT)
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "needs the Barnes-Hut backend");
    if (e->count == 0) return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "no bodies");
    int rc = cm::ensure_mlp(*e);
    if (rc != CM_OK) return rc;
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    // tree with second moments, target groups
    const int saved = e->cfg.learned_correction;
    e->cfg.learned_correction = 1;
    const void* keys = nullptr;
    const int32_t* order = nullptr;
    rc = cm::build_tree(*e, &keys, &order);
    e->cfg.learned_correction = saved;
    if (rc != CM_OK) return rc;
    rc = cm::form_groups(*e, order);
    if (rc != CM_OK) return rc;
    CM_HIP_TRY(*e, hipMemcpyAsync(e->hdr_host, e->hdr, sizeof(cm::TreeHeader),
                                  hipMemcpyDeviceToHost, e->stream));
    CM_HIP_TRY(*e, hipStreamSynchronize(e->stream));
    if (e->hdr_host->overflow)
        return cm::set_error(*e, CM_ERR_POOL_EXHAUSTED, "node pool exhausted");
    const int ng = e->hdr_host->n_groups;
    // groups are emitted in a device-dependent order: rank them by first
    // body, then draw a seeded sample (without replacement) in that order
    std::vector<int2> gr(ng);
    CM_HIP_TRY(*e, hipMemcpy(gr.data(), e->groups, (size_t)ng * sizeof(int2),
                             hipMemcpyDeviceToHost));
    std::vector<int32_t> all(ng);
    for (int g = 0; g < ng; g++) all[g] = g;
    std::sort(all.begin(), all.end(), [&](int a, int b) { return gr[a].x < gr[b].x; });
    std::vector<int32_t> sel;
    uint64_t s = seed;
    if (max_groups >= ng) {
        sel = all;
    } else {
        for (int k = 0; k < max_groups; k++) {
            int j = k + (int)(cm::splitmix64(s) % (uint64_t)(ng - k));
            std::swap(all[k], all[j]);
            sel.push_back(all[k]);
        }
    }
    rc = cm::ensure_pairs(*e, max_pairs, (int)sel.size());
    if (rc != CM_OK) return rc;
    cm::Mlp* m = cm::mlp_of(*e);
    CM_HIP_TRY(*e, hipMemcpyAsync(m->gsel, sel.data(), sel.size() * 4,
                                  hipMemcpyHostToDevice, e->stream));
    CM_HIP_TRY(*e, hipMemsetAsync(m->counter, 0, 4, e->stream));
    double th = e->cfg.theta > 0.2 ? e->cfg.theta : 0.2;
    int nsel = (int)sel.size();
    hipLaunchKernelGGL(cm::k_mlp_collect, dim3(std::min(nsel, 4096)), dim3(cm::kWave), 0,
                       e->stream, e->hdr, e->nd_info, e->nd_cm, e->groups, m->gsel, nsel,
                       e->sorted, e->capacity, e->cfg.morton_levels, th * th,
                       e->cfg.shell_theorem, m->pairs,
                       m->pair_c, max_pairs, m->counter);
    CM_HIP_TRY(*e, hipGetLastError());
    int32_t found = 0;
    CM_HIP_TRY(*e, hipMemcpyAsync(&found, m->counter, 4, hipMemcpyDeviceToHost, e->stream));
    CM_HIP_TRY(*e, hipStreamSynchronize(e->stream));
    int np = std::min(found, max_pairs);
    if (np > 0) {
        // pairs arrive in a device-dependent order: sort by (sample, node)
        std::vector<int2> pr(np);
        std::vector<double4> pc(np);
        CM_HIP_TRY(*e, hipMemcpy(pr.data(), m->pairs, (size_t)np * sizeof(int2),
                                 hipMemcpyDeviceToHost));
        CM_HIP_TRY(*e, hipMemcpy(pc.data(), m->pair_c, (size_t)np * sizeof(double4),
                                 hipMemcpyDeviceToHost));
        std::vector<int32_t> ord(np);
        for (int k = 0; k < np; k++) ord[k] = k;
        std::sort(ord.begin(), ord.end(), [&](int a, int b) {
            return pr[a].x != pr[b].x ? pr[a].x < pr[b].x : pr[a].y < pr[b].y;
        });
        std::vector<int2> pr2(np);
        std::vector<double4> pc2(np);
        for (int k = 0; k < np; k++) { pr2[k] = pr[ord[k]]; pc2[k] = pc[ord[k]]; }
        CM_HIP_TRY(*e, hipMemcpy(m->pairs, pr2.data(), (size_t)np * sizeof(int2),
                                 hipMemcpyHostToDevice));
        CM_HIP_TRY(*e, hipMemcpy(m->pair_c, pc2.data(), (size_t)np * sizeof(double4),
                                 hipMemcpyHostToDevice));
        hipLaunchKernelGGL(cm::k_mlp_targets, dim3(std::min(np, 16384)), dim3(cm::kWave), 0,
                           e->stream, e->hdr, e->nd_info, e->nd_cm, e->nd_quad, m->pairs,
                           m->pair_c, np, e->sorted, e->capacity, softening * softening,
                           m->px, m->pt);
        CM_HIP_TRY(*e, hipGetLastError());
        CM_HIP_TRY(*e, hipStreamSynchronize(e->stream));
    }
    m->n_pairs = np;
    m->n_sel = nsel;
    *n_pairs = np;
    return CM_OK;
}

int32_t cm_mlp_get_pairs(void* h, float* x, float* t, int32_t* pair_info, double* centre,
                         int32_t cap) {
    Engine* e = static_cast<Engine*>(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (!e->mlp) return cm::set_error(*e, CM_ERR_NOT_INITIALIZED, "nothing collected");
    cm::Mlp* m = cm::mlp_of(*e);
    int n = std::min(cap, m->n_pairs);
    if (n <= 0) return CM_OK;
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    if (x)
        CM_HIP_TRY(*e, hipMemcpyAsync(x, m->px, (size_t)n * cm::kMlpIn * 4,
                                      hipMemcpyDeviceToHost, e->stream));
    if (t)
        CM_HIP_TRY(*e, hipMemcpyAsync(t, m->pt, (size_t)n * cm::kMlpUsed * 4,
                                      hipMemcpyDeviceToHost, e->stream));
    if (pair_info) {
        std::vector<int2> pr(n), gr(e->hdr_host->n_groups);
        std::vector<int32_t> gs(m->n_sel);
        CM_HIP_TRY(*e, hipMemcpy(pr.data(), m->pairs, (size_t)n * sizeof(int2),
                                 hipMemcpyDeviceToHost));
        CM_HIP_TRY(*e, hipMemcpy(gs.data(), m->gsel, gs.size() * 4, hipMemcpyDeviceToHost));
        CM_HIP_TRY(*e, hipMemcpy(gr.data(), e->groups, gr.size() * sizeof(int2),
                                 hipMemcpyDeviceToHost));
        for (int k = 0; k < n; k++) {
            const int2 g = gr[gs[pr[k].x]];
            pair_info[4 * k] = pr[k].x;
            pair_info[4 * k + 1] = pr[k].y;
            pair_info[4 * k + 2] = g.x;
            pair_info[4 * k + 3] = g.y;
        }
    }
    if (centre) {
        std::vector<double4> c(n);
        CM_HIP_TRY(*e, hipMemcpyAsync(c.data(), m->pair_c, (size_t)n * sizeof(double4),
                                      hipMemcpyDeviceToHost, e->stream));
        CM_HIP_TRY(*e, hipStreamSynchronize(e->stream));
        for (int k = 0; k < n; k++) {
            centre[3 * k] = c[k].x; centre[3 * k + 1] = c[k].y; centre[3 * k + 2] = c[k].z;
        }
    }
    CM_HIP_TRY(*e, hipStreamSynchronize(e->stream));
    return CM_OK;
}

int32_t cm_mlp_train(void* h, int32_t epochs, int32_t batch, float lr, uint64_t seed,
                     float* loss_per_epoch) {
    Engine* e = static_cast<Engine*>(h);
    if (!e) return CM_ERR_NOT_INITIALIZED;
    if (epochs <= 0 || batch <= 0 || !(lr > 0.0f))
        return cm::set_error(*e, CM_ERR_INVALID_ARGUMENT, "bad training parameters");
    if (!cm::mlp_ready(*e)) return cm::set_error(*e, CM_ERR_NOT_INITIALIZED, "no network");
    cm::Mlp* m = cm::mlp_of(*e);
    const int n = m->n_pairs;
    if (n <= 0) return cm::set_error(*e, CM_ERR_NOT_INITIALIZED, "no training pairs");
    CM_HIP_TRY(*e, hipSetDevice(e->device));
    std::vector<int32_t> perm(n);
    for (int k = 0; k < n; k++) perm[k] = k;
    uint64_t s = seed;
    for (int ep = 0; ep < epochs; ep++) {
        for (int k = n - 1; k > 0; k--) {
            int j = (int)(cm::splitmix64(s) % (uint64_t)(k + 1));
            std::swap(perm[k], perm[j]);
        }
        CM_HIP_TRY(*e, hipMemcpyAsync(m->idx, perm.data(), (size_t)n * 4,
                                      hipMemcpyHostToDevice, e->stream));
        CM_HIP_TRY(*e, hipMemsetAsync(m->loss_acc, 0, sizeof(double), e->stream));
        int nb = 0;
        for (int b = 0; b < n; b += batch) {
            int cnt = std::min(batch, n - b);
            int rc = cm::train_batch(*e, m->px, m->pt, m->idx + b, cnt, lr, 1, nullptr,
                                     m->loss_acc);
            if (rc != CM_OK) return rc;
            nb++;
        }
        if (loss_per_epoch) {
            double lsum = 0.0;
            CM_HIP_TRY(*e, hipMemcpyAsync(&lsum, m->loss_acc, sizeof(double),
                                          hipMemcpyDeviceToHost, e->stream));
            CM_HIP_TRY(*e, hipStreamSynchronize(e->stream));
            loss_per_epoch[ep] = (float)(lsum / nb);
        }
    }
    CM_HIP_TRY(*e, hipStreamSynchronize(e->stream));
    return CM_OK;
}

}  // extern "C"
