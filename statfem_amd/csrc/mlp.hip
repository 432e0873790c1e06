// Learned surrogate of the hyperparameter log-likelihood and delayed-acceptance
// MCMC. No reference counterpart: the reference samples with plain random-walk
// Metropolis (statX/infer/MCMC.hpp:123-158) on the exact likelihood
// (statX/infer/LikelihoodKOH.ipp:74-120); this adds a cheap screen in front of it.
//
// Network: theta (3, log space, standardised) -> 64 tanh -> 64 tanh -> 1
// (standardised log-likelihood). Parameters live as FP32 masters (Adam) and
// BF16 copies; every forward and backward pass reads the BF16 copies and
// accumulates in FP32. Training is full-batch MSE with deterministic
// per-work-group gradient partials reduced in a fixed order.
//
// Sampling (surrogate transition / two-stage delayed acceptance, Christen & Fox
// 2005; Liu 2001): from the current state each chain runs L random-walk steps on
// the surrogate target pi~ = NN + log prior; the end point is accepted with
// min(1, pi(y) pi~(x) / (pi(x) pi~(y))) using the exact likelihood. The
// surrogate sub-chain is reversible for pi~, so the composite kernel leaves the
// exact posterior pi invariant whatever the surrogate's error.
#include <cmath>
#include <cstring>
#include <random>

#include "common.hpp"
#include "mcmc.hpp"

namespace sfh {
namespace {

constexpr int HID = 64;
constexpr int NP = 3 * HID + HID + HID * HID + HID + HID + 1;   // 4481
constexpr int OFF_W1 = 0, OFF_B1 = 3 * HID, OFF_W2 = OFF_B1 + HID, OFF_B2 = OFF_W2 + HID * HID,
              OFF_W3 = OFF_B2 + HID, OFF_B3 = OFF_W3 + HID;
constexpr int SPB = 64;        // training samples per work-group
constexpr int RS = SPB / 16;   // samples per thread of the 16 x 16 gradient tile
static_assert(SPB % 16 == 0, "the gradient tile spreads SPB samples over 16 thread rows");
constexpr int LS = HID + 4;    // LDS row stride: 16-byte rows, successive rows on other banks
constexpr int PSTRIDE = (NP + 1 + 3) / 4 * 4;   // per-work-group partial row (gradient, SSE), 16-byte aligned
// small parameters staged in LDS by the gradient kernel
constexpr int S_B1 = 3 * HID, S_B2 = S_B1 + HID, S_W3 = S_B2 + HID, S_B3 = S_W3 + HID, SMALL = S_B3 + 4;

__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

__device__ inline void load_weights(const uint16_t* __restrict__ bf, float* W) {
    for (int p = threadIdx.x; p < NP; p += blockDim.x) W[p] = bf2f(bf[p]);
}

// forward pass for one standardised input, h1/h2 kept for the backward pass
__device__ inline float mlp_forward(const float* W, const float x[3], float* h1, float* h2) {
    for (int j = 0; j < HID; ++j) {
        float a = W[OFF_B1 + j];
        a += W[OFF_W1 + 3 * j] * x[0] + W[OFF_W1 + 3 * j + 1] * x[1] + W[OFF_W1 + 3 * j + 2] * x[2];
        h1[j] = tanhf(a);
    }
    float y = W[OFF_B3];
    for (int j = 0; j < HID; ++j) {
        float a = W[OFF_B2 + j];
        const float* w2 = W + OFF_W2 + HID * j;
        for (int k = 0; k < HID; ++k) a += w2[k] * h1[k];
        h2[j] = tanhf(a);
        y += W[OFF_W3 + j] * h2[j];
    }
    return y;
}

// Per-work-group gradient of sum_s w (y_s - t_s)^2 and the SSE, written to
// part[blk][PSTRIDE] in parameter order (SSE at index NP). The three 64-wide
// products per sample block (layer-2 forward, layer-2 backward, layer-2 weight
// gradient) run as register-tiled FP32 products over LDS: the 256 threads form
// a 16 x 16 grid, thread (r, c) owning samples r*RS .. r*RS+RS-1 and hidden
// units 4c .. 4c+3 (rows 4r .. 4r+3 of the weight gradient). Weight tiles are
// read as 16-byte vectors that are contiguous across c; activation rows are
// shared by the 16 threads of one r. Every sum runs in a fixed order.
__global__ __launch_bounds__(256) void k_mlp_grad(int n, const float* __restrict__ X, const float* __restrict__ T,
                                                  float w, const uint16_t* __restrict__ bf, float* __restrict__ part) {
    extern __shared__ float4 sh4[];
    float* const W2 = reinterpret_cast<float*>(sh4);   // HID x LS, W2[j][k]
    float* const W2T = W2 + HID * LS;                   // HID x LS, W2T[k][j]
    float* const Ssm = W2T + HID * LS;                  // W1 (3 HID), b1, b2, W3, b3
    float* const H1 = Ssm + SMALL;                      // SPB x LS
    float* const H2 = H1 + SPB * LS;
    float* const D2 = H2 + SPB * LS;                    // dL/da2
    float* const D1 = D2 + SPB * LS;                    // dL/da1
    float* const Sx = D1 + SPB * LS;                    // SPB x 4
    float* const Sdy = Sx + SPB * 4;                    // SPB
    float* const Sse = Sdy + SPB;                       // SPB
    const int t = threadIdx.x, r = t >> 4, c = t & 15;
    const int s0 = blockIdx.x * SPB;
    for (int p = t; p < HID * HID; p += 256) {
        const float v = bf2f(bf[OFF_W2 + p]);
        const int j = p / HID, k = p % HID;
        W2[j * LS + k] = v;
        W2T[k * LS + j] = v;
    }
    for (int p = t; p < 3 * HID; p += 256) Ssm[p] = bf2f(bf[OFF_W1 + p]);
    for (int p = t; p < HID; p += 256) {
        Ssm[S_B1 + p] = bf2f(bf[OFF_B1 + p]);
        Ssm[S_B2 + p] = bf2f(bf[OFF_B2 + p]);
        Ssm[S_W3 + p] = bf2f(bf[OFF_W3 + p]);
    }
    if (t == 0) Ssm[S_B3] = bf2f(bf[OFF_B3]);
    if (t < SPB) {
        const int gs = s0 + t;
        const bool ok = gs < n;
        Sx[4 * t] = ok ? X[3 * gs] : 0.f;
        Sx[4 * t + 1] = ok ? X[3 * gs + 1] : 0.f;
        Sx[4 * t + 2] = ok ? X[3 * gs + 2] : 0.f;
        Sx[4 * t + 3] = 0.f;
    }
    __syncthreads();

    // layer 1: h1[s][j] = tanh(b1[j] + W1[j] . x[s])
    float h1r[RS][4];
#pragma unroll
    for (int a = 0; a < RS; ++a) {
        const int s = r * RS + a;
        const float x0 = Sx[4 * s], x1 = Sx[4 * s + 1], x2 = Sx[4 * s + 2];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int j = 4 * c + b;
            h1r[a][b] = tanhf(Ssm[S_B1 + j] + Ssm[3 * j] * x0 + Ssm[3 * j + 1] * x1 + Ssm[3 * j + 2] * x2);
        }
        *reinterpret_cast<float4*>(H1 + s * LS + 4 * c) = make_float4(h1r[a][0], h1r[a][1], h1r[a][2], h1r[a][3]);
    }
    __syncthreads();

    // layer 2: a2[s][j] = b2[j] + sum_k h1[s][k] W2[j][k]
    float acc[RS][4];
#pragma unroll
    for (int a = 0; a < RS; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = Ssm[S_B2 + 4 * c + b];
    for (int k = 0; k < HID; k += 4) {
        float4 wv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) wv[q] = *reinterpret_cast<const float4*>(W2T + (k + q) * LS + 4 * c);
#pragma unroll
        for (int a = 0; a < RS; ++a) {
            const float4 hv = *reinterpret_cast<const float4*>(H1 + (r * RS + a) * LS + k);
            const float hq[4] = {hv.x, hv.y, hv.z, hv.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                acc[a][0] += hq[q] * wv[q].x;
                acc[a][1] += hq[q] * wv[q].y;
                acc[a][2] += hq[q] * wv[q].z;
                acc[a][3] += hq[q] * wv[q].w;
            }
        }
    }
    // h2, output layer, error, dL/da2
    const float b3 = Ssm[S_B3];
#pragma unroll
    for (int a = 0; a < RS; ++a) {
        const int s = r * RS + a, gs = s0 + s;
        float h2[4], yp = 0.f;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            h2[b] = tanhf(acc[a][b]);
            yp += Ssm[S_W3 + 4 * c + b] * h2[b];
        }
        // sum over the 16 threads of this r (lanes 16m .. 16m+15 of the wave)
#pragma unroll
        for (int m = 8; m > 0; m >>= 1) yp += __shfl_xor(yp, m, 64);
        const bool ok = gs < n;
        const float e = ok ? yp + b3 - T[gs] : 0.f;
        const float dy = 2.0f * w * e;
        *reinterpret_cast<float4*>(H2 + s * LS + 4 * c) = make_float4(h2[0], h2[1], h2[2], h2[3]);
        float d2[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) d2[b] = dy * Ssm[S_W3 + 4 * c + b] * (1.0f - h2[b] * h2[b]);
        *reinterpret_cast<float4*>(D2 + s * LS + 4 * c) = make_float4(d2[0], d2[1], d2[2], d2[3]);
        if (c == 0) {
            Sdy[s] = dy;
            Sse[s] = w * e * e;
        }
    }
    __syncthreads();

    // dL/da1[s][k] = (1 - h1^2) sum_j d2[s][j] W2[j][k]
#pragma unroll
    for (int a = 0; a < RS; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.f;
    for (int j = 0; j < HID; j += 4) {
        float4 wv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) wv[q] = *reinterpret_cast<const float4*>(W2 + (j + q) * LS + 4 * c);
#pragma unroll
        for (int a = 0; a < RS; ++a) {
            const float4 dv = *reinterpret_cast<const float4*>(D2 + (r * RS + a) * LS + j);
            const float dq[4] = {dv.x, dv.y, dv.z, dv.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                acc[a][0] += dq[q] * wv[q].x;
                acc[a][1] += dq[q] * wv[q].y;
                acc[a][2] += dq[q] * wv[q].z;
                acc[a][3] += dq[q] * wv[q].w;
            }
        }
    }
#pragma unroll
    for (int a = 0; a < RS; ++a) {
        float d1[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) d1[b] = acc[a][b] * (1.0f - h1r[a][b] * h1r[a][b]);
        *reinterpret_cast<float4*>(D1 + (r * RS + a) * LS + 4 * c) = make_float4(d1[0], d1[1], d1[2], d1[3]);
    }
    __syncthreads();

    // weight gradient of layer 2: gW2[j][k] = sum_s d2[s][j] h1[s][k], j = 4r + a, k = 4c + b
    float* out = part + (size_t)blockIdx.x * PSTRIDE;
    {
        float g[4][4] = {};
        for (int s = 0; s < SPB; ++s) {
            const float4 dv = *reinterpret_cast<const float4*>(D2 + s * LS + 4 * r);
            const float4 hv = *reinterpret_cast<const float4*>(H1 + s * LS + 4 * c);
            const float dq[4] = {dv.x, dv.y, dv.z, dv.w};
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                g[a][0] += dq[a] * hv.x;
                g[a][1] += dq[a] * hv.y;
                g[a][2] += dq[a] * hv.z;
                g[a][3] += dq[a] * hv.w;
            }
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
            *reinterpret_cast<float4*>(out + OFF_W2 + (4 * r + a) * HID + 4 * c) =
                make_float4(g[a][0], g[a][1], g[a][2], g[a][3]);
    }
    // the small sums: W1 and b1 (threads 0-63), b2 and W3 (64-127), b3 and SSE (128)
    if (t < HID) {
        const int j = t;
        float g0 = 0.f, g1 = 0.f, g2 = 0.f, gb = 0.f;
        for (int s = 0; s < SPB; ++s) {
            const float d = D1[s * LS + j];
            g0 += d * Sx[4 * s];
            g1 += d * Sx[4 * s + 1];
            g2 += d * Sx[4 * s + 2];
            gb += d;
        }
        out[OFF_W1 + 3 * j] = g0;
        out[OFF_W1 + 3 * j + 1] = g1;
        out[OFF_W1 + 3 * j + 2] = g2;
        out[OFF_B1 + j] = gb;
    } else if (t < 2 * HID) {
        const int j = t - HID;
        float gb = 0.f, gw = 0.f;
        for (int s = 0; s < SPB; ++s) {
            gb += D2[s * LS + j];
            gw += Sdy[s] * H2[s * LS + j];
        }
        out[OFF_B2 + j] = gb;
        out[OFF_W3 + j] = gw;
    } else if (t == 2 * HID) {
        float gb = 0.f, e = 0.f;
        for (int s = 0; s < SPB; ++s) {
            gb += Sdy[s];
            e += Sse[s];
        }
        out[OFF_B3] = gb;
        out[NP] = e;
    }
}

// fixed-order sum of the partials; optional Adam step on the FP32 masters and
// refresh of the BF16 copies
__global__ void k_mlp_update(int nblk, const float* __restrict__ part, float* __restrict__ grad, float* master,
                             float* m, float* v, uint16_t* bf, int adam, float lr, float b1, float b2, float eps,
                             float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > NP) return;
    // loads issued eight at a time, added in block order
    float g = 0.f;
    int b = 0;
    for (; b + 8 <= nblk; b += 8) {
        float v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = part[(size_t)(b + q) * PSTRIDE + p];
#pragma unroll
        for (int q = 0; q < 8; ++q) g += v[q];
    }
    for (; b < nblk; ++b) g += part[(size_t)b * PSTRIDE + p];
    grad[p] = g;
    if (!adam || p == NP) return;
    const float mm = b1 * m[p] + (1.f - b1) * g;
    const float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float w = master[p] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    master[p] = w;
    bf[p] = f2bf(w);
}

__global__ void k_bf_from_master(const float* master, uint16_t* bf) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < NP) bf[p] = f2bf(master[p]);
}

// log-likelihood predictions for theta (n x 3, log space)
__global__ __launch_bounds__(256) void k_mlp_infer(int n, const double* __restrict__ th, const double* __restrict__ norm,
                                                   const uint16_t* __restrict__ bf, double* __restrict__ out) {
    __shared__ float W[NP];
    load_weights(bf, W);
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float x[3], h1[HID], h2[HID];
    for (int i = 0; i < 3; ++i) x[i] = (float)((th[3 * s + i] - norm[i]) / norm[3 + i]);
    const float y = mlp_forward(W, x, h1, h2);
    out[s] = (double)y * norm[7] + norm[6];
}

// one wave per chain: lane j owns hidden unit j and holds row j of the second
// layer's weights in registers (w2r); returns the surrogate log target in
// every lane. The layer-2 sums run over k in ascending order.
__device__ inline double surrogate_target(const float* W, const float (&w2r)[HID], float* h1s, const double* th,
                                          const double* norm, int mode, const double* pm, const double* pv) {
    const int j = threadIdx.x;
    float x[3];
    for (int i = 0; i < 3; ++i) x[i] = (float)((th[i] - norm[i]) / norm[3 + i]);
    float a = W[OFF_B1 + j] + W[OFF_W1 + 3 * j] * x[0] + W[OFF_W1 + 3 * j + 1] * x[1] + W[OFF_W1 + 3 * j + 2] * x[2];
    h1s[j] = tanhf(a);
    __syncthreads();
    float a2 = W[OFF_B2 + j];
#pragma unroll
    for (int k = 0; k < HID; ++k) a2 += w2r[k] * h1s[k];   // h1s[k]: one address for the wave (broadcast)
    float c = W[OFF_W3 + j] * tanhf(a2);
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    __syncthreads();
    const double ll = (double)(c + W[OFF_B3]) * norm[7] + norm[6];
    return ll + log_prior(th, mode, pm, pv);
}

// L-step random-walk sub-chain on the surrogate target per chain (L at most
// kMaxSub). The draws of all L steps do not depend on the chain's state: the
// wave forms them first, one step per lane (the same Philox counters and
// transforms as one step after another), and keeps z and log u in LDS, so the
// sequential part is the network and the accept test only.
constexpr int kMaxSub = 256;
__global__ __launch_bounds__(64) void k_da_subchain(int C, int L, int step, uint64_t seed, const double* __restrict__ cur,
                                                    double p0, double p1, double p2, const double* __restrict__ norm,
                                                    const uint16_t* __restrict__ bf, int mode, double m0, double m1,
                                                    double m2, double v0, double v1, double v2, double* __restrict__ prop,
                                                    double* __restrict__ ys0, double* __restrict__ ysL,
                                                    int* __restrict__ nsub) {
    __shared__ float W[NP];
    __shared__ float h1s[HID];
    __shared__ double zs[kMaxSub][3], lus[kMaxSub];
    const int c = blockIdx.x, j = threadIdx.x;
    load_weights(bf, W);
    for (int s = j; s < L; s += blockDim.x) {
        double z[3], u;
        draws3(seed, (uint32_t)step, (uint32_t)c, 0x1000u + 3u * (uint32_t)s, z, &u);
        zs[s][0] = z[0]; zs[s][1] = z[1]; zs[s][2] = z[2];
        lus[s] = log(u);
    }
    // row j of the second layer straight from the BF16 weights (16-byte loads)
    float w2r[HID];
    {
        const uint4* row = reinterpret_cast<const uint4*>(bf + OFF_W2 + HID * j);
#pragma unroll
        for (int q = 0; q < HID / 8; ++q) {
            const uint4 v = row[q];
            const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                w2r[8 * q + 2 * h] = bf2f((uint16_t)(w4[h] & 0xFFFFu));
                w2r[8 * q + 2 * h + 1] = bf2f((uint16_t)(w4[h] >> 16));
            }
        }
    }
    __syncthreads();
    const double pm[3] = {m0, m1, m2}, pv[3] = {v0, v1, v2}, sd[3] = {p0, p1, p2};
    double th[3] = {cur[3 * c], cur[3 * c + 1], cur[3 * c + 2]};
    const double y0 = surrogate_target(W, w2r, h1s, th, norm, mode, pm, pv);
    double y = y0;
    int acc = 0;
    for (int s = 0; s < L; ++s) {
        double tp[3];
        for (int i = 0; i < 3; ++i) tp[i] = th[i] + sd[i] * zs[s][i];
        const double yp = surrogate_target(W, w2r, h1s, tp, norm, mode, pm, pv);
        if (lus[s] <= fmin(0.0, yp - y)) {
            th[0] = tp[0]; th[1] = tp[1]; th[2] = tp[2];
            y = yp;
            ++acc;
        }
    }
    if (threadIdx.x == 0) {
        prop[3 * c] = th[0]; prop[3 * c + 1] = th[1]; prop[3 * c + 2] = th[2];
        ys0[c] = y0;
        ysL[c] = y;
        nsub[c] += acc;
    }
}

__global__ void k_da_nat(int C, const double* __restrict__ th, double* __restrict__ nat) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 3 * C) nat[t] = exp(th[t]);
}

__global__ void k_da_init(int C, const double* __restrict__ cur, const double* __restrict__ ll, const int* __restrict__ st,
                          int mode, double m0, double m1, double m2, double v0, double v1, double v2,
                          double* __restrict__ ycur) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const double pm[3] = {m0, m1, m2}, pv[3] = {v0, v1, v2};
    ycur[c] = st[c] == SFH_OK ? ll[c] + log_prior(cur + 3 * c, mode, pm, pv) : -INFINITY;
}

// exact correction: accept the sub-chain end point y from x with
// min(1, pi(y) pi~(x) / (pi(x) pi~(y)))
__global__ void k_da_accept(int C, int step, int steps, uint64_t seed, double* __restrict__ cur, double* __restrict__ ycur,
                            const double* __restrict__ prop, const double* __restrict__ ll, const int* __restrict__ st,
                            const double* __restrict__ ys0, const double* __restrict__ ysL, int mode, double m0, double m1,
                            double m2, double v0, double v1, double v2, int* __restrict__ nacc,
                            double* __restrict__ samples) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const double pm[3] = {m0, m1, m2}, pv[3] = {v0, v1, v2};
    const double yp = st[c] == SFH_OK ? ll[c] + log_prior(prop + 3 * c, mode, pm, pv) : -INFINITY;
    double z[3], u;
    draws3(seed, (uint32_t)step, (uint32_t)c, 0x7000u, z, &u);
    const double la = (yp - ycur[c]) - (ysL[c] - ys0[c]);
    if (log(u) <= fmin(0.0, la)) {
        for (int i = 0; i < 3; ++i) cur[3 * c + i] = prop[3 * c + i];
        ycur[c] = yp;
        nacc[c] += 1;
    }
    if (samples) {
        double* s = samples + ((size_t)c * steps + step) * 4;
        s[0] = cur[3 * c]; s[1] = cur[3 * c + 1]; s[2] = cur[3 * c + 2]; s[3] = ycur[c];
    }
}

inline unsigned nb(long long n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace
}  // namespace sfh

struct sfh_mlp {
    sfh_ctx* ctx = nullptr;
    sfh::DevBuf master, m, v, bf, part, grad, norm, X, T, th, out;
    int step = 0;
    double norm_h[8] = {0, 0, 0, 1, 1, 1, 0, 1};
};

using namespace sfh;

namespace {

const size_t kGradLds = sizeof(float) * (2 * HID * LS + SMALL + 4 * SPB * LS + 4 * SPB + 2 * SPB);

int upload_training(sfh_mlp* M, int n, const double* th, const double* y) {
    std::vector<float> xf((size_t)n * 3), tf(n);
    for (int s = 0; s < n; ++s) {
        for (int i = 0; i < 3; ++i) xf[3 * s + i] = (float)((th[3 * s + i] - M->norm_h[i]) / M->norm_h[3 + i]);
        tf[s] = (float)((y[s] - M->norm_h[6]) / M->norm_h[7]);
    }
    hipStream_t st = M->ctx->stream;
    SFH_TRY(M->X.ensure(sizeof(float) * 3 * (size_t)n));
    SFH_TRY(M->T.ensure(sizeof(float) * (size_t)n));
    SFH_CHECK_HIP(hipMemcpyAsync(M->X.p, xf.data(), sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice, st));
    SFH_CHECK_HIP(hipMemcpyAsync(M->T.p, tf.data(), sizeof(float) * (size_t)n, hipMemcpyHostToDevice, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    return SFH_OK;
}

int grad_pass(sfh_mlp* M, int n, int adam, float lr) {
    hipStream_t st = M->ctx->stream;
    const int nblk = (n + SPB - 1) / SPB;
    if (kGradLds > M->ctx->lds_per_block) { set_error("MLP gradient needs %zu B of LDS", kGradLds); return SFH_EARG; }
    SFH_TRY(M->part.ensure(sizeof(float) * (size_t)nblk * PSTRIDE));
    k_mlp_grad<<<nblk, 256, kGradLds, st>>>(n, M->X.as<float>(), M->T.as<float>(), 1.0f / n, M->bf.as<uint16_t>(),
                                            M->part.as<float>());
    SFH_LAUNCH_CHECK();
    float c1 = 1.f, c2 = 1.f;
    if (adam) {
        ++M->step;
        c1 = 1.f - std::pow(0.9f, (float)M->step);
        c2 = 1.f - std::pow(0.999f, (float)M->step);
    }
    k_mlp_update<<<nb(NP + 1, 256), 256, 0, st>>>(nblk, M->part.as<float>(), M->grad.as<float>(), M->master.as<float>(),
                                                  M->m.as<float>(), M->v.as<float>(), M->bf.as<uint16_t>(), adam, lr,
                                                  0.9f, 0.999f, 1e-8f, c1, c2);
    SFH_LAUNCH_CHECK();
    return SFH_OK;
}

}  // namespace

extern "C" {

int sfh_mlp_create(sfh_ctx* c, uint64_t seed, sfh_mlp** out) {
    if (!c || !out) { set_error("bad argument"); return SFH_EARG; }
    *out = nullptr;
    sfh_mlp* M = new sfh_mlp;
    M->ctx = c;
    int rc = M->master.ensure(sizeof(float) * (NP + 1));
    if (!rc) rc = M->m.ensure(sizeof(float) * (NP + 1));
    if (!rc) rc = M->v.ensure(sizeof(float) * (NP + 1));
    if (!rc) rc = M->bf.ensure(sizeof(uint16_t) * (NP + 1));
    if (!rc) rc = M->grad.ensure(sizeof(float) * (NP + 1));
    if (!rc) rc = M->norm.ensure(sizeof(double) * 8);
    if (rc) { delete M; return rc; }
    // Glorot-uniform weights, zero biases
    std::vector<float> w(NP + 1, 0.f);
    std::mt19937_64 g(seed);
    auto fill = [&](int off, int fan_in, int fan_out, int cnt) {
        const double a = std::sqrt(6.0 / (fan_in + fan_out));
        std::uniform_real_distribution<double> U(-a, a);
        for (int i = 0; i < cnt; ++i) w[off + i] = (float)U(g);
    };
    fill(OFF_W1, 3, HID, 3 * HID);
    fill(OFF_W2, HID, HID, HID * HID);
    fill(OFF_W3, HID, 1, HID);
    hipStream_t st = c->stream;
    if (hipMemcpyAsync(M->master.p, w.data(), sizeof(float) * (NP + 1), hipMemcpyHostToDevice, st) ||
        hipMemsetAsync(M->m.p, 0, sizeof(float) * (NP + 1), st) || hipMemsetAsync(M->v.p, 0, sizeof(float) * (NP + 1), st) ||
        hipMemcpyAsync(M->norm.p, M->norm_h, sizeof(double) * 8, hipMemcpyHostToDevice, st)) {
        set_error("upload failed");
        delete M;
        return SFH_EDEVICE;
    }
    k_bf_from_master<<<nb(NP, 256), 256, 0, st>>>(M->master.as<float>(), M->bf.as<uint16_t>());
    if (hipStreamSynchronize(st) != hipSuccess) { set_error("init failed"); delete M; return SFH_EDEVICE; }
    *out = M;
    return SFH_OK;
}

int sfh_mlp_destroy(sfh_mlp* M) {
    if (!M) return SFH_OK;
    (void)hipStreamSynchronize(M->ctx->stream);
    delete M;
    return SFH_OK;
}

int sfh_mlp_set_normalization(sfh_mlp* M, const double* norm8) {
    if (!M || !norm8 || !(norm8[3] > 0) || !(norm8[4] > 0) || !(norm8[5] > 0) || !(norm8[7] > 0)) {
        set_error("bad normalisation");
        return SFH_EARG;
    }
    std::memcpy(M->norm_h, norm8, sizeof(double) * 8);
    SFH_CHECK_HIP(hipMemcpyAsync(M->norm.p, M->norm_h, sizeof(double) * 8, hipMemcpyHostToDevice, M->ctx->stream));
    SFH_CHECK_HIP(hipStreamSynchronize(M->ctx->stream));
    return SFH_OK;
}

int sfh_mlp_params(sfh_mlp* M, float* master, uint16_t* bf, double* norm8) {
    if (!M) { set_error("bad argument"); return SFH_EARG; }
    hipStream_t st = M->ctx->stream;
    if (master) SFH_CHECK_HIP(hipMemcpyAsync(master, M->master.p, sizeof(float) * NP, hipMemcpyDeviceToHost, st));
    if (bf) SFH_CHECK_HIP(hipMemcpyAsync(bf, M->bf.p, sizeof(uint16_t) * NP, hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    if (norm8) std::memcpy(norm8, M->norm_h, sizeof(double) * 8);
    return SFH_OK;
}

int sfh_mlp_set_params(sfh_mlp* M, const float* master) {
    if (!M || !master) { set_error("bad argument"); return SFH_EARG; }
    hipStream_t st = M->ctx->stream;
    SFH_CHECK_HIP(hipMemcpyAsync(M->master.p, master, sizeof(float) * NP, hipMemcpyHostToDevice, st));
    k_bf_from_master<<<nb(NP, 256), 256, 0, st>>>(M->master.as<float>(), M->bf.as<uint16_t>());
    SFH_LAUNCH_CHECK();
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    return SFH_OK;
}

int sfh_mlp_grad(sfh_mlp* M, int n, const double* th, const double* y, float* grad, double* loss) {
    if (!M || n <= 0 || !th || !y) { set_error("bad argument"); return SFH_EARG; }
    SFH_TRY(upload_training(M, n, th, y));
    SFH_TRY(grad_pass(M, n, 0, 0.f));
    std::vector<float> g(NP + 1);
    SFH_CHECK_HIP(hipMemcpyAsync(g.data(), M->grad.p, sizeof(float) * (NP + 1), hipMemcpyDeviceToHost, M->ctx->stream));
    SFH_CHECK_HIP(hipStreamSynchronize(M->ctx->stream));
    if (grad) std::memcpy(grad, g.data(), sizeof(float) * NP);
    if (loss) *loss = g[NP];
    return SFH_OK;
}

int sfh_mlp_train(sfh_mlp* M, int n, const double* th, const double* y, int steps, double lr, double* loss) {
    if (!M || n <= 0 || !th || !y || steps <= 0 || !(lr > 0)) { set_error("bad argument"); return SFH_EARG; }
    // standardise inputs and targets from the training set
    double norm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        double mu = 0, s2 = 0;
        for (int s = 0; s < n; ++s) mu += th[3 * s + i];
        mu /= n;
        for (int s = 0; s < n; ++s) s2 += (th[3 * s + i] - mu) * (th[3 * s + i] - mu);
        norm[i] = mu;
        norm[3 + i] = std::sqrt(s2 / n) > 0 ? std::sqrt(s2 / n) : 1.0;
    }
    double mu = 0, s2 = 0;
    for (int s = 0; s < n; ++s) mu += y[s];
    mu /= n;
    for (int s = 0; s < n; ++s) s2 += (y[s] - mu) * (y[s] - mu);
    norm[6] = mu;
    norm[7] = std::sqrt(s2 / n) > 0 ? std::sqrt(s2 / n) : 1.0;
    SFH_TRY(sfh_mlp_set_normalization(M, norm));
    SFH_TRY(upload_training(M, n, th, y));
    for (int t = 0; t < steps; ++t) {
        const float a = (float)(lr * 0.5 * (1.0 + std::cos(M_PI * t / steps)));   // cosine decay
        SFH_TRY(grad_pass(M, n, 1, a));
    }
    float g[1];
    SFH_CHECK_HIP(hipMemcpyAsync(g, M->grad.as<float>() + NP, sizeof(float), hipMemcpyDeviceToHost, M->ctx->stream));
    SFH_CHECK_HIP(hipStreamSynchronize(M->ctx->stream));
    if (loss) *loss = g[0];
    return SFH_OK;
}

int sfh_mlp_predict(sfh_mlp* M, int n, const double* th, double* out) {
    if (!M || n <= 0 || !th || !out) { set_error("bad argument"); return SFH_EARG; }
    hipStream_t st = M->ctx->stream;
    SFH_TRY(M->th.ensure(sizeof(double) * 3 * (size_t)n));
    SFH_TRY(M->out.ensure(sizeof(double) * (size_t)n));
    SFH_CHECK_HIP(hipMemcpyAsync(M->th.p, th, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, st));
    k_mlp_infer<<<nb(n, 256), 256, 0, st>>>(n, M->th.as<double>(), M->norm.as<double>(), M->bf.as<uint16_t>(),
                                            M->out.as<double>());
    SFH_LAUNCH_CHECK();
    SFH_CHECK_HIP(hipMemcpyAsync(out, M->out.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    return SFH_OK;
}

int sfh_mcmc_da(sfh_lik* L, sfh_mlp* M, int C, int steps, int sub, const double* theta0, const double* propsd,
                uint64_t seed, int mode, const double* pm, const double* pv, double* samples, double* acc,
                double* sub_acc) {
    if (!L || !M || C <= 0 || steps <= 0 || sub <= 0 || !theta0 || !propsd) { set_error("bad argument"); return SFH_EARG; }
    if (sub > kMaxSub) { set_error("at most %d surrogate steps per outer step", kMaxSub); return SFH_EARG; }
    if (mode < 0 || mode > 2 || (mode && (!pm || !pv))) { set_error("bad prior"); return SFH_EARG; }
    hipStream_t st = L->ctx->stream;
    SFH_TRY(L->mc_theta.ensure(sizeof(double) * 3 * C));
    SFH_TRY(L->mc_prop.ensure(sizeof(double) * 3 * C));
    SFH_TRY(L->theta.ensure(sizeof(double) * 3 * C));
    SFH_TRY(L->mc_y.ensure(sizeof(double) * C));
    SFH_TRY(L->out.ensure(sizeof(double) * C));
    SFH_TRY(L->status.ensure(sizeof(int) * C));
    SFH_TRY(L->mc_acc.ensure(sizeof(int) * 2 * C));
    SFH_TRY(L->mc_yprop.ensure(sizeof(double) * 2 * C));
    double* dsamp = nullptr;
    if (samples) {
        SFH_TRY(L->mc_samples.ensure(sizeof(double) * 4 * (size_t)C * steps));
        dsamp = L->mc_samples.as<double>();
    }
    const double m0 = pm ? pm[0] : 0, m1 = pm ? pm[1] : 0, m2 = pm ? pm[2] : 0;
    const double v0 = pv ? pv[0] : 1, v1 = pv ? pv[1] : 1, v2 = pv ? pv[2] : 1;
    double* ys0 = L->mc_yprop.as<double>();
    double* ysL = ys0 + C;
    int* nacc = L->mc_acc.as<int>();
    int* nsub = nacc + C;
    SFH_CHECK_HIP(hipMemcpyAsync(L->mc_theta.p, theta0, sizeof(double) * 3 * C, hipMemcpyHostToDevice, st));
    SFH_CHECK_HIP(hipMemsetAsync(L->mc_acc.p, 0, sizeof(int) * 2 * C, st));
    k_da_nat<<<nb(3 * C, 64), 64, 0, st>>>(C, L->mc_theta.as<double>(), L->theta.as<double>());
    SFH_TRY(loglik_batch(L, C, L->theta.as<double>(), L->out.as<double>(), L->status.as<int>()));
    k_da_init<<<nb(C, 64), 64, 0, st>>>(C, L->mc_theta.as<double>(), L->out.as<double>(), L->status.as<int>(), mode, m0,
                                        m1, m2, v0, v1, v2, L->mc_y.as<double>());
    SFH_LAUNCH_CHECK();
    for (int t = 0; t < steps; ++t) {
        k_da_subchain<<<C, 64, 0, st>>>(C, sub, t, seed, L->mc_theta.as<double>(), propsd[0], propsd[1], propsd[2],
                                        M->norm.as<double>(), M->bf.as<uint16_t>(), mode, m0, m1, m2, v0, v1, v2,
                                        L->mc_prop.as<double>(), ys0, ysL, nsub);
        k_da_nat<<<nb(3 * C, 64), 64, 0, st>>>(C, L->mc_prop.as<double>(), L->theta.as<double>());
        SFH_TRY(loglik_batch(L, C, L->theta.as<double>(), L->out.as<double>(), L->status.as<int>()));
        k_da_accept<<<nb(C, 64), 64, 0, st>>>(C, t, steps, seed, L->mc_theta.as<double>(), L->mc_y.as<double>(),
                                              L->mc_prop.as<double>(), L->out.as<double>(), L->status.as<int>(), ys0, ysL,
                                              mode, m0, m1, m2, v0, v1, v2, nacc, dsamp);
        if ((t & 255) == 255) SFH_LAUNCH_CHECK();
    }
    SFH_LAUNCH_CHECK();
    std::vector<int> na(2 * C);
    SFH_CHECK_HIP(hipMemcpyAsync(na.data(), L->mc_acc.p, sizeof(int) * 2 * C, hipMemcpyDeviceToHost, st));
    if (samples)
        SFH_CHECK_HIP(hipMemcpyAsync(samples, dsamp, sizeof(double) * 4 * (size_t)C * steps, hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    for (int c = 0; c < C; ++c) {
        if (acc) acc[c] = na[c] / (double)steps;
        if (sub_acc) sub_acc[c] = na[C + c] / (double)(steps * (double)sub);
    }
    return SFH_OK;
}

}  // extern "C"
