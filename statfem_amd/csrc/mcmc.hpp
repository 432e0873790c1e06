// Random streams and target-density terms shared by the MCMC, reading and
// surrogate kernels. Philox4x32-10
// (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11). Counters are {step or reading, chain or index, purpose, 0}, the key is
// the 64-bit seed.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sfh {

__device__ inline void philox10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                                uint32_t* out) {
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0; c1 = lo1; c2 = hi0 ^ c3 ^ k1; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// 53-bit uniform in (0,1)
__device__ inline double to_unit(uint32_t hi, uint32_t lo) {
    const uint64_t w = ((uint64_t)hi << 32) | lo;
    return ((double)(w >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// three N(0,1) and one U(0,1) for counter words {c0, c1} and purposes p, p+1, p+2
__device__ inline void draws3(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t p, double* z, double* u) {
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    uint32_t a[4], b[4], c[4];
    philox10(c0, c1, p, 0u, k0, k1, a);
    philox10(c0, c1, p + 1u, 0u, k0, k1, b);
    philox10(c0, c1, p + 2u, 0u, k0, k1, c);
    const double two_pi = 6.283185307179586;
    const double u1 = to_unit(a[0], a[1]), u2 = to_unit(a[2], a[3]);
    const double r1 = sqrt(-2.0 * log(u1));
    z[0] = r1 * cos(two_pi * u2);
    z[1] = r1 * sin(two_pi * u2);
    const double u3 = to_unit(b[0], b[1]), u4 = to_unit(b[2], b[3]);
    z[2] = sqrt(-2.0 * log(u3)) * cos(two_pi * u4);
    *u = to_unit(c[0], c[1]);
}

// plain random-walk chain: counter {step, chain, 0..2, 0}
__device__ inline void mcmc_draws(uint64_t seed, int chain, int step, double* z, double* u) {
    draws3(seed, (uint32_t)step, (uint32_t)chain, 0u, z, u);
}

// squared-exponential kernel sigma^2 exp(-|x - y|^2 / (2 ell^2)) of
// ExpKernel::giveValue (statX/infer/ExpKernel.hpp:74-92)
__host__ __device__ inline double se_kernel(int dim, const double* x, const double* y, double sigma, double ell) {
    double sq = 0.0;
    for (int i = 0; i < dim; ++i) {
        const double d = x[i] - y[i];
        sq += d * d;
    }
    return sigma * sigma * exp(-sq / (2. * ell * ell));
}

// prior / Jacobian terms of the log target, hyperParamLearning.hpp:172-190 in
// log space: mode 0 flat in x (+ sum theta), 1 the reference's enrichWithPrior
// (normal density evaluated at e^theta, lines 60-82), 2 log-normal N(theta; mu, var)
__device__ inline double log_prior(const double* th, int mode, const double* pm, const double* pv) {
    double l = 0.0;
    if (mode == 0 || mode == 1) l += th[0] + th[1] + th[2];
    if (mode == 1) {
        double q = 0.0, det = 1.0;
        for (int i = 0; i < 3; ++i) {
            const double x = exp(th[i]);
            q += (x - pm[i]) * (x - pm[i]) / pv[i];
            det *= pv[i];
        }
        l += log(1. / (sqrt(pow(2. * M_PI, 3.0) * det)) * exp(-0.5 * q));
    } else if (mode == 2) {
        for (int i = 0; i < 3; ++i)
            l += -0.5 * (th[i] - pm[i]) * (th[i] - pm[i]) / pv[i] - 0.5 * log(2. * M_PI * pv[i]);
    }
    return l;
}

}  // namespace sfh
