// CPU FP64 oracle — statistics half (test infrastructure only).
//
// Restates the statX algorithms the HIP library implements:
//   statX/infer/ExpKernel.hpp:74-92               squared-exponential kernel
//   statX/infer/CovarianceForcing.ipp:37-109      C_f = diag(covVec) K diag(covVec), R C_f R^T
//   statX/infer/MismatchCovariance.hpp:76-99      K_s = K_d + eps^2 I
//   statX/infer/LikelihoodKOH.ipp:18-39,74-120    total covariance and log-likelihood (LLT)
//   statX/infer/LikelihoodKOH.ipp:310-361         sensor-space posterior (three FullPivLU inverses)
//   statX/infer/LikelihoodKOH.ipp:441-506         Woodbury full-field posterior
//   statX/infer/GaussianProcess.hpp:117-242       dense GP fit / logL / mean / variance
//   statX/infer/MCMC.hpp:123-158,186-197          random-walk Metropolis
//   del2/apps/reference/laplaceRHSinfer/hyperParamLearning.hpp:60-82,172-190   target density
//   del2/apps/reference/laplaceInverse/hyperParamLearning.hpp:172-190         (draws of its chain)
// The random streams are Philox4x32-10 (Salmon et al., SC'11), not the
// reference's Boost mt19937 stream, which cannot be rebuilt here.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace {

typedef long double LD;

// --------------------------------------------------------------- Philox4x32-10
void philox(const uint32_t* ctr_in, const uint32_t* key_in, uint32_t* out) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// 53-bit uniform in (0,1) from two 32-bit words
double to_unit(uint32_t hi, uint32_t lo) {
    const uint64_t w = ((uint64_t)hi << 32) | lo;
    return ((double)(w >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// Per (chain, step): three N(0,1) proposal increments and one U(0,1) for the
// accept test. Counter {step, chain, purpose, 0}, key {seed lo, seed hi}.
void mcmc_draws(uint64_t seed, int chain, int step, double* z, double* u) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t a[4], b[4], c[4];
    uint32_t ca[4] = {(uint32_t)step, (uint32_t)chain, 0u, 0u};
    uint32_t cb[4] = {(uint32_t)step, (uint32_t)chain, 1u, 0u};
    uint32_t cc[4] = {(uint32_t)step, (uint32_t)chain, 2u, 0u};
    philox(ca, key, a);
    philox(cb, key, b);
    philox(cc, key, c);
    const double two_pi = 6.283185307179586;
    const double u1 = to_unit(a[0], a[1]), u2 = to_unit(a[2], a[3]);
    const double r1 = std::sqrt(-2.0 * std::log(u1));
    z[0] = r1 * std::cos(two_pi * u2);
    z[1] = r1 * std::sin(two_pi * u2);
    const double u3 = to_unit(b[0], b[1]), u4 = to_unit(b[2], b[3]);
    z[2] = std::sqrt(-2.0 * std::log(u3)) * std::cos(two_pi * u4);
    *u = to_unit(c[0], c[1]);
}

// one N(0,1) per (reading, index): counter {reading, index, 3, 0}, Box-Muller (cos)
double reading_normal(uint64_t seed, int reading, int idx) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t ctr[4] = {(uint32_t)reading, (uint32_t)idx, 3u, 0u};
    uint32_t a[4];
    philox(ctr, key, a);
    const double u1 = to_unit(a[0], a[1]), u2 = to_unit(a[2], a[3]);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

double se(int dim, const double* x, const double* y, double sigma, double ell) {
    double sq = 0.0;
    for (int i = 0; i < dim; ++i) {
        const double d = x[i] - y[i];
        sq += d * d;
    }
    return sigma * sigma * std::exp(-sq / (2. * ell * ell));
}

// Cholesky (Eigen::LLT semantics: fails if a pivot is not positive). Lower L in place.
bool llt(int n, std::vector<double>& a) {
    for (int j = 0; j < n; ++j) {
        double s = a[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) s -= a[(size_t)j * n + k] * a[(size_t)j * n + k];
        if (!(s > 0.0)) return false;
        const double d = std::sqrt(s);
        a[(size_t)j * n + j] = d;
        for (int i = j + 1; i < n; ++i) {
            double t = a[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) t -= a[(size_t)i * n + k] * a[(size_t)j * n + k];
            a[(size_t)i * n + j] = t / d;
        }
    }
    return true;
}

void llt_solve(int n, const std::vector<double>& L, double* x) {
    for (int i = 0; i < n; ++i) {
        double s = x[i];
        for (int k = 0; k < i; ++k) s -= L[(size_t)i * n + k] * x[k];
        x[i] = s / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = x[i];
        for (int k = i + 1; k < n; ++k) s -= L[(size_t)k * n + i] * x[k];
        x[i] = s / L[(size_t)i * n + i];
    }
}

// Inverse by LU with full pivoting in long double (Eigen::FullPivLU::inverse).
bool fullpiv_inverse(int n, const double* A, std::vector<LD>& inv) {
    std::vector<LD> a((size_t)n * n);
    for (size_t i = 0; i < (size_t)n * n; ++i) a[i] = A[i];
    std::vector<int> rp(n), cp(n);
    for (int i = 0; i < n; ++i) rp[i] = cp[i] = i;
    for (int k = 0; k < n; ++k) {
        int pi = k, pj = k;
        LD best = -1;
        for (int i = k; i < n; ++i)
            for (int j = k; j < n; ++j) {
                const LD v = std::fabs(a[(size_t)i * n + j]);
                if (v > best) { best = v; pi = i; pj = j; }
            }
        if (!(best > 0)) return false;
        if (pi != k) {
            for (int j = 0; j < n; ++j) std::swap(a[(size_t)k * n + j], a[(size_t)pi * n + j]);
            std::swap(rp[k], rp[pi]);
        }
        if (pj != k) {
            for (int i = 0; i < n; ++i) std::swap(a[(size_t)i * n + k], a[(size_t)i * n + pj]);
            std::swap(cp[k], cp[pj]);
        }
        const LD piv = a[(size_t)k * n + k];
        for (int i = k + 1; i < n; ++i) {
            const LD l = a[(size_t)i * n + k] / piv;
            a[(size_t)i * n + k] = l;
            for (int j = k + 1; j < n; ++j) a[(size_t)i * n + j] -= l * a[(size_t)k * n + j];
        }
    }
    // P A Q = L U  =>  A^-1 = Q U^-1 L^-1 P
    inv.assign((size_t)n * n, 0);
    std::vector<LD> col(n);
    for (int c = 0; c < n; ++c) {
        for (int i = 0; i < n; ++i) col[i] = (rp[i] == c) ? 1 : 0;
        for (int i = 0; i < n; ++i) {
            LD s = col[i];
            for (int k = 0; k < i; ++k) s -= a[(size_t)i * n + k] * col[k];
            col[i] = s;
        }
        for (int i = n - 1; i >= 0; --i) {
            LD s = col[i];
            for (int k = i + 1; k < n; ++k) s -= a[(size_t)i * n + k] * col[k];
            col[i] = s / a[(size_t)i * n + i];
        }
        for (int i = 0; i < n; ++i) inv[(size_t)cp[i] * n + c] = col[i];
    }
    return true;
}

int loglik_impl(int ns, int no, const double* Cu, const double* mean, const double* Y,
                const double* Ks, double rho, double* out) {
    std::vector<double> C((size_t)ns * ns);
    for (size_t i = 0; i < (size_t)ns * ns; ++i) C[i] = Ks[i] + rho * rho * Cu[i];
    if (!llt(ns, C)) return OR_ENOTSPD;
    const double tmp1 = -0.5 * ns * std::log(2. * M_PI);
    double ld = 0.0;
    for (int i = 0; i < ns; ++i) ld += std::log(C[(size_t)i * ns + i]);
    const double tmp2 = -0.5 * (2. * ld);
    double tmp3 = 0.0;
    std::vector<double> d(ns), s(ns);
    for (int k = 0; k < no; ++k) {
        for (int i = 0; i < ns; ++i) d[i] = Y[(size_t)k * ns + i] - rho * mean[i];
        s = d;
        llt_solve(ns, C, s.data());
        double q = 0.0;
        for (int i = 0; i < ns; ++i) q += d[i] * s[i];
        const double c = -0.5 * q;
        if (!(c < 0.)) return OR_EQUADFORM;
        tmp3 += c;
    }
    *out = no * (tmp1 + tmp2) + tmp3;
    return OR_OK;
}

}  // namespace

extern "C" {

void or_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    philox(ctr, key, out);
}

int or_mcmc_normals(uint64_t seed, int chain, int step, double* z, double* u) {
    mcmc_draws(seed, chain, step, z, u);
    return OR_OK;
}

// Draws of the conductivity-uncertainty chain (np proposal increments and the
// accept uniform): counters {step, chain, 32 + i, 0} give the Box-Muller pair
// (cosine, sine) for components 2i, 2i+1; {step, chain, 31, 0} the uniform.
int or_lhs_draws(uint64_t seed, int chain, int step, int np, double* z, double* u) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t a[4];
    const double two_pi = 6.283185307179586;
    for (int i = 0; 2 * i < np; ++i) {
        const uint32_t ctr[4] = {(uint32_t)step, (uint32_t)chain, 32u + (uint32_t)i, 0u};
        philox(ctr, key, a);
        const double u1 = to_unit(a[0], a[1]), u2 = to_unit(a[2], a[3]);
        const double r = std::sqrt(-2.0 * std::log(u1));
        z[2 * i] = r * std::cos(two_pi * u2);
        if (2 * i + 1 < np) z[2 * i + 1] = r * std::sin(two_pi * u2);
    }
    const uint32_t ctr[4] = {(uint32_t)step, (uint32_t)chain, 31u, 0u};
    philox(ctr, key, a);
    *u = to_unit(a[0], a[1]);
    return OR_OK;
}

double or_se_kernel(int dim, const double* x, const double* y, double sigma, double ell) {
    return se(dim, x, y, sigma, ell);
}

// CovarianceForcing::giveProjectedCovar (statX/infer/CovarianceForcing.ipp:37-109):
// C_f(i,j) = covvec_i k(x_i,x_j) covvec_j (entries below cut_eps dropped),
// Cu = R C_f R^T, R is ns x n row-major.
int or_covforcing_project(int n, int dim, const double* coords, const double* covvec,
                          double sigma, double ell, double cut_eps, int ns, const double* R,
                          double* Cu) {
    std::vector<double> T((size_t)ns * n, 0.0), row(n);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
            const double v =
                covvec[i] * se(dim, coords + (size_t)i * dim, coords + (size_t)j * dim, sigma, ell) *
                covvec[j];
            row[j] = (v < cut_eps) ? 0.0 : v;
        }
        for (int s = 0; s < ns; ++s) {
            const double r = R[(size_t)s * n + i];
            if (r == 0.0) continue;
            double* t = T.data() + (size_t)s * n;
            for (int j = 0; j < n; ++j) t[j] += r * row[j];
        }
    }
    for (int s = 0; s < ns; ++s)
        for (int t = 0; t < ns; ++t) {
            double acc = 0.0;
            for (int j = 0; j < n; ++j) acc += T[(size_t)s * n + j] * R[(size_t)t * n + j];
            Cu[(size_t)s * ns + t] = acc;
        }
    return OR_OK;
}

// V = C_f W for W n x k row-major (same C_f as above, cut_eps = 0).
int or_covforcing_apply(int n, int dim, const double* coords, const double* covvec,
                        double sigma, double ell, int k, const double* W, double* V) {
    std::vector<double> row(n);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j)
            row[j] = covvec[i] *
                     se(dim, coords + (size_t)i * dim, coords + (size_t)j * dim, sigma, ell) *
                     covvec[j];
        for (int c = 0; c < k; ++c) {
            double acc = 0.0;
            for (int j = 0; j < n; ++j) acc += row[j] * W[(size_t)j * k + c];
            V[(size_t)i * k + c] = acc;
        }
    }
    return OR_OK;
}

// Rows `rows` (nr indices) of V = C_f W, each an n-term sum of the same C_f
// entries as or_covforcing_apply (CovarianceForcing.ipp:37-109): spot checks
// of the product at sizes where the whole of C_f is out of reach.
int or_covforcing_rows(int n, int dim, const double* coords, const double* covvec, double sigma,
                       double ell, int k, const double* W, int nr, const int* rows, double* V) {
    std::vector<double> acc((size_t)k);
    for (int q = 0; q < nr; ++q) {
        const int i = rows[q];
        if (i < 0 || i >= n) return OR_EARG;
        std::fill(acc.begin(), acc.end(), 0.0);
        for (int j = 0; j < n; ++j) {
            const double cij =
                covvec[i] * se(dim, coords + (size_t)i * dim, coords + (size_t)j * dim, sigma, ell) * covvec[j];
            if (cij == 0.0) continue;
            const double* w = W + (size_t)j * k;
            for (int c = 0; c < k; ++c) acc[c] += cij * w[c];
        }
        for (int c = 0; c < k; ++c) V[(size_t)q * k + c] = acc[c];
    }
    return OR_OK;
}

// MismatchCovariance::giveCovariance (statX/infer/MismatchCovariance.hpp:76-99)
int or_mismatch_cov(int ns, int dim, const double* xs, double sigma, double ell, double eps,
                    double* Ks) {
    for (int k = 0; k < ns; ++k)
        for (int l = 0; l < ns; ++l)
            Ks[(size_t)k * ns + l] = se(dim, xs + (size_t)k * dim, xs + (size_t)l * dim, sigma, ell);
    for (int k = 0; k < ns; ++k) Ks[(size_t)k * ns + k] += eps * eps;
    return OR_OK;
}

// LikelihoodKOH::giveLogLikelihood (statX/infer/LikelihoodKOH.ipp:74-120); Y is
// no x ns (snapshot-major, i.e. the column-major ns x no sensor matrix).
int or_loglik(int ns, int no, const double* Cu, const double* mean, const double* Y,
              const double* Ks, double rho, double* out) {
    return loglik_impl(ns, no, Cu, mean, Y, Ks, rho, out);
}

// LikelihoodKOH::giveProjFEconditionedData (statX/infer/LikelihoodKOH.ipp:310-361)
int or_post_proj(int ns, int no, const double* Cu, const double* mean, const double* Y,
                 const double* Ks, double rho, double* mean_out, double* cov_out) {
    std::vector<LD> KsInv, CuInv, Post;
    if (!fullpiv_inverse(ns, Ks, KsInv)) return OR_ESINGULAR;
    if (!fullpiv_inverse(ns, Cu, CuInv)) return OR_ESINGULAR;
    std::vector<double> sum((size_t)ns * ns);
    for (size_t i = 0; i < (size_t)ns * ns; ++i)
        sum[i] = (double)((LD)(rho * rho * no) * KsInv[i] + CuInv[i]);
    if (!fullpiv_inverse(ns, sum.data(), Post)) return OR_ESINGULAR;
    std::vector<LD> ysum(ns, 0), rhs(ns, 0);
    for (int k = 0; k < no; ++k)
        for (int i = 0; i < ns; ++i) ysum[i] += Y[(size_t)k * ns + i];
    for (int i = 0; i < ns; ++i) {
        LD a = 0, b = 0;
        for (int j = 0; j < ns; ++j) {
            a += KsInv[(size_t)i * ns + j] * ysum[j];
            b += CuInv[(size_t)i * ns + j] * (LD)mean[j];
        }
        rhs[i] = (LD)rho * a + b;
    }
    for (int i = 0; i < ns; ++i) {
        LD m = 0;
        for (int j = 0; j < ns; ++j) m += Post[(size_t)i * ns + j] * rhs[j];
        mean_out[i] = (double)m;
        for (int j = 0; j < ns; ++j) cov_out[(size_t)i * ns + j] = (double)Post[(size_t)i * ns + j];
    }
    return OR_OK;
}

// LikelihoodKOH::giveFEconditionedDataWoodbury (statX/infer/LikelihoodKOH.ipp:441-506).
// CuFull: n x n unprojected C_u; ubar: n; P: ns x n.
int or_post_woodbury(int n, int ns, int no, const double* CuFull, const double* ubar,
                     const double* P, const double* Ks, const double* Y, double rho,
                     double* mean_out, double* cov_out) {
    std::vector<LD> KsInv, Sinv;
    if (!fullpiv_inverse(ns, Ks, KsInv)) return OR_ESINGULAR;
    // CuPt = CuFull P^T (n x ns); covU = P CuFull P^T
    std::vector<LD> CuPt((size_t)n * ns, 0);
    for (int i = 0; i < n; ++i)
        for (int s = 0; s < ns; ++s) {
            LD acc = 0;
            for (int j = 0; j < n; ++j) acc += (LD)CuFull[(size_t)i * n + j] * P[(size_t)s * n + j];
            CuPt[(size_t)i * ns + s] = acc;
        }
    std::vector<double> S((size_t)ns * ns);
    for (int s = 0; s < ns; ++s)
        for (int t = 0; t < ns; ++t) {
            LD acc = 0;
            for (int i = 0; i < n; ++i) acc += (LD)P[(size_t)s * n + i] * CuPt[(size_t)i * ns + t];
            S[(size_t)s * ns + t] = (double)(acc + (LD)Ks[(size_t)s * ns + t] / (LD)(rho * rho * no));
        }
    if (!fullpiv_inverse(ns, S.data(), Sinv)) return OR_ESINGULAR;
    // W = CuPt S^-1 (n x ns); D = I - W P
    std::vector<LD> W((size_t)n * ns, 0);
    for (int i = 0; i < n; ++i)
        for (int t = 0; t < ns; ++t) {
            LD acc = 0;
            for (int s = 0; s < ns; ++s) acc += CuPt[(size_t)i * ns + s] * Sinv[(size_t)s * ns + t];
            W[(size_t)i * ns + t] = acc;
        }
    std::vector<LD> ysum(ns, 0), kv(ns, 0);
    for (int k = 0; k < no; ++k)
        for (int i = 0; i < ns; ++i) ysum[i] += Y[(size_t)k * ns + i];
    for (int s = 0; s < ns; ++s) {
        LD acc = 0;
        for (int t = 0; t < ns; ++t) acc += KsInv[(size_t)s * ns + t] * ysum[t];
        kv[s] = acc;
    }
    // v = rho CuPt kv + ubar ; mean = v - W (P v)
    std::vector<LD> v(n), Pv(ns, 0);
    for (int i = 0; i < n; ++i) {
        LD acc = 0;
        for (int s = 0; s < ns; ++s) acc += CuPt[(size_t)i * ns + s] * kv[s];
        v[i] = (LD)rho * acc + (LD)ubar[i];
    }
    for (int s = 0; s < ns; ++s) {
        LD acc = 0;
        for (int i = 0; i < n; ++i) acc += (LD)P[(size_t)s * n + i] * v[i];
        Pv[s] = acc;
    }
    for (int i = 0; i < n; ++i) {
        LD acc = 0;
        for (int s = 0; s < ns; ++s) acc += W[(size_t)i * ns + s] * Pv[s];
        mean_out[i] = (double)(v[i] - acc);
    }
    if (cov_out) {
        // cov = CuFull - W (P CuFull) = CuFull - W CuPt^T
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                LD acc = 0;
                for (int s = 0; s < ns; ++s) acc += W[(size_t)i * ns + s] * CuPt[(size_t)j * ns + s];
                cov_out[(size_t)i * n + j] = (double)((LD)CuFull[(size_t)i * n + j] - acc);
            }
    }
    return OR_OK;
}

// GaussianProcess (statX/infer/GaussianProcess.hpp:117-242)
int or_gp(int n, int dim, const double* X, const double* y, double sigma, double ell, double eps,
          int np, const double* Xp, double* logl, double* mean, double* var) {
    std::vector<double> K((size_t)n * n);
    for (int i = 0; i < n; ++i) {
        K[(size_t)i * n + i] = se(dim, X + (size_t)i * dim, X + (size_t)i * dim, sigma, ell);
        for (int j = i + 1; j < n; ++j) {
            K[(size_t)i * n + j] = se(dim, X + (size_t)i * dim, X + (size_t)j * dim, sigma, ell);
            K[(size_t)j * n + i] = K[(size_t)i * n + j];
        }
    }
    for (int i = 0; i < n; ++i) K[(size_t)i * n + i] += eps * eps;
    if (!llt(n, K)) return OR_ENOTSPD;
    std::vector<double> alpha(y, y + n);
    llt_solve(n, K, alpha.data());
    if (logl) {
        double t1 = 0.0, t2 = 0.0;
        for (int i = 0; i < n; ++i) t1 += y[i] * alpha[i];
        for (int i = 0; i < n; ++i) t2 += std::log(K[(size_t)i * n + i]);
        *logl = -0.5 * t1 - t2 - 0.5 * n * std::log(2.0 * M_PI);
    }
    std::vector<double> ks(n), w(n);
    for (int p = 0; p < np; ++p) {
        const double* xp = Xp + (size_t)p * dim;
        for (int i = 0; i < n; ++i) ks[i] = se(dim, X + (size_t)i * dim, xp, sigma, ell);
        double m = 0.0;
        for (int i = 0; i < n; ++i) m += ks[i] * alpha[i];
        if (mean) mean[p] = m;
        w = ks;
        llt_solve(n, K, w.data());
        double q = 0.0;
        for (int i = 0; i < n; ++i) q += ks[i] * w[i];
        if (var) var[p] = se(dim, xp, xp, sigma, ell) - q;
    }
    return OR_OK;
}

// GaussianProcess::predictCovariance (statX/infer/GaussianProcess.hpp:160-190):
// cov = k(X*, X*) - K*^T (K + eps^2 I)^-1 K*, K* = k(X, X*), np x np row-major.
int or_gp_cov(int n, int dim, const double* X, double sigma, double ell, double eps, int np,
              const double* Xp, double* cov) {
    std::vector<double> K((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            K[(size_t)i * n + j] = se(dim, X + (size_t)i * dim, X + (size_t)j * dim, sigma, ell) + (i == j ? eps * eps : 0.0);
    if (!llt(n, K)) return OR_ENOTSPD;
    // W = K^-1 K* column by column
    std::vector<double> ks((size_t)n * np), w(n);
    for (int p = 0; p < np; ++p)
        for (int i = 0; i < n; ++i) ks[(size_t)p * n + i] = se(dim, X + (size_t)i * dim, Xp + (size_t)p * dim, sigma, ell);
    for (int q = 0; q < np; ++q) {
        for (int i = 0; i < n; ++i) w[i] = ks[(size_t)q * n + i];
        llt_solve(n, K, w.data());
        for (int p = 0; p < np; ++p) {
            double acc = 0.0;
            for (int i = 0; i < n; ++i) acc += ks[(size_t)p * n + i] * w[i];
            cov[(size_t)p * np + q] = se(dim, Xp + (size_t)p * dim, Xp + (size_t)q * dim, sigma, ell) - acc;
        }
    }
    return OR_OK;
}

// Random-walk Metropolis on theta = log(rho, sigma_d, ell_d) (statX/infer/MCMC.hpp:123-158)
// with the target of hyperParamLearning.hpp:172-190:
//   prior_mode 0: logL(e^theta) + sum(theta)                (flat prior on the hyperparameters)
//   prior_mode 1: ... + log N(e^theta; prior_mu, diag(prior_var))   (enrichWithPrior, lines 60-82)
//   prior_mode 2: logL(e^theta) + log N(theta; prior_mu, diag(prior_var))  (log-normal prior)
// A likelihood that cannot be evaluated (non-SPD covariance) counts as -inf.
// samples: nsteps x 4 = (theta0, theta1, theta2, log target) after each step.
static double koh_target(int ns, int no, int dim, const double* Cu, const double* mean,
                         const double* Y, const double* xs, double eps, const double* th,
                         int prior_mode, const double* pm, const double* pv,
                         std::vector<double>& Ks) {
    const double rho = std::exp(th[0]), sd = std::exp(th[1]), ld = std::exp(th[2]);
    or_mismatch_cov(ns, dim, xs, sd, ld, eps, Ks.data());
    double l;
    if (loglik_impl(ns, no, Cu, mean, Y, Ks.data(), rho, &l) != OR_OK) return -INFINITY;
    if (prior_mode == 0 || prior_mode == 1) l += th[0] + th[1] + th[2];
    if (prior_mode == 1) {
        const double x[3] = {rho, sd, ld};
        double q = 0.0, det = 1.0;
        for (int i = 0; i < 3; ++i) {
            q += (x[i] - pm[i]) * (x[i] - pm[i]) / pv[i];
            det *= pv[i];
        }
        l += std::log(1. / (std::sqrt(std::pow(2. * M_PI, 3) * det)) * std::exp(-0.5 * q));
    } else if (prior_mode == 2) {
        for (int i = 0; i < 3; ++i)
            l += -0.5 * (th[i] - pm[i]) * (th[i] - pm[i]) / pv[i] - 0.5 * std::log(2. * M_PI * pv[i]);
    }
    return l;
}

int or_mcmc_koh(int ns, int no, int dim, const double* Cu, const double* mean, const double* Y,
                const double* xs, double eps, const double* theta0, const double* propsd,
                int nsteps, uint64_t seed, int chain, int prior_mode, const double* prior_mu,
                const double* prior_var, double* samples, double* acc) {
    std::vector<double> Ks((size_t)ns * ns);
    double x[3] = {theta0[0], theta0[1], theta0[2]};
    double y = koh_target(ns, no, dim, Cu, mean, Y, xs, eps, x, prior_mode, prior_mu, prior_var, Ks);
    int accepted = 0;
    for (int t = 0; t < nsteps; ++t) {
        double z[3], u;
        mcmc_draws(seed, chain, t, z, &u);
        double xp[3];
        for (int i = 0; i < 3; ++i) xp[i] = x[i] + propsd[i] * z[i];
        const double yp =
            koh_target(ns, no, dim, Cu, mean, Y, xs, eps, xp, prior_mode, prior_mu, prior_var, Ks);
        const double ratio = std::fmin(0.0, yp - y);
        if (std::log(u) <= ratio) {
            x[0] = xp[0]; x[1] = xp[1]; x[2] = xp[2];
            y = yp;
            ++accepted;
        }
        double* s = samples + (size_t)t * 4;
        s[0] = x[0]; s[1] = x[1]; s[2] = x[2]; s[3] = y;
    }
    *acc = accepted / (double)nsteps;
    return OR_OK;
}

// LikelihoodKOH::giveTypicalReadings + NormalSampler::giveSample
// (statX/infer/LikelihoodKOH.ipp:42-71, statX/infer/NormalSampler.hpp:46-72):
// y_r = rho mean + L z_r with L L^T = K_s + rho^2 Cu; Y is nr x ns.
int or_typical_readings(int ns, int nr, int dim, const double* Cu, const double* mean, const double* xs,
                        double eps, double rho, double sd, double ld, uint64_t seed, double* Y) {
    std::vector<double> C((size_t)ns * ns);
    or_mismatch_cov(ns, dim, xs, sd, ld, eps, C.data());
    for (size_t i = 0; i < (size_t)ns * ns; ++i) C[i] += rho * rho * Cu[i];
    if (!llt(ns, C)) return OR_ENOTSPD;
    std::vector<double> z(ns);
    for (int r = 0; r < nr; ++r) {
        for (int j = 0; j < ns; ++j) z[j] = reading_normal(seed, r, j);
        for (int i = 0; i < ns; ++i) {
            double acc = 0.0;
            for (int j = 0; j <= i; ++j) acc += C[(size_t)i * ns + j] * z[j];
            Y[(size_t)r * ns + i] = rho * mean[i] + acc;
        }
    }
    return OR_OK;
}

// PrecisionMatrixRational::giveTau_ (statX/inferQ/PrecisionMatrixRational.ipp:258-275)
double or_matern_tau(double sigma, double kappa, double nu, int dim) {
    const double t2 = std::tgamma(nu) / ((sigma * sigma) * std::tgamma(nu + dim / 2.) *
                                          std::pow(4. * M_PI, dim / 2.) * std::pow(kappa, 2. * nu));
    return std::sqrt(t2);
}

// givePrecisionMatrixInteger_ applied to X (PrecisionMatrixRational.ipp:177-209,278-298):
// Q X = tau^2 kappa^{4 beta} (A M^-1)^{2 alpha - 1} A X, alpha = max(1, floor(beta)),
// beta = (nu + d/2)/2; A is the (scaled) Matern matrix in CSR, M the lumped mass.
int or_precision_apply(int n, const int* rp, const int* ci, const double* va, const double* mass,
                       double sigma, double kappa, double nu, int dim, int k, const double* X, double* Y) {
    const double beta = (nu + dim / 2.) / 2.;
    const int alpha = (int)std::max(1., std::floor(beta));
    std::vector<double> t((size_t)n * k), u((size_t)n * k);
    or_spmv(n, rp, ci, va, k, X, t.data());
    for (int rep = 0; rep < 2 * alpha - 1; ++rep) {
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < k; ++c) u[(size_t)i * k + c] = t[(size_t)i * k + c] / mass[i];
        or_spmv(n, rp, ci, va, k, u.data(), t.data());
    }
    const double tau = or_matern_tau(sigma, kappa, nu, dim);
    const double sc = tau * tau * std::pow(kappa, 4. * beta);
    for (size_t i = 0; i < (size_t)n * k; ++i) Y[i] = sc * t[i];
    return OR_OK;
}

// statX/apps/reference/mcmcGaussian: RW-MH on log N(x; mu, sd^2); samples nsteps x 2.
int or_mcmc_gauss(double mu, double sd, double x0, double propsd, int nsteps, uint64_t seed,
                  int chain, double* samples, double* acc) {
    auto lp = [&](double x) {
        return std::log(1. / (std::sqrt(2. * M_PI) * sd) * std::exp(-(x - mu) * (x - mu) / (2. * sd * sd)));
    };
    double x = x0, y = lp(x0);
    int accepted = 0;
    for (int t = 0; t < nsteps; ++t) {
        double z[3], u;
        mcmc_draws(seed, chain, t, z, &u);
        const double xp = x + propsd * z[0];
        const double yp = lp(xp);
        if (std::log(u) <= std::fmin(0.0, yp - y)) {
            x = xp; y = yp; ++accepted;
        }
        samples[(size_t)t * 2] = x;
        samples[(size_t)t * 2 + 1] = y;
    }
    *acc = accepted / (double)nsteps;
    return OR_OK;
}

}  // extern "C"
