// ForwardGaussLHS (statX/infer/ForwardGaussLHS.hpp:41-105) with the solve
// function of laplaceInverse (del2/apps/reference/laplaceInverse/inference.hpp:72-120)
// and the MCMC of its hyperParamLearning
// (del2/apps/reference/laplaceInverse/hyperParamLearning.hpp:110-236), for many
// conductivity fields at once.
//
// The reference re-interpolates the conductivity, re-assembles and re-solves
// one FE system per likelihood evaluation. Here `nch` fields (MCMC chains, or
// any batch of hyperparameter sets) are one system: the disjoint union of nch
// copies of the mesh, copy c owning rows seg[c] .. seg[c+1]-1 and elements
// c nel .. (c+1) nel - 1. One step is a handful of launches for all copies:
//   conductivity   kappa_e = GaussianProcess::predictMean at the centroid of
//                  element e (GaussianProcess.hpp:117-135) given anchor values
//                  w = exp(theta[0 .. na-1]): kappa = Kc w, Kc = k(centroids,
//                  X_a) (K_a + eps_gp^2 I)^-1 formed once;
//   assembly       the Poisson operator with per-element conductivity, the
//                  constraints re-applied (fem.hpp:43-98);
//   solve          one block-diagonal multi-segment PCG, one column per copy;
//   projection     mean_c = P u_c (ForwardGaussLHS.hpp:83-92);
//   likelihood     LikelihoodKOH::giveLogLikelihood (LikelihoodKOH.ipp:74-120)
//                  with C_u = 0 (giveProjectedCovarU, ForwardGaussLHS.hpp:71-79)
//                  and per-copy mean and sensor noise.
// A field whose interpolated conductivity is not positive cannot be assembled
// into an SPD operator: its copy is assembled with unit conductivity and its
// log target is -inf (the proposal is rejected).
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "mcmc.hpp"

struct sfh_lhs {
    sfh_system* sys = nullptr;
    int nch = 0, nel = 0, na = 0, np = 0, dim = 0, nq1d = 2;
    int ns = 0, pn = 0, no = 0, sdim = 0;
    int nbc = 0;
    bool have_data = false, have_fixed = false;
    int prior = 1;
    double rho = 1.0, sd = 0.0, ld = 1.0;
    std::vector<int> seg;
    std::vector<double> pm, pv;
    sfh::DevBuf Kc, pcols, pvals, seg_d, bc_d;
    sfh::DevBuf kappa, bad, U, mean, mean_fixed;
    sfh::DevBuf xs, Cu0, Y, theta3, pm_d, pv_d;
    sfh::DevBuf cur, prop, nat, eps_nat, ll, llst, ycur, nacc, samples, work, tgt;
};

namespace sfh {
namespace {

constexpr int kMaxNP = 16;   // hyperparameters per set (anchor values + sensor noise)

// chain c's N(0,1) proposal increments and U(0,1) acceptance draw at `step`:
// Philox counters {step, chain, 32 + i, 0}, Box-Muller pairs (cosine, sine)
// for components 2i, 2i+1, and {step, chain, 31, 0} for the uniform
__host__ __device__ inline void lhs_draws(uint64_t seed, int chain, int step, int np, double* z, double* u) {
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    uint32_t a[4];
    const double two_pi = 6.283185307179586;
    for (int i = 0; 2 * i < np; ++i) {
        philox10((uint32_t)step, (uint32_t)chain, 32u + (uint32_t)i, 0u, k0, k1, a);
        const double u1 = to_unit(a[0], a[1]), u2 = to_unit(a[2], a[3]);
        const double r = sqrt(-2.0 * log(u1));
        z[2 * i] = r * cos(two_pi * u2);
        if (2 * i + 1 < np) z[2 * i + 1] = r * sin(two_pi * u2);
    }
    philox10((uint32_t)step, (uint32_t)chain, 31u, 0u, k0, k1, a);
    *u = to_unit(a[0], a[1]);
}

// log target of posteriorGivenData (hyperParamLearning.hpp:172-190): the
// log-likelihood, the Jacobian of the sensor noise's log transform
// theta[na], and with `prior` the log of the normal prior density of
// (theta[0 .. na-1], e^theta[na]) with diagonal covariance (lines 60-82,
// 126-150)
__device__ inline double lhs_target(double ll, int st, int bad, const double* th, int na, int prior,
                                    const double* pm, const double* pv) {
    if (st != SFH_OK || bad) return -INFINITY;
    double y = ll + th[na];
    if (prior) {
        const int d = na + 1;
        double q = 0.0, ldet = 0.0;
        for (int i = 0; i < d; ++i) {
            const double x = i < na ? th[i] : exp(th[na]);
            q += (x - pm[i]) * (x - pm[i]) / pv[i];
            ldet += log(pv[i]);
        }
        y += -0.5 * q - 0.5 * (d * log(2.0 * M_PI) + ldet);
    }
    return y;
}

// Kc[e, i] = sum_j k(centroid_e, X_j) Kinv[j, i] for the elements of copy 0
__global__ void k_lhs_kc(int nel, int nen, int dim, int na, const double* __restrict__ X,
                         const int* __restrict__ conn, const double* __restrict__ Xa,
                         const double* __restrict__ Kinv, double sigma, double ell, double* __restrict__ Kc) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nel) return;
    double c[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < nen; ++a) {
        const int v = conn[(size_t)e * nen + a];
        for (int d = 0; d < dim; ++d) c[d] += X[(size_t)v * dim + d];
    }
    for (int d = 0; d < dim; ++d) c[d] /= (double)nen;
    double ks[kMaxNP];
    for (int j = 0; j < na; ++j) ks[j] = se_kernel(dim, c, Xa + (size_t)j * dim, sigma, ell);
    for (int i = 0; i < na; ++i) {
        double s = 0.0;
        for (int j = 0; j < na; ++j) s += ks[j] * Kinv[(size_t)j * na + i];
        Kc[(size_t)e * na + i] = s;
    }
}

// kappa[c nel + e] = sum_i Kc[e, i] w_c[i]; bad[c] when a value is not positive
__global__ void k_lhs_kappa(int nch, int nel, int na, int np, const double* __restrict__ Kc,
                            const double* __restrict__ nat, double* __restrict__ kappa, int* __restrict__ bad) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nch * nel) return;
    const int c = (int)(t / nel), e = (int)(t % nel);
    const double* w = nat + (size_t)c * np;
    double k = 0.0;
    for (int i = 0; i < na; ++i) k += Kc[(size_t)e * na + i] * w[i];
    if (!(k > 0.0)) {   // also NaN
        bad[c] = 1;
        k = 1.0;
    }
    kappa[t] = k;
}

// mean[c ns + s] = sum_j pvals[s pn + j] u[seg[c] + pcols[s pn + j]]
__global__ void k_lhs_project(int nch, int ns, int pn, const int* __restrict__ pcols,
                              const double* __restrict__ pvals, const int* __restrict__ seg,
                              const double* __restrict__ U, double* __restrict__ mean) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nch * ns) return;
    const int c = t / ns, s = t % ns;
    const double* u = U + seg[c];
    double m = 0.0;
    for (int j = 0; j < pn; ++j) m += pvals[(size_t)s * pn + j] * u[pcols[(size_t)s * pn + j]];
    mean[t] = m;
}

// natural values and sensor noise of the sets in th (M x np)
__global__ void k_lhs_natural(int M, int np, int na, const double* __restrict__ th, double* __restrict__ nat,
                              double* __restrict__ eps) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M * np) return;
    const double v = exp(th[t]);
    nat[t] = v;
    if (t % np == na) eps[t / np] = v;
}

__global__ void k_lhs_propose(int C, int np, int na, int step, uint64_t seed, const double* __restrict__ cur,
                              const double* __restrict__ sd, double* __restrict__ prop, double* __restrict__ nat,
                              double* __restrict__ eps, int* __restrict__ bad) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    double z[kMaxNP], u;
    lhs_draws(seed, c, step, np, z, &u);
    for (int i = 0; i < np; ++i) {
        const double v = cur[(size_t)c * np + i] + sd[i] * z[i];
        prop[(size_t)c * np + i] = v;
        nat[(size_t)c * np + i] = exp(v);
    }
    eps[c] = exp(prop[(size_t)c * np + na]);
    bad[c] = 0;
}

__global__ void k_lhs_targets(int M, int np, int na, int prior, const double* __restrict__ th,
                              const double* __restrict__ ll, const int* __restrict__ st,
                              const int* __restrict__ bad, const double* __restrict__ pm,
                              const double* __restrict__ pv, double* __restrict__ y) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= M) return;
    y[c] = lhs_target(ll[c], st[c], bad ? bad[c] : 0, th + (size_t)c * np, na, prior, pm, pv);
}

__global__ void k_lhs_accept(int C, int np, int na, int prior, int step, int steps, uint64_t seed,
                             double* __restrict__ cur, double* __restrict__ ycur, const double* __restrict__ prop,
                             const double* __restrict__ ll, const int* __restrict__ st,
                             const int* __restrict__ bad, const double* __restrict__ pm,
                             const double* __restrict__ pv, int* __restrict__ nacc, double* __restrict__ samples) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    double z[kMaxNP], u;
    lhs_draws(seed, c, step, np, z, &u);
    const double yp = lhs_target(ll[c], st[c], bad[c], prop + (size_t)c * np, na, prior, pm, pv);
    const double ratio = fmin(0.0, yp - ycur[c]);
    if (log(u) <= ratio) {
        for (int i = 0; i < np; ++i) cur[(size_t)c * np + i] = prop[(size_t)c * np + i];
        ycur[c] = yp;
        nacc[c] += 1;
    }
    if (samples) {
        double* s = samples + ((size_t)c * steps + step) * (np + 1);
        for (int i = 0; i < np; ++i) s[i] = cur[(size_t)c * np + i];
        s[np] = ycur[c];
    }
}

__global__ void k_fill_theta3(int M, double rho, double sd, double ld, double* __restrict__ t3) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= M) return;
    t3[3 * c] = rho;
    t3[3 * c + 1] = sd;
    t3[3 * c + 2] = ld;
}

inline unsigned nb(long long n, int b) { return (unsigned)((n + b - 1) / b); }

// dense SPD inverse on the host (set-up, na <= kMaxNP)
bool spd_inverse(int n, std::vector<double> A, std::vector<double>& inv) {
    for (int j = 0; j < n; ++j) {
        double d = A[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
        if (!(d > 0.0)) return false;
        d = std::sqrt(d);
        A[(size_t)j * n + j] = d;
        for (int i = j + 1; i < n; ++i) {
            double s = A[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) s -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
            A[(size_t)i * n + j] = s / d;
        }
    }
    inv.assign((size_t)n * n, 0.0);
    std::vector<double> x(n);
    for (int c = 0; c < n; ++c) {
        for (int i = 0; i < n; ++i) {
            double s = i == c ? 1.0 : 0.0;
            for (int k = 0; k < i; ++k) s -= A[(size_t)i * n + k] * x[k];
            x[i] = s / A[(size_t)i * n + i];
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = x[i];
            for (int k = i + 1; k < n; ++k) s -= A[(size_t)k * n + i] * x[k];
            x[i] = s / A[(size_t)i * n + i];
        }
        for (int i = 0; i < n; ++i) inv[(size_t)i * n + c] = x[i];
    }
    return true;
}

// the solve function for the sets in L->nat (one per copy): conductivity,
// assembly, constraints, the batched solve and the projection into L->mean
int lhs_solve(sfh_lhs* L, const sfh_solve_opts* o, sfh_solve_stats* st) {
    sfh_system* s = L->sys;
    hipStream_t stm = s->ctx->stream;
    k_lhs_kappa<<<nb((long long)L->nch * L->nel, 256), 256, 0, stm>>>(
        L->nch, L->nel, L->na, L->np, L->Kc.as<double>(), L->nat.as<double>(), L->kappa.as<double>(),
        L->bad.as<int>());
    SFH_LAUNCH_CHECK();
    SFH_TRY(assemble_matrix(s, SFH_POISSON, L->nq1d, L->kappa.as<double>(), nullptr, 0.0, 0.0, 1.0));
    SFH_TRY(apply_constraints_dev(s, L->nbc, L->bc_d.as<int>(), 1.0, 1.0));
    s->nc = L->nbc;
    SFH_TRY(solve(s, 0, 1, s->F.as<double>(), 1, L->U.as<double>(), 1, o, st, L->nch, L->seg.data()));
    k_lhs_project<<<nb((long long)L->nch * L->ns, 256), 256, 0, stm>>>(
        L->nch, L->ns, L->pn, L->pcols.as<int>(), L->pvals.as<double>(), L->seg_d.as<int>(), L->U.as<double>(),
        L->mean.as<double>());
    SFH_LAUNCH_CHECK();
    return SFH_OK;
}

// log-likelihood of M sets (natural noise in L->eps_nat) with per-copy means
// (resolve) or the fixed mean
int lhs_loglik(sfh_lhs* L, int M, bool resolve) {
    const double* mean = resolve ? L->mean.as<double>() : L->mean_fixed.as<double>();
    return loglik_sets(L->sys->ctx, L->ns, L->no, L->sdim, L->xs.as<double>(), L->Cu0.as<double>(), mean,
                       resolve ? L->ns : 0, L->Y.as<double>(), 0.0, L->eps_nat.as<double>(), M,
                       L->theta3.as<double>(), L->ll.as<double>(), L->llst.as<int>(), nullptr, L->work);
}

int lhs_sized(sfh_lhs* L, int M) {
    const size_t m = (size_t)std::max(M, L->nch);
    SFH_TRY(L->cur.ensure(sizeof(double) * m * L->np));
    SFH_TRY(L->prop.ensure(sizeof(double) * m * L->np));
    SFH_TRY(L->nat.ensure(sizeof(double) * m * L->np));
    SFH_TRY(L->eps_nat.ensure(sizeof(double) * m));
    SFH_TRY(L->ll.ensure(sizeof(double) * m));
    SFH_TRY(L->llst.ensure(sizeof(int) * m));
    SFH_TRY(L->ycur.ensure(sizeof(double) * m));
    SFH_TRY(L->tgt.ensure(sizeof(double) * m));
    SFH_TRY(L->nacc.ensure(sizeof(int) * m));
    SFH_TRY(L->theta3.ensure(sizeof(double) * 3 * m));
    // (rho, mismatch sigma, mismatch length) of every set, for the likelihood
    k_fill_theta3<<<nb((long long)m, 256), 256, 0, L->sys->ctx->stream>>>((int)m, L->rho, L->sd, L->ld,
                                                                          L->theta3.as<double>());
    SFH_LAUNCH_CHECK();
    return SFH_OK;
}

}  // namespace
}  // namespace sfh

using namespace sfh;

extern "C" {

int sfh_lhs_create(sfh_system* s, int nch, const int* seg_rows, int nel, int na, const double* Xa,
                   double gp_sigma, double gp_ell, double gp_eps, int nq1d, int nbc, const int* bc, int ns,
                   int pn, const int* pcols, const double* pvals, sfh_lhs** out) {
    if (!s || !out || !seg_rows || !Xa || nch < 1 || nel < 1 || na < 1 || na + 1 > kMaxNP || ns < 1 || pn < 1 ||
        !pcols || !pvals || nbc < 0 || (nbc > 0 && !bc)) {
        set_error("bad argument");
        return SFH_EARG;
    }
    if (s->kind == 0 || s->dofpn != 1) { set_error("the conductivity path needs a scalar mesh system"); return SFH_EARG; }
    if ((long long)nch * nel != s->nel) { set_error("%d copies of %d elements do not make the system's %d", nch, nel, s->nel); return SFH_EARG; }
    if (seg_rows[0] != 0 || seg_rows[nch] != s->ndof) { set_error("segments must cover all rows"); return SFH_EARG; }
    const int n0 = seg_rows[1] - seg_rows[0];
    for (int c = 0; c < nch; ++c)
        if (seg_rows[c + 1] - seg_rows[c] != n0) { set_error("copy %d has a different size", c); return SFH_EARG; }
    for (int i = 0; i < ns * pn; ++i)
        if (pcols[i] < 0 || pcols[i] >= n0) { set_error("projection column %d out of range", pcols[i]); return SFH_EARG; }
    for (int i = 0; i < nbc; ++i)
        if (bc[i] < 0 || bc[i] >= s->ndof) { set_error("constraint dof %d out of range", bc[i]); return SFH_EARG; }
    const int dim = s->dim;
    std::vector<double> K((size_t)na * na), Kinv;
    for (int i = 0; i < na; ++i)
        for (int j = 0; j < na; ++j) {
            K[(size_t)i * na + j] = se_kernel(dim, Xa + (size_t)i * dim, Xa + (size_t)j * dim, gp_sigma, gp_ell);
            if (i == j) K[(size_t)i * na + j] += gp_eps * gp_eps;
        }
    if (!spd_inverse(na, K, Kinv)) { set_error("anchor covariance is not positive definite"); return SFH_ENOTSPD; }
    sfh_lhs* L = new sfh_lhs;
    L->sys = s;
    L->nch = nch;
    L->nel = nel;
    L->na = na;
    L->np = na + 1;
    L->dim = dim;
    L->nq1d = nq1d;
    L->ns = ns;
    L->pn = pn;
    L->nbc = nbc;
    L->seg.assign(seg_rows, seg_rows + nch + 1);
    hipStream_t st = s->ctx->stream;
    DevBuf xa, ki;
    int rc = SFH_OK;
#define LTRY(e)                  \
    do {                         \
        rc = (e);                \
        if (rc != SFH_OK) break; \
    } while (0)
    do {
        LTRY(L->Kc.ensure(sizeof(double) * (size_t)nel * na));
        LTRY(xa.ensure(sizeof(double) * (size_t)na * dim));
        LTRY(ki.ensure(sizeof(double) * (size_t)na * na));
        LTRY(L->pcols.ensure(sizeof(int) * (size_t)ns * pn));
        LTRY(L->pvals.ensure(sizeof(double) * (size_t)ns * pn));
        LTRY(L->seg_d.ensure(sizeof(int) * (size_t)(nch + 1)));
        LTRY(L->bc_d.ensure(sizeof(int) * (size_t)std::max(nbc, 1)));
        LTRY(L->kappa.ensure(sizeof(double) * (size_t)nch * nel));
        LTRY(L->bad.ensure(sizeof(int) * (size_t)nch));
        LTRY(L->U.ensure(sizeof(double) * (size_t)s->ndof));
        LTRY(L->mean.ensure(sizeof(double) * (size_t)nch * ns));
        LTRY(L->mean_fixed.ensure(sizeof(double) * (size_t)ns));
        LTRY(L->Cu0.ensure(sizeof(double) * (size_t)ns * ns));
        if (hipMemcpyAsync(xa.p, Xa, sizeof(double) * na * dim, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(ki.p, Kinv.data(), sizeof(double) * na * na, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(L->pcols.p, pcols, sizeof(int) * ns * pn, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(L->pvals.p, pvals, sizeof(double) * ns * pn, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(L->seg_d.p, seg_rows, sizeof(int) * (nch + 1), hipMemcpyHostToDevice, st) != hipSuccess ||
            (nbc > 0 && hipMemcpyAsync(L->bc_d.p, bc, sizeof(int) * nbc, hipMemcpyHostToDevice, st) != hipSuccess) ||
            hipMemsetAsync(L->Cu0.p, 0, sizeof(double) * ns * ns, st) != hipSuccess) {
            set_error("upload failed");
            rc = SFH_EDEVICE;
            break;
        }
        k_lhs_kc<<<nb(nel, 256), 256, 0, st>>>(nel, s->nen, dim, na, s->coords.as<double>(), s->conn.as<int>(),
                                               xa.as<double>(), ki.as<double>(), gp_sigma, gp_ell,
                                               L->Kc.as<double>());
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
            set_error("conductivity weights failed");
            rc = SFH_EDEVICE;
            break;
        }
    } while (0);
#undef LTRY
    if (rc != SFH_OK) {
        delete L;
        return rc;
    }
    *out = L;
    return SFH_OK;
}

int sfh_lhs_destroy(sfh_lhs* L) {
    delete L;
    return SFH_OK;
}

int sfh_lhs_set_data(sfh_lhs* L, int no, const double* Y, int sdim, const double* xs, double rho,
                     double mis_sigma, double mis_ell, int prior, const double* prior_mu, const double* prior_var) {
    if (!L || no < 1 || !Y || sdim < 1 || sdim > 3 || !xs || (prior && (!prior_mu || !prior_var))) {
        set_error("bad argument");
        return SFH_EARG;
    }
    if (L->ns + no > 224 && (size_t)sizeof(double) * (L->ns + (size_t)L->ns * no + 256) > L->sys->ctx->lds_per_block) {
        set_error("%d sensors x %d readings too large", L->ns, no);
        return SFH_EARG;
    }
    hipStream_t st = L->sys->ctx->stream;
    L->no = no;
    L->sdim = sdim;
    L->rho = rho;
    L->sd = mis_sigma;
    L->ld = mis_ell;
    L->prior = prior;
    L->pm.assign(L->np, 0.0);
    L->pv.assign(L->np, 1.0);
    if (prior) {
        for (int i = 0; i < L->np; ++i) {
            if (!(prior_var[i] > 0.0)) { set_error("prior variance %d not positive", i); return SFH_EARG; }
            L->pm[i] = prior_mu[i];
            L->pv[i] = prior_var[i];
        }
    }
    SFH_TRY(L->Y.ensure(sizeof(double) * (size_t)no * L->ns));
    SFH_TRY(L->xs.ensure(sizeof(double) * (size_t)L->ns * sdim));
    SFH_TRY(L->pm_d.ensure(sizeof(double) * L->np));
    SFH_TRY(L->pv_d.ensure(sizeof(double) * L->np));
    SFH_CHECK_HIP(hipMemcpyAsync(L->Y.p, Y, sizeof(double) * (size_t)no * L->ns, hipMemcpyHostToDevice, st));
    SFH_CHECK_HIP(hipMemcpyAsync(L->xs.p, xs, sizeof(double) * (size_t)L->ns * sdim, hipMemcpyHostToDevice, st));
    SFH_CHECK_HIP(hipMemcpyAsync(L->pm_d.p, L->pm.data(), sizeof(double) * L->np, hipMemcpyHostToDevice, st));
    SFH_CHECK_HIP(hipMemcpyAsync(L->pv_d.p, L->pv.data(), sizeof(double) * L->np, hipMemcpyHostToDevice, st));
    SFH_TRY(lhs_sized(L, L->nch));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    L->have_data = true;
    return SFH_OK;
}

int sfh_lhs_update(sfh_lhs* L, const double* theta, const sfh_solve_opts* o, double* mean, sfh_solve_stats* stats) {
    if (!L || !theta) { set_error("bad argument"); return SFH_EARG; }
    hipStream_t st = L->sys->ctx->stream;
    SFH_TRY(lhs_sized(L, L->nch));
    SFH_CHECK_HIP(hipMemcpyAsync(L->cur.p, theta, sizeof(double) * L->nch * L->np, hipMemcpyHostToDevice, st));
    SFH_CHECK_HIP(hipMemsetAsync(L->bad.p, 0, sizeof(int) * L->nch, st));
    k_lhs_natural<<<nb((long long)L->nch * L->np, 256), 256, 0, st>>>(L->nch, L->np, L->na, L->cur.as<double>(),
                                                                      L->nat.as<double>(), L->eps_nat.as<double>());
    SFH_LAUNCH_CHECK();
    SFH_TRY(lhs_solve(L, o, stats));
    std::vector<int> bad(L->nch);
    SFH_CHECK_HIP(hipMemcpyAsync(bad.data(), L->bad.p, sizeof(int) * L->nch, hipMemcpyDeviceToHost, st));
    if (mean)
        SFH_CHECK_HIP(hipMemcpyAsync(mean, L->mean.p, sizeof(double) * L->nch * L->ns, hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    for (int c = 0; c < L->nch; ++c)
        if (bad[c]) { set_error("set %d interpolates a non-positive conductivity", c); return SFH_ENOTSPD; }
    return SFH_OK;
}

int sfh_lhs_fix_mean(sfh_lhs* L, int copy) {
    if (!L || copy < 0 || copy >= L->nch) { set_error("bad argument"); return SFH_EARG; }
    SFH_CHECK_HIP(hipMemcpyAsync(L->mean_fixed.p, L->mean.as<double>() + (size_t)copy * L->ns,
                                 sizeof(double) * L->ns, hipMemcpyDeviceToDevice, L->sys->ctx->stream));
    SFH_CHECK_HIP(hipStreamSynchronize(L->sys->ctx->stream));
    L->have_fixed = true;
    return SFH_OK;
}

int sfh_lhs_log_target(sfh_lhs* L, int M, const double* theta, int resolve, const sfh_solve_opts* o, double* out,
                       int* status, sfh_solve_stats* stats) {
    if (!L || !theta || !out || M < 1) { set_error("bad argument"); return SFH_EARG; }
    if (!L->have_data) { set_error("sfh_lhs_set_data first"); return SFH_ESTATE; }
    if (resolve && M != L->nch) { set_error("a re-solving evaluation takes one set per copy (%d)", L->nch); return SFH_EARG; }
    if (!resolve && !L->have_fixed) { set_error("sfh_lhs_fix_mean first"); return SFH_ESTATE; }
    hipStream_t st = L->sys->ctx->stream;
    SFH_TRY(lhs_sized(L, M));
    SFH_CHECK_HIP(hipMemcpyAsync(L->cur.p, theta, sizeof(double) * M * L->np, hipMemcpyHostToDevice, st));
    SFH_CHECK_HIP(hipMemsetAsync(L->bad.p, 0, sizeof(int) * L->nch, st));
    k_lhs_natural<<<nb((long long)M * L->np, 256), 256, 0, st>>>(M, L->np, L->na, L->cur.as<double>(),
                                                                 L->nat.as<double>(), L->eps_nat.as<double>());
    SFH_LAUNCH_CHECK();
    if (resolve) SFH_TRY(lhs_solve(L, o, stats));
    SFH_TRY(lhs_loglik(L, M, resolve));
    k_lhs_targets<<<nb(M, 64), 64, 0, st>>>(M, L->np, L->na, L->prior, L->cur.as<double>(), L->ll.as<double>(),
                                            L->llst.as<int>(), resolve ? L->bad.as<int>() : nullptr,
                                            L->pm_d.as<double>(), L->pv_d.as<double>(), L->tgt.as<double>());
    SFH_LAUNCH_CHECK();
    SFH_CHECK_HIP(hipMemcpyAsync(out, L->tgt.p, sizeof(double) * M, hipMemcpyDeviceToHost, st));
    if (status) SFH_CHECK_HIP(hipMemcpyAsync(status, L->llst.p, sizeof(int) * M, hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    return SFH_OK;
}

int sfh_lhs_mcmc(sfh_lhs* L, int steps, const double* theta0, const double* propsd, uint64_t seed, int resolve,
                 const sfh_solve_opts* o, double* samples, double* acc, sfh_solve_stats* stats) {
    if (!L || !theta0 || !propsd || steps < 0) { set_error("bad argument"); return SFH_EARG; }
    if (!L->have_data) { set_error("sfh_lhs_set_data first"); return SFH_ESTATE; }
    if (!resolve && !L->have_fixed) { set_error("sfh_lhs_fix_mean first"); return SFH_ESTATE; }
    hipStream_t st = L->sys->ctx->stream;
    const int C = L->nch, np = L->np;
    SFH_TRY(lhs_sized(L, C));
    DevBuf sd;
    SFH_TRY(sd.ensure(sizeof(double) * np));
    double* dsamp = nullptr;
    if (samples) {
        SFH_TRY(L->samples.ensure(sizeof(double) * (size_t)C * steps * (np + 1)));
        dsamp = L->samples.as<double>();
    }
    SFH_CHECK_HIP(hipMemcpyAsync(sd.p, propsd, sizeof(double) * np, hipMemcpyHostToDevice, st));
    SFH_CHECK_HIP(hipMemcpyAsync(L->cur.p, theta0, sizeof(double) * C * np, hipMemcpyHostToDevice, st));
    SFH_CHECK_HIP(hipMemsetAsync(L->nacc.p, 0, sizeof(int) * C, st));
    SFH_CHECK_HIP(hipMemsetAsync(L->bad.p, 0, sizeof(int) * C, st));
    sfh_solve_stats total{}, one{};
    total.converged = 1;
    auto add = [&](const sfh_solve_stats& a) {
        total.inner_iters += a.inner_iters;
        total.outer += a.outer;
        total.max_relres = std::max(total.max_relres, a.max_relres);
        total.ms += a.ms;
        total.converged = total.converged && a.converged;
        total.layout = a.layout;
        total.precond = a.precond;
        total.levels = a.levels;
    };
    // target of the initial sets (posteriorGivenData( initHP ), hyperParamLearning.hpp:221-223)
    k_lhs_natural<<<nb((long long)C * np, 256), 256, 0, st>>>(C, np, L->na, L->cur.as<double>(), L->nat.as<double>(),
                                                              L->eps_nat.as<double>());
    SFH_LAUNCH_CHECK();
    if (resolve) {
        SFH_TRY(lhs_solve(L, o, &one));
        add(one);
    }
    SFH_TRY(lhs_loglik(L, C, resolve));
    k_lhs_targets<<<nb(C, 64), 64, 0, st>>>(C, np, L->na, L->prior, L->cur.as<double>(), L->ll.as<double>(),
                                            L->llst.as<int>(), resolve ? L->bad.as<int>() : nullptr,
                                            L->pm_d.as<double>(), L->pv_d.as<double>(), L->ycur.as<double>());
    SFH_LAUNCH_CHECK();
    for (int t = 0; t < steps; ++t) {
        k_lhs_propose<<<nb(C, 64), 64, 0, st>>>(C, np, L->na, t, seed, L->cur.as<double>(), sd.as<double>(),
                                                L->prop.as<double>(), L->nat.as<double>(), L->eps_nat.as<double>(),
                                                L->bad.as<int>());
        SFH_LAUNCH_CHECK();
        if (resolve) {
            SFH_TRY(lhs_solve(L, o, &one));
            add(one);
        }
        SFH_TRY(lhs_loglik(L, C, resolve));
        k_lhs_accept<<<nb(C, 64), 64, 0, st>>>(C, np, L->na, L->prior, t, steps, seed, L->cur.as<double>(),
                                               L->ycur.as<double>(), L->prop.as<double>(), L->ll.as<double>(),
                                               L->llst.as<int>(), L->bad.as<int>(), L->pm_d.as<double>(),
                                               L->pv_d.as<double>(), L->nacc.as<int>(), dsamp);
        if ((t & 255) == 255) SFH_LAUNCH_CHECK();
    }
    SFH_LAUNCH_CHECK();
    std::vector<int> na(C);
    SFH_CHECK_HIP(hipMemcpyAsync(na.data(), L->nacc.p, sizeof(int) * C, hipMemcpyDeviceToHost, st));
    if (samples)
        SFH_CHECK_HIP(hipMemcpyAsync(samples, dsamp, sizeof(double) * (size_t)C * steps * (np + 1),
                                     hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    if (acc)
        for (int c = 0; c < C; ++c) acc[c] = steps > 0 ? na[c] / (double)steps : 0.0;
    if (stats) *stats = total;
    return SFH_OK;
}

}  // extern "C"
