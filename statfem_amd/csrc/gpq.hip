// GaussianProcessQ (statX/inferQ/GaussianProcessQ.hpp:65-159,
// GaussianProcessQ.ipp:37-186) with the SPDE precision that
// DriverMaxLikelihood assembles (statX/inferQ/DriverMaxLikelihood.hpp:202-312,
// PrecisionMatrixRational.ipp:24-39,177-209,278-298):
//   Q = c (A M^-1)^(2 alpha - 1) A,  c = tau^2 kappa^(4 beta),  alpha = max(1, floor beta),
// A the Matern matrix M + kappa^-2 S (SFH_MATERN with factor 1/kappa^2), M
// its lumped mass, tau from (sigma, kappa, nu).
//
// A fractional beta adds the rational factors of PrecisionMatrixRational.ipp:41-161
// (x^(beta - alpha) ~ a prod (x - c_i) / (b prod (x - d_j))):
//   Q_v = P_l^T Q P_l, P_l = b prod_j (I - d_j M^-1 A), P_r = a prod_i (I - c_i M^-1 A),
// and GaussianProcessQ works with u = P_r v, P_v = P P_r. P_l and P_r are
// polynomials in T = M^-1 A, so they commute and
//   P_v Q_v^-1 P_v^T = (a/b)^2 P r~(T)^2 Q^-1 P^T,  r~(T) = prod_i (I - c_i T) prod_j (I - d_j T)^-1,
// with (I - d_j T)^-1 = (M - d_j A)^-1 M: one more multi-right-hand-side
// solve per pole (M - d_j A is SPD for the negative poles of the best
// approximation) and one sparse product per zero.
//
// The reference factorises sigma_e^2 Q + n_o P^T P and Q with a sparse
// Cholesky per likelihood evaluation. Here everything goes through the
// sensor-space matrix G = P Q^-1 P^T:
//   Q^-1 = (1/c) (A^-1 M)^(2 alpha - 1) A^-1, so with X = r~(T) (A^-1 M)^(alpha - 1) A^-1 P^T
//   (alpha multi-right-hand-side solves with A, one column per sensor; r~ = I
//   for integer beta)
//   G = ((a/b)^2 / c) X^T M X,
// the log-determinant difference and the quadratic form follow from the
// determinant lemma and the push-through identity (k_gpq_loglik), and the
// posterior uses P_r Q_v^-1 P_v^T = ((a/b)^2 / c) r~(T) (A^-1 M)^alpha X. The solves are the
// library's PCG (multigrid on declared grids); no sparse factorisation is
// formed. All noise levels that share (sigma, kappa, nu) share G.
#include <cmath>
#include <map>
#include <tuple>
#include <vector>

#include "common.hpp"

struct sfh_gpq {
    struct Rational {
        double exponent = 0.0;            // beta - alpha
        std::vector<double> top, bot;     // a, c_0 .. c_{m-1};  b, d_0 .. d_{m-1}
    };
    sfh_system* sys = nullptr;
    std::vector<Rational> rats;
    // the operator the system holds: A (d = 0) or M - d A, for kappa
    bool held = false;
    double held_kappa = 0.0, held_d = 0.0;
    int nq1d = 2, ns = 0, no = 0, N = 0;
    double yy = 0.0, ss = 0.0;   // sum of squared readings, |s|^2 (s: the readings' sum)
    sfh::DevBuf PT, mass, Y, X, T, G, eps, out, st, work, w, v;
    sfh::DevBuf srow, zeros, th3, ld;   // inputs of the register-resident factorisation
};

namespace sfh {
namespace {

__global__ void k_gpq_scatter_pt(int ns, int pn, int N, const int* __restrict__ pcols,
                                 const double* __restrict__ pvals, double* __restrict__ PT) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    for (int j = 0; j < pn; ++j) {
        const int c = pcols[(size_t)s * pn + j];
        PT[(size_t)c * ns + s] += pvals[(size_t)s * pn + j];   // column s: one thread
    }
}

__global__ void k_gpq_identity(long long n, double* __restrict__ I) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n * n) I[t] = t / n == t % n ? 1.0 : 0.0;
}

// out = s * in (N x k), or out -= in when s is NaN-free and sub != 0
__global__ void k_gpq_scale(long long total, double s, const double* __restrict__ in, double* __restrict__ out,
                            int sub) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    out[t] = sub ? out[t] - in[t] : s * in[t];
}

__global__ void k_gpq_transpose(int rows, int cols, const double* __restrict__ in, double* __restrict__ out) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)rows * cols) return;
    const int r = (int)(t / cols), c = (int)(t % cols);
    out[(size_t)c * rows + r] = in[t];
}

inline unsigned nb(long long n, int b) { return (unsigned)((n + b - 1) / b); }

// rho = sqrt(n_o g_scale), no mismatch kernel: the register factorisation then
// works on B = sigma_e^2 I + n_o G with the data row s
__global__ void k_gpq_theta(int m, double rho, double* __restrict__ t3) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m) return;
    t3[3 * b] = rho;
    t3[3 * b + 1] = 0.0;
    t3[3 * b + 2] = 1.0;
}

// log p(Y) from log|B| and the factorisation's value ll = -ns log(2 pi)/2 -
// log|B|/2 - |L^-1 s|^2/2: s^T B^-1 s = |L^-1 s|^2 and s^T G B^-1 s =
// (|s|^2 - sigma_e^2 s^T B^-1 s) / n_o. A non-negative quadratic form flag is
// not a failure here (s may vanish).
__global__ void k_gpq_combine(int m, int ns, int no, double yy, double ss, const double* __restrict__ eps,
                              const double* ll, const double* __restrict__ logdet, double* out,
                              int* __restrict__ st) {   // ll and out may be one buffer
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m) return;
    if (st[b] == SFH_ENOTSPD) { out[b] = -INFINITY; return; }
    const double e2 = eps[b] * eps[b];
    const double ld = logdet[b];
    const double qs = -2.0 * (ll[b] + 0.5 * ns * log(2.0 * M_PI) + 0.5 * ld);
    const double tmp1 = -0.5 * (yy - (ss - e2 * qs) / no) / e2;
    const double tmp2 = -0.5 * (ns * (no - 1) * log(e2) + ld);
    out[b] = tmp1 + tmp2 - 0.5 * no * ns * log(2.0 * M_PI);
    st[b] = SFH_OK;
}

struct Prec {
    double scale = 0.0;   // (a/b)^2 / c  (1 / c for integer beta)
    int beta = 0;         // alpha = max(1, floor beta)
    double kappa = 0.0;
    const sfh_gpq::Rational* rat = nullptr;
};

// V -= c T / mass (N x k): one zero factor (I - c M^-1 A) after T = A V
__global__ void k_gpq_zero_factor(long long total, int k, double c, const double* __restrict__ mass,
                                  const double* __restrict__ T, double* __restrict__ V) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    V[t] -= c * T[t] / mass[t / k];
}

void add_stats(sfh_solve_stats& t, const sfh_solve_stats& a) {
    t.inner_iters += a.inner_iters;
    t.outer += a.outer;
    t.max_relres = std::max(t.max_relres, a.max_relres);
    t.ms += a.ms;
    t.converged = t.converged && a.converged;
    t.layout = a.layout;
    t.precond = a.precond;
    t.levels = a.levels;
}

// Puts A = M + kappa^-2 S (d = 0) or M - d A = (1 - d) M + (-d) kappa^-2 S
// (d < 0) into the system, unless it already holds it.
int gpq_operator(sfh_gpq* g, double kappa, double d) {
    if (g->held && g->held_kappa == kappa && g->held_d == d) return SFH_OK;
    g->held = false;
    if (d == 0.0)
        SFH_TRY(assemble_matrix(g->sys, SFH_MATERN, g->nq1d, nullptr, nullptr, kappa, 1.0, 1.0 / (kappa * kappa)));
    else
        SFH_TRY(assemble_matrix(g->sys, SFH_MATERN, g->nq1d, nullptr, nullptr, std::sqrt(1.0 - d),
                                -d / (kappa * kappa), 1.0));
    g->held = true;
    g->held_kappa = kappa;
    g->held_d = d;
    return SFH_OK;
}

// V <- r~(T)^times V for k columns (scratch T: N x k); leading coefficients
// are in Prec::scale
int gpq_rational(sfh_gpq* g, const Prec& pr, int times, int k, double* V, double* T, const sfh_solve_opts* o,
                 sfh_solve_stats& acc) {
    const sfh_gpq::Rational& r = *pr.rat;
    const int m = (int)r.top.size() - 1;
    const long long total = (long long)g->N * k;
    hipStream_t st = g->sys->ctx->stream;
    SFH_TRY(gpq_operator(g, pr.kappa, 0.0));
    for (int t = 0; t < times; ++t)
        for (int i = 0; i < m; ++i) {
            SFH_TRY(spmv(g->sys, k, V, T));
            k_gpq_zero_factor<<<nb(total, 256), 256, 0, st>>>(total, k, r.top[i + 1], g->mass.as<double>(), T, V);
            SFH_LAUNCH_CHECK();
        }
    sfh_solve_stats one{};
    for (int j = 0; j < (int)r.bot.size() - 1; ++j) {
        SFH_TRY(gpq_operator(g, pr.kappa, r.bot[j + 1]));
        for (int t = 0; t < times; ++t) {
            SFH_TRY(sfh_row_scale(g->sys->ctx, g->N, k, g->mass.as<double>(), V, T));
            SFH_TRY(solve(g->sys, 0, k, T, k, V, k, o, &one));
            add_stats(acc, one);
        }
    }
    return SFH_OK;
}

// A(kappa), then X = r~(T) (A^-1 M)^(alpha-1) A^-1 P^T (N x ns) in g->X and
// the raw sensor matrix X^T M X in g->G
int gpq_precision(sfh_gpq* g, double sigma, double kappa, double nu, const sfh_solve_opts* o, sfh_solve_stats& acc,
                  Prec& pr) {
    sfh_system* s = g->sys;
    const int N = g->N, ns = g->ns;
    if (!(kappa > 0.0) || !(sigma > 0.0) || !(nu > 0.0)) { set_error("sigma, kappa and nu must be positive"); return SFH_EARG; }
    const double beta = (nu + s->dim / 2.) / 2.;
    const double alpha = std::max(1., std::floor(beta));
    pr.rat = nullptr;
    if (std::fabs(beta - alpha) >= 1e-12) {
        for (const auto& r : g->rats)
            if (std::fabs(r.exponent - (beta - alpha)) < 1e-12) pr.rat = &r;
        if (!pr.rat) {
            set_error("fractional SPDE exponent %.17g (beta = %g) needs its rational factors", beta - alpha, beta);
            return SFH_EARG;
        }
    }
    double tau = 0.0;
    SFH_TRY(sfh_matern_tau(sigma, kappa, nu, s->dim, &tau));
    pr.beta = (int)alpha;
    pr.kappa = kappa;
    pr.scale = 1.0 / (tau * tau * std::pow(kappa, 4.0 * beta));
    if (pr.rat) {
        const double ab = pr.rat->top[0] / pr.rat->bot[0];
        pr.scale *= ab * ab;
    }
    g->held = false;   // the caller may have re-assembled the system since the last call
    SFH_TRY(gpq_operator(g, kappa, 0.0));
    sfh_solve_stats one{};
    SFH_TRY(solve(s, 0, ns, g->PT.as<double>(), ns, g->X.as<double>(), ns, o, &one));
    add_stats(acc, one);
    for (int r = 1; r < pr.beta; ++r) {
        SFH_TRY(sfh_row_scale(s->ctx, N, ns, g->mass.as<double>(), g->X.as<double>(), g->T.as<double>()));
        SFH_TRY(solve(s, 0, ns, g->T.as<double>(), ns, g->X.as<double>(), ns, o, &one));
        add_stats(acc, one);
    }
    if (pr.rat) SFH_TRY(gpq_rational(g, pr, 1, ns, g->X.as<double>(), g->T.as<double>(), o, acc));
    SFH_TRY(sfh_row_scale(s->ctx, N, ns, g->mass.as<double>(), g->X.as<double>(), g->T.as<double>()));
    SFH_TRY(gram(s->ctx, N, ns, ns, g->X.as<double>(), ns, g->T.as<double>(), ns, g->G.as<double>(), s->scratch3));
    return SFH_OK;
}

// V <- (A^-1 M)^times V for k columns (V, scratch T: N x k)
int gpq_apply_inv_m(sfh_gpq* g, const Prec& pr, int times, int k, double* V, double* T, const sfh_solve_opts* o,
                    sfh_solve_stats& acc) {
    sfh_solve_stats one{};
    SFH_TRY(gpq_operator(g, pr.kappa, 0.0));
    for (int r = 0; r < times; ++r) {
        SFH_TRY(sfh_row_scale(g->sys->ctx, g->N, k, g->mass.as<double>(), V, T));
        SFH_TRY(solve(g->sys, 0, k, T, k, V, k, o, &one));
        add_stats(acc, one);
    }
    return SFH_OK;
}

}  // namespace
}  // namespace sfh

using namespace sfh;

extern "C" {

int sfh_gpq_create(sfh_system* s, int nq1d, int ns, int pn, const int* pcols, const double* pvals, int no,
                   const double* Y, sfh_gpq** out) {
    if (!s || !out || ns < 1 || pn < 1 || !pcols || !pvals || no < 1 || !Y || nq1d < 1 || nq1d > 3) {
        set_error("bad argument");
        return SFH_EARG;
    }
    if (s->kind == 0 || s->dofpn != 1) { set_error("the sparse-precision GP needs a scalar mesh system"); return SFH_EARG; }
    for (int i = 0; i < ns * pn; ++i)
        if (pcols[i] < 0 || pcols[i] >= s->ndof) { set_error("projection column %d out of range", pcols[i]); return SFH_EARG; }
    sfh_gpq* g = new sfh_gpq;
    g->sys = s;
    g->nq1d = nq1d;
    g->ns = ns;
    g->no = no;
    g->N = s->ndof;
    const int N = g->N;
    hipStream_t st = s->ctx->stream;
    DevBuf pc, pv;
    int rc = SFH_OK;
    do {
        const double one[3] = {1.0, 1.0, 1.0};
        if ((rc = g->PT.ensure(sizeof(double) * (size_t)N * ns)) || (rc = g->mass.ensure(sizeof(double) * N)) ||
            (rc = g->Y.ensure(sizeof(double) * (size_t)no * ns)) || (rc = g->X.ensure(sizeof(double) * (size_t)N * ns)) ||
            (rc = g->T.ensure(sizeof(double) * (size_t)N * ns)) || (rc = g->G.ensure(sizeof(double) * (size_t)ns * ns)) ||
            (rc = pc.ensure(sizeof(int) * (size_t)ns * pn)) || (rc = pv.ensure(sizeof(double) * (size_t)ns * pn)))
            break;
        // the lumped mass (DriverMaxLikelihood::computeLumpedMassVector_, DriverMaxLikelihood.hpp:234-262)
        if ((rc = assemble_rhs(s, nq1d, one, nullptr, 1.0, 0, g->mass.as<double>()))) break;
        if (hipMemsetAsync(g->PT.p, 0, sizeof(double) * (size_t)N * ns, st) != hipSuccess ||
            hipMemcpyAsync(pc.p, pcols, sizeof(int) * ns * pn, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(pv.p, pvals, sizeof(double) * ns * pn, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(g->Y.p, Y, sizeof(double) * (size_t)no * ns, hipMemcpyHostToDevice, st) != hipSuccess) {
            set_error("upload failed");
            rc = SFH_EDEVICE;
            break;
        }
        // the readings' sum s and the squared norms the likelihood needs (input data)
        std::vector<double> sv(ns, 0.0), zv(ns, 0.0);
        for (int k = 0; k < no; ++k)
            for (int i = 0; i < ns; ++i) {
                const double y = Y[(size_t)k * ns + i];
                sv[i] += y;
                g->yy += y * y;
            }
        for (int i = 0; i < ns; ++i) g->ss += sv[i] * sv[i];
        if ((rc = g->srow.ensure(sizeof(double) * ns)) || (rc = g->zeros.ensure(sizeof(double) * ns))) break;
        if (hipMemcpyAsync(g->srow.p, sv.data(), sizeof(double) * ns, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(g->zeros.p, zv.data(), sizeof(double) * ns, hipMemcpyHostToDevice, st) != hipSuccess) {
            set_error("upload failed");
            rc = SFH_EDEVICE;
            break;
        }
        k_gpq_scatter_pt<<<nb(ns, 64), 64, 0, st>>>(ns, pn, N, pc.as<int>(), pv.as<double>(), g->PT.as<double>());
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
            set_error("projection set-up failed");
            rc = SFH_EDEVICE;
        }
    } while (0);
    if (rc != SFH_OK) {
        delete g;
        return rc;
    }
    *out = g;
    return SFH_OK;
}

int sfh_gpq_destroy(sfh_gpq* g) {
    delete g;
    return SFH_OK;
}

int sfh_gpq_set_rational(sfh_gpq* g, double exponent, int degree, const double* top, const double* bot) {
    if (!g || !top || !bot || degree < 1 || degree > 32) { set_error("bad argument"); return SFH_EARG; }
    if (!(std::fabs(exponent) > 0.0 && std::fabs(exponent) < 1.0)) {
        set_error("the exponent must lie in (-1, 0) or (0, 1)");
        return SFH_EARG;
    }
    if (!(top[0] != 0.0) || !(bot[0] != 0.0) || !std::isfinite(top[0] / bot[0])) {
        set_error("leading coefficients must be finite and nonzero");
        return SFH_EARG;
    }
    for (int j = 1; j <= degree; ++j) {
        if (!std::isfinite(top[j])) { set_error("zero %d is not finite", j - 1); return SFH_EARG; }
        if (!(bot[j] < 0.0) || !std::isfinite(bot[j])) {
            set_error("pole %d (%g) must be negative: M - d A is then positive definite", j - 1, bot[j]);
            return SFH_EARG;
        }
    }
    sfh_gpq::Rational r;
    r.exponent = exponent;
    r.top.assign(top, top + degree + 1);
    r.bot.assign(bot, bot + degree + 1);
    for (auto& q : g->rats)
        if (std::fabs(q.exponent - exponent) < 1e-12) {
            q = r;
            return SFH_OK;
        }
    g->rats.push_back(r);
    return SFH_OK;
}

int sfh_gpq_loglik(sfh_gpq* g, int M, const double* hps, const sfh_solve_opts* o, double* out, int* status,
                   sfh_solve_stats* stats) {
    if (!g || M < 1 || !hps || !out) { set_error("bad argument"); return SFH_EARG; }
    hipStream_t st = g->sys->ctx->stream;
    for (int b = 0; b < M; ++b)
        if (!(hps[4 * b + 3] > 0.0)) { set_error("set %d: sigma_e must be positive", b); return SFH_EARG; }
    // sets sharing (sigma, kappa, nu) share the precision and G
    std::map<std::tuple<double, double, double>, std::vector<int>> groups;
    for (int b = 0; b < M; ++b) groups[{hps[4 * b], hps[4 * b + 1], hps[4 * b + 2]}].push_back(b);
    SFH_TRY(g->eps.ensure(sizeof(double) * M));
    SFH_TRY(g->out.ensure(sizeof(double) * M));
    SFH_TRY(g->st.ensure(sizeof(int) * M));
    SFH_TRY(g->th3.ensure(sizeof(double) * 3 * M));
    SFH_TRY(g->ld.ensure(sizeof(double) * M));
    std::vector<double> ev(M), ov(M);
    std::vector<int> sv(M);
    sfh_solve_stats acc{};
    acc.converged = 1;
    for (const auto& kv : groups) {
        const std::vector<int>& idx = kv.second;
        Prec pr;
        SFH_TRY(gpq_precision(g, std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first), o, acc, pr));
        const int m = (int)idx.size();
        for (int i = 0; i < m; ++i) ev[i] = hps[4 * idx[i] + 3];
        SFH_CHECK_HIP(hipMemcpyAsync(g->eps.p, ev.data(), sizeof(double) * m, hipMemcpyHostToDevice, st));
        k_gpq_theta<<<nb(m, 64), 64, 0, st>>>(m, std::sqrt(g->no * pr.scale), g->th3.as<double>());
        SFH_LAUNCH_CHECK();
        // one data row (s), zero mean, mismatch sigma 0: B's log-determinant and |L^-1 s|^2
        SFH_TRY(loglik_sets(g->sys->ctx, g->ns, 1, 1, g->zeros.as<double>(), g->G.as<double>(), g->zeros.as<double>(),
                            0, g->srow.as<double>(), 0.0, g->eps.as<double>(), m, g->th3.as<double>(),
                            g->out.as<double>(), g->st.as<int>(), g->ld.as<double>(), g->work));
        k_gpq_combine<<<nb(m, 64), 64, 0, st>>>(m, g->ns, g->no, g->yy, g->ss, g->eps.as<double>(),
                                                g->out.as<double>(), g->ld.as<double>(), g->out.as<double>(),
                                                g->st.as<int>());
        SFH_LAUNCH_CHECK();
        SFH_CHECK_HIP(hipMemcpyAsync(ov.data(), g->out.p, sizeof(double) * m, hipMemcpyDeviceToHost, st));
        SFH_CHECK_HIP(hipMemcpyAsync(sv.data(), g->st.p, sizeof(int) * m, hipMemcpyDeviceToHost, st));
        SFH_CHECK_HIP(hipStreamSynchronize(st));
        for (int i = 0; i < m; ++i) {
            out[idx[i]] = ov[i];
            if (status) status[idx[i]] = sv[i];
        }
    }
    if (stats) *stats = acc;
    return SFH_OK;
}

int sfh_gpq_predict(sfh_gpq* g, const double* hp, const sfh_solve_opts* o, double* mean_dev, double* cov_dev,
                    double* logl, sfh_solve_stats* stats) {
    if (!g || !hp || (!mean_dev && !cov_dev)) { set_error("bad argument"); return SFH_EARG; }
    if (!(hp[3] > 0.0)) { set_error("sigma_e must be positive"); return SFH_EARG; }
    sfh_ctx* c = g->sys->ctx;
    hipStream_t st = c->stream;
    const int N = g->N, ns = g->ns;
    sfh_solve_stats acc{};
    acc.converged = 1;
    Prec pr;
    SFH_TRY(gpq_precision(g, hp[0], hp[1], hp[2], o, acc, pr));
    SFH_TRY(g->eps.ensure(sizeof(double)));
    SFH_TRY(g->out.ensure(sizeof(double)));
    SFH_TRY(g->st.ensure(sizeof(int)));
    SFH_TRY(g->w.ensure(sizeof(double) * ns));
    SFH_CHECK_HIP(hipMemcpyAsync(g->eps.p, hp + 3, sizeof(double), hipMemcpyHostToDevice, st));
    SFH_TRY(gpq_sensor_loglik(c, ns, g->no, g->G.as<double>(), pr.scale, 1, g->eps.as<double>(), g->Y.as<double>(),
                              g->out.as<double>(), g->st.as<int>(), g->w.as<double>(), g->work));
    int s0 = 0;
    double l0 = 0.0;
    SFH_CHECK_HIP(hipMemcpyAsync(&s0, g->st.p, sizeof(int), hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipMemcpyAsync(&l0, g->out.p, sizeof(double), hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    if (s0 != SFH_OK) { set_error("sigma_e^2 I + n_o G is not positive definite"); return s0; }
    if (logl) *logl = l0;
    if (mean_dev) {
        // u = ((a/b)^2 / c) r~(T) (A^-1 M)^alpha X B^-1 s
        SFH_TRY(g->v.ensure(sizeof(double) * 2 * (size_t)N));
        double* v = g->v.as<double>();
        SFH_TRY(dgemm_shared_a(c, N, 1, ns, g->X.as<double>(), ns, g->w.as<double>(), 1, 0, v, 1, 0, 1));
        SFH_TRY(gpq_apply_inv_m(g, pr, pr.beta, 1, v, v + N, o, acc));
        if (pr.rat) SFH_TRY(gpq_rational(g, pr, 1, 1, v, v + N, o, acc));
        k_gpq_scale<<<nb(N, 256), 256, 0, st>>>(N, pr.scale, v, mean_dev, 0);
        SFH_LAUNCH_CHECK();
    }
    if (cov_dev) {
        // P_r sigma_e^2 (sigma_e^2 Q_v + n_o P_v^T P_v)^-1 P_r^T = C - Z H^-1 Z^T with
        // C = P_r Q_v^-1 P_r^T = ((a/b)^2 / c) r~(T)^2 (A^-1 M)^(2 alpha - 1) A^-1,
        // Z = P_r Q_v^-1 P_v^T = ((a/b)^2 / c) r~(T) (A^-1 M)^alpha X, H = (sigma_e^2 / n_o) I + G
        DevBuf Z, Zt, H, big, bigT, hs;
        SFH_TRY(Z.ensure(sizeof(double) * (size_t)N * ns));
        SFH_TRY(Zt.ensure(sizeof(double) * (size_t)N * ns));
        SFH_TRY(H.ensure(sizeof(double) * (size_t)ns * ns));
        SFH_TRY(big.ensure(sizeof(double) * (size_t)N * N));
        SFH_TRY(bigT.ensure(sizeof(double) * (size_t)N * N));
        SFH_TRY(hs.ensure(sizeof(int)));
        SFH_CHECK_HIP(hipMemcpyAsync(Z.p, g->X.p, sizeof(double) * (size_t)N * ns, hipMemcpyDeviceToDevice, st));
        SFH_TRY(gpq_apply_inv_m(g, pr, pr.beta, ns, Z.as<double>(), g->T.as<double>(), o, acc));
        if (pr.rat) SFH_TRY(gpq_rational(g, pr, 1, ns, Z.as<double>(), g->T.as<double>(), o, acc));
        k_gpq_scale<<<nb((long long)N * ns, 256), 256, 0, st>>>((long long)N * ns, pr.scale, Z.as<double>(),
                                                               Z.as<double>(), 0);
        SFH_LAUNCH_CHECK();
        // C on the identity, into cov_dev
        k_gpq_identity<<<nb((long long)N * N, 256), 256, 0, st>>>(N, big.as<double>());
        SFH_LAUNCH_CHECK();
        sfh_solve_stats one{};
        SFH_TRY(gpq_operator(g, pr.kappa, 0.0));
        SFH_TRY(solve(g->sys, 0, N, big.as<double>(), N, cov_dev, N, o, &one));
        add_stats(acc, one);
        SFH_TRY(gpq_apply_inv_m(g, pr, 2 * pr.beta - 1, N, cov_dev, bigT.as<double>(), o, acc));
        if (pr.rat) SFH_TRY(gpq_rational(g, pr, 2, N, cov_dev, bigT.as<double>(), o, acc));
        k_gpq_scale<<<nb((long long)N * N, 256), 256, 0, st>>>((long long)N * N, pr.scale, cov_dev, cov_dev, 0);
        SFH_LAUNCH_CHECK();
        SFH_TRY(gpq_hmat(c, ns, g->no, g->G.as<double>(), pr.scale, hp[3], H.as<double>()));
        k_gpq_transpose<<<nb((long long)N * ns, 256), 256, 0, st>>>(N, ns, Z.as<double>(), Zt.as<double>());
        SFH_LAUNCH_CHECK();
        SFH_TRY(dense_spd_solve(c, ns, H.as<double>(), N, Zt.as<double>(), hs.as<int>()));   // Zt <- H^-1 Z^T
        SFH_TRY(dgemm_shared_a(c, N, N, ns, Z.as<double>(), ns, Zt.as<double>(), N, 0, big.as<double>(), N, 0, 1));
        k_gpq_scale<<<nb((long long)N * N, 256), 256, 0, st>>>((long long)N * N, 0.0, big.as<double>(), cov_dev, 1);
        SFH_LAUNCH_CHECK();
        int hst = 0;
        SFH_CHECK_HIP(hipMemcpyAsync(&hst, hs.p, sizeof(int), hipMemcpyDeviceToHost, st));
        SFH_CHECK_HIP(hipStreamSynchronize(st));
        if (hst != SFH_OK) { set_error("(sigma_e^2 / n_o) I + G is not positive definite"); return hst; }
    }
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    if (stats) *stats = acc;
    return SFH_OK;
}

}  // extern "C"
