// ForwardPerturbation (statX/infer/ForwardPerturbation.hpp:43-215): first-order
// perturbation density of the FE solution under LHS (material) and RHS
// (forcing) uncertainty, projected onto the sensors,
//   cov = R C_f R^T + sum_e sum_d C_k[e, d] R D_e F D_d^T R^T + (reg sigma_k)^2 P P^T,
//   F = A^-1 C_f A^-T + u u^T,  R = P A^-1,  u = A^-1 f,
// with C_k the CovarianceMaterial kernel matrix of the parameters'
// coordinates (CovarianceMaterial.hpp:85-101) and D_e = dA/dk_e given as a
// dense block on a few DOFs per parameter (the element stiffness derivative
// of setStiffnessDerivatives, ForwardPerturbation.hpp:84-89).
//
// The reference loops over parameter pairs with sparse-dense triple products
// (n_k^2 products of size n_s x N x N). Here the double sum is one structured
// product: with H_e = R[:, dofs_e] D_e (n_s x nb) stacked into H
// (n_s x n_k nb) and Psi = (C_k (x) 1) o F[dofs, dofs] (n_k nb x n_k nb),
// the material part is H Psi H^T: two FP64 MFMA GEMMs after a gather.
// A^-1 comes from one multi-RHS solve with the identity (the reference fills
// the first n_s rows of its aInv, lines 127-141; the two agree when n_s = N).
// Dense N x N intermediates: for moderate N.
#include <cmath>
#include <vector>

#include "common.hpp"
#include "mcmc.hpp"

struct sfh_fpert {
    sfh_system* sys = nullptr;
    int ns = 0, nk = 0, nb = 0, pdim = 0, N = 0, nq1d = 2;
    double reg = 1e-4;
    sfh::DevBuf Pd, PT, dofs, blk, pc;
    // work matrices, kept between updates
    sfh::DevBuf cv, u, I, Ainv, T1, T2, R, HT, Ck, Psi, TT, Cm, RCR, PPt, md, work;
};

namespace sfh {
namespace {

inline unsigned nb_(long long n, int b) { return (unsigned)((n + b - 1) / b); }

__global__ void k_fp_dense_p(int ns, int pn, int N, const int* __restrict__ pcols, const double* __restrict__ pvals,
                             double* __restrict__ Pd) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    for (int j = 0; j < pn; ++j) Pd[(size_t)s * N + pcols[(size_t)s * pn + j]] += pvals[(size_t)s * pn + j];
}

__global__ void k_fp_transpose(int rows, int cols, const double* __restrict__ in, double* __restrict__ out) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)rows * cols) return;
    const int r = (int)(t / cols), c = (int)(t % cols);
    out[(size_t)c * rows + r] = in[t];
}

__global__ void k_fp_identity(long long n, double* __restrict__ I) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n * n) I[t] = t / n == t % n ? 1.0 : 0.0;
}

// covariance weights: the lumped mass, zero on constrained DOFs
__global__ void k_fp_cv(int N, const unsigned char* __restrict__ cmask, double* __restrict__ cv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N && cmask[i]) cv[i] = 0.0;
}

__global__ void k_fp_rank1(int N, const double* __restrict__ u, double* __restrict__ F) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)N * N) return;
    F[t] += u[t / N] * u[t % N];
}

// H^T[(e nb + b), s] = sum_a R[s, dofs_e[a]] D_e[a][b]
__global__ void k_fp_ht(int nk, int nb, int ns, int N, const int* __restrict__ dofs, const double* __restrict__ blk,
                        const double* __restrict__ R, double* __restrict__ HT) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nk * nb * ns) return;
    const int s = (int)(t % ns);
    const long long eb = t / ns;
    const int e = (int)(eb / nb), b = (int)(eb % nb);
    double h = 0.0;
    for (int a = 0; a < nb; ++a)
        h += R[(size_t)s * N + dofs[(size_t)e * nb + a]] * blk[((size_t)e * nb + a) * nb + b];
    HT[t] = h;
}

__global__ void k_fp_ck(int nk, int pdim, const double* __restrict__ pc, double sigma, double ell,
                        double* __restrict__ Ck) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nk * nk) return;
    const int e = (int)(t / nk), d = (int)(t % nk);
    Ck[t] = se_kernel(pdim, pc + (size_t)e * pdim, pc + (size_t)d * pdim, sigma, ell);
}

// Psi[(e nb + a), (d nb + b)] = C_k[e, d] F[dofs_e[a], dofs_d[b]]
__global__ void k_fp_psi(int nk, int nb, int N, const int* __restrict__ dofs, const double* __restrict__ Ck,
                         const double* __restrict__ F, double* __restrict__ Psi) {
    const long long K = (long long)nk * nb;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= K * K) return;
    const long long r = t / K, c = t % K;
    const int e = (int)(r / nb), d = (int)(c / nb);
    Psi[t] = Ck[(size_t)e * nk + d] * F[(size_t)dofs[r] * N + dofs[c]];
}

__global__ void k_fp_combine(int n, const double* __restrict__ a, const double* __restrict__ b, double reg2,
                             const double* __restrict__ ppt, double* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = a[t] + b[t] + reg2 * ppt[t];
}

}  // namespace
}  // namespace sfh

using namespace sfh;

extern "C" {

int sfh_fpert_create(sfh_system* s, int nq1d, int ns, int pn, const int* pcols, const double* pvals, int nk, int nb,
                     const int* dofs, const double* blk, int pdim, const double* pcoords, double reg_factor,
                     sfh_fpert** out) {
    if (!s || !out || ns < 1 || pn < 1 || !pcols || !pvals || nk < 1 || nb < 1 || !dofs || !blk || pdim < 1 ||
        pdim > 3 || !pcoords || nq1d < 1 || nq1d > 3) {
        set_error("bad argument");
        return SFH_EARG;
    }
    if (s->dofpn != 1 || s->kind == 0) { set_error("the perturbation path needs a scalar mesh system"); return SFH_EARG; }
    const int N = s->ndof;
    for (int i = 0; i < ns * pn; ++i)
        if (pcols[i] < 0 || pcols[i] >= N) { set_error("projection column %d out of range", pcols[i]); return SFH_EARG; }
    for (long long i = 0; i < (long long)nk * nb; ++i)
        if (dofs[i] < 0 || dofs[i] >= N) { set_error("derivative DOF %d out of range", dofs[i]); return SFH_EARG; }
    if ((long long)nk * nb > 65536) { set_error("%d x %d derivative DOFs exceed 65536", nk, nb); return SFH_EARG; }
    sfh_fpert* f = new sfh_fpert;
    f->sys = s;
    f->nq1d = nq1d;
    f->ns = ns;
    f->nk = nk;
    f->nb = nb;
    f->pdim = pdim;
    f->N = N;
    f->reg = reg_factor;
    hipStream_t st = s->ctx->stream;
    DevBuf pc, pv;
    int rc = SFH_OK;
    do {
        if ((rc = f->Pd.ensure(sizeof(double) * (size_t)ns * N)) || (rc = f->PT.ensure(sizeof(double) * (size_t)ns * N)) ||
            (rc = f->dofs.ensure(sizeof(int) * (size_t)nk * nb)) ||
            (rc = f->blk.ensure(sizeof(double) * (size_t)nk * nb * nb)) ||
            (rc = f->pc.ensure(sizeof(double) * (size_t)nk * pdim)) || (rc = pc.ensure(sizeof(int) * (size_t)ns * pn)) ||
            (rc = pv.ensure(sizeof(double) * (size_t)ns * pn)))
            break;
        if (hipMemsetAsync(f->Pd.p, 0, sizeof(double) * (size_t)ns * N, st) != hipSuccess ||
            hipMemcpyAsync(pc.p, pcols, sizeof(int) * ns * pn, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(pv.p, pvals, sizeof(double) * ns * pn, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(f->dofs.p, dofs, sizeof(int) * (size_t)nk * nb, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(f->blk.p, blk, sizeof(double) * (size_t)nk * nb * nb, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(f->pc.p, pcoords, sizeof(double) * (size_t)nk * pdim, hipMemcpyHostToDevice, st) != hipSuccess) {
            set_error("upload failed");
            rc = SFH_EDEVICE;
            break;
        }
        k_fp_dense_p<<<nb_(ns, 64), 64, 0, st>>>(ns, pn, N, pc.as<int>(), pv.as<double>(), f->Pd.as<double>());
        k_fp_transpose<<<nb_((long long)ns * N, 256), 256, 0, st>>>(ns, N, f->Pd.as<double>(), f->PT.as<double>());
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
            set_error("projection set-up failed");
            rc = SFH_EDEVICE;
        }
    } while (0);
    if (rc != SFH_OK) {
        delete f;
        return rc;
    }
    *out = f;
    return SFH_OK;
}

int sfh_fpert_destroy(sfh_fpert* f) {
    delete f;
    return SFH_OK;
}

int sfh_fpert_update(sfh_fpert* f, double sigma_f, double ell_f, double sigma_k, double ell_k,
                     const sfh_solve_opts* o, double* mean, double* cov, double* u_dev, sfh_solve_stats* stats) {
    if (!f || !mean || !cov) { set_error("bad argument"); return SFH_EARG; }
    sfh_system* s = f->sys;
    if (!s->assembled) { set_error("assemble and constrain the system first"); return SFH_ESTATE; }
    sfh_ctx* c = s->ctx;
    hipStream_t st = c->stream;
    const int N = f->N, ns = f->ns, nk = f->nk, nb = f->nb;
    const long long K = (long long)nk * nb;
    DevBuf &cv = f->cv, &u = f->u, &I = f->I, &Ainv = f->Ainv, &T1 = f->T1, &T2 = f->T2, &R = f->R, &HT = f->HT,
           &Ck = f->Ck, &Psi = f->Psi, &TT = f->TT, &Cm = f->Cm, &RCR = f->RCR, &PPt = f->PPt, &md = f->md,
           &work = f->work;
    SFH_TRY(cv.ensure(sizeof(double) * N));
    SFH_TRY(u.ensure(sizeof(double) * N));
    SFH_TRY(I.ensure(sizeof(double) * (size_t)N * N));
    SFH_TRY(Ainv.ensure(sizeof(double) * (size_t)N * N));
    SFH_TRY(T1.ensure(sizeof(double) * (size_t)N * N));
    SFH_TRY(T2.ensure(sizeof(double) * (size_t)N * N));
    SFH_TRY(R.ensure(sizeof(double) * (size_t)ns * N));
    SFH_TRY(HT.ensure(sizeof(double) * (size_t)K * ns));
    SFH_TRY(Ck.ensure(sizeof(double) * (size_t)nk * nk));
    SFH_TRY(Psi.ensure(sizeof(double) * (size_t)K * K));
    SFH_TRY(TT.ensure(sizeof(double) * (size_t)K * ns));
    SFH_TRY(Cm.ensure(sizeof(double) * (size_t)ns * ns));
    SFH_TRY(RCR.ensure(sizeof(double) * (size_t)ns * ns));
    SFH_TRY(PPt.ensure(sizeof(double) * (size_t)ns * ns));
    SFH_TRY(md.ensure(sizeof(double) * ns));
    // forcing weights (CovarianceForcing.hpp:101-114)
    const double one[3] = {1.0, 1.0, 1.0};
    SFH_TRY(assemble_rhs(s, f->nq1d, one, nullptr, 1.0, 0, cv.as<double>()));
    k_fp_cv<<<nb_(N, 256), 256, 0, st>>>(N, s->cmask.as<unsigned char>(), cv.as<double>());
    SFH_LAUNCH_CHECK();
    // u = A^-1 f and A^-1
    sfh_solve_stats acc{}, one_st{};
    SFH_TRY(solve(s, 0, 1, s->F.as<double>(), 1, u.as<double>(), 1, o, &one_st));
    acc = one_st;
    k_fp_identity<<<nb_((long long)N * N, 256), 256, 0, st>>>(N, I.as<double>());
    SFH_LAUNCH_CHECK();
    SFH_TRY(solve(s, 0, N, I.as<double>(), N, Ainv.as<double>(), N, o, &one_st));
    acc.inner_iters += one_st.inner_iters;
    acc.ms += one_st.ms;
    acc.max_relres = std::max(acc.max_relres, one_st.max_relres);
    acc.converged = acc.converged && one_st.converged;
    // R = P A^-1
    SFH_TRY(dgemm_shared_a(c, ns, N, N, f->Pd.as<double>(), N, Ainv.as<double>(), N, 0, R.as<double>(), N, 0, 1));
    // F = A^-1 (C_f A^-T) + u u^T   (T1 = A^-T, T2 = C_f A^-T, then I holds F)
    k_fp_transpose<<<nb_((long long)N * N, 256), 256, 0, st>>>(N, N, Ainv.as<double>(), T1.as<double>());
    SFH_LAUNCH_CHECK();
    SFH_TRY(covforcing_apply(s, cv.as<double>(), sigma_f, ell_f, N, T1.as<double>(), T2.as<double>(), 0));
    SFH_TRY(dgemm_shared_a(c, N, N, N, Ainv.as<double>(), N, T2.as<double>(), N, 0, I.as<double>(), N, 0, 1));
    k_fp_rank1<<<nb_((long long)N * N, 256), 256, 0, st>>>(N, u.as<double>(), I.as<double>());
    SFH_LAUNCH_CHECK();
    // material part H Psi H^T
    k_fp_ht<<<nb_(K * ns, 256), 256, 0, st>>>(nk, nb, ns, N, f->dofs.as<int>(), f->blk.as<double>(), R.as<double>(),
                                              HT.as<double>());
    k_fp_ck<<<nb_((long long)nk * nk, 256), 256, 0, st>>>(nk, f->pdim, f->pc.as<double>(), sigma_k, ell_k,
                                                          Ck.as<double>());
    k_fp_psi<<<nb_(K * K, 256), 256, 0, st>>>(nk, nb, N, f->dofs.as<int>(), Ck.as<double>(), I.as<double>(),
                                              Psi.as<double>());
    SFH_LAUNCH_CHECK();
    SFH_TRY(dgemm_shared_a(c, (int)K, ns, (int)K, Psi.as<double>(), (int)K, HT.as<double>(), ns, 0, TT.as<double>(), ns,
                           0, 1));
    SFH_TRY(gram(c, (int)K, ns, ns, TT.as<double>(), ns, HT.as<double>(), ns, Cm.as<double>(), work));
    // forcing part R C_f R^T (T1 = R^T, T2 = C_f R^T)
    k_fp_transpose<<<nb_((long long)ns * N, 256), 256, 0, st>>>(ns, N, R.as<double>(), T1.as<double>());
    SFH_LAUNCH_CHECK();
    SFH_TRY(covforcing_apply(s, cv.as<double>(), sigma_f, ell_f, ns, T1.as<double>(), T2.as<double>(), 0));
    SFH_TRY(gram(c, N, ns, ns, T1.as<double>(), ns, T2.as<double>(), ns, RCR.as<double>(), work));
    SFH_TRY(gram(c, N, ns, ns, f->PT.as<double>(), ns, f->PT.as<double>(), ns, PPt.as<double>(), work));
    const double rs = f->reg * sigma_k;
    k_fp_combine<<<nb_(ns * ns, 256), 256, 0, st>>>(ns * ns, RCR.as<double>(), Cm.as<double>(), rs * rs,
                                                    PPt.as<double>(), Cm.as<double>());
    SFH_LAUNCH_CHECK();
    SFH_TRY(dgemm_shared_a(c, ns, 1, N, f->Pd.as<double>(), N, u.as<double>(), 1, 0, md.as<double>(), 1, 0, 1));
    SFH_CHECK_HIP(hipMemcpyAsync(mean, md.p, sizeof(double) * ns, hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipMemcpyAsync(cov, Cm.p, sizeof(double) * (size_t)ns * ns, hipMemcpyDeviceToHost, st));
    if (u_dev) SFH_CHECK_HIP(hipMemcpyAsync(u_dev, u.p, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    if (stats) *stats = acc;
    return SFH_OK;
}

}  // extern "C"
