// CPU FP64 oracle for the statFEM hot path — test infrastructure only.
// Loaded by the test suite, the smoke check and the CPU-baseline leg of the
// benchmark; never by the product path.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_OK = 0, OR_EARG = -1, OR_ENEGJAC = -2, OR_ESINGULAR = -3, OR_ENOTSPD = -4,
       OR_EQUADFORM = -5 };
enum { OR_LINE2 = 1, OR_TRI3 = 2, OR_QUAD4 = 3, OR_HEX8 = 4 };
enum { OR_POISSON = 0, OR_ELASTICITY = 1, OR_MATERN = 2 };

int or_quad_rule(int kind, int nq1d, double* w, double* xi);
int or_element_matrices(int physics, int kind, int nq1d, int nnodes, const double* coords,
                        int nel, const int* conn, const double* p1, const double* p2,
                        double factor, double* Ke);
int or_body_force(int kind, int nq1d, int nnodes, const double* coords, int nel,
                  const int* conn, int dofpn, const double* fconst, const double* fnodal,
                  double factor, double* F);
void* or_triplet_new(int n);
void or_triplet_free(void* h);
int or_triplet_insert_elements(void* h, int nel, int ned, const int* edofs, const double* Ke);
int or_triplet_apply_constraints(void* h, double* F, int nc, const int* dofs,
                                 const double* values, double factor, double scalar);
long long or_triplet_nnz(void* h);
int or_triplet_csr(void* h, int* rowptr, int* colind, double* vals);
void* or_lu_factor(int n, const int* rowptr, const int* colind, const double* vals, int* status);
void or_lu_free(void* h);
int or_lu_solve(void* h, int trans, int k, const double* B, double* X);
void* or_chol_factor(int n, const int* rowptr, const int* colind, const double* vals,
                     const unsigned char* constrained, int* status);
void or_chol_free(void* h);
int or_chol_solve(void* h, int k, const double* B, double* X);
int or_chol_solve_mt(void* h, int k, const double* B, double* X, int nthreads);
int or_spmv(int n, const int* rowptr, const int* colind, const double* vals, int k,
            const double* X, double* Y);

// statistics half (oracle_stat.cpp)
void or_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out);
int or_mcmc_normals(uint64_t seed, int chain, int step, double* z /*3*/, double* u /*1*/);
int or_lhs_draws(uint64_t seed, int chain, int step, int np, double* z /*np*/, double* u /*1*/);
double or_se_kernel(int dim, const double* x, const double* y, double sigma, double ell);
int or_covforcing_project(int n, int dim, const double* coords, const double* covvec,
                          double sigma, double ell, double cut_eps, int ns, const double* R,
                          double* Cu);
int or_covforcing_apply(int n, int dim, const double* coords, const double* covvec,
                        double sigma, double ell, int k, const double* W, double* V);
int or_covforcing_rows(int n, int dim, const double* coords, const double* covvec, double sigma,
                       double ell, int k, const double* W, int nr, const int* rows, double* V);
int or_mismatch_cov(int ns, int dim, const double* xs, double sigma, double ell, double eps,
                    double* Ks);
int or_loglik(int ns, int no, const double* Cu, const double* mean, const double* Y,
              const double* Ks, double rho, double* out);
int or_post_proj(int ns, int no, const double* Cu, const double* mean, const double* Y,
                 const double* Ks, double rho, double* mean_out, double* cov_out);
int or_post_woodbury(int n, int ns, int no, const double* CuFull, const double* ubar,
                     const double* P, const double* Ks, const double* Y, double rho,
                     double* mean_out, double* cov_out);
int or_gp(int n, int dim, const double* X, const double* y, double sigma, double ell, double eps,
          int np, const double* Xp, double* logl, double* mean, double* var);
int or_gp_cov(int n, int dim, const double* X, double sigma, double ell, double eps, int np,
              const double* Xp, double* cov);
int or_mcmc_koh(int ns, int no, int dim, const double* Cu, const double* mean, const double* Y,
                const double* xs, double eps, const double* theta0, const double* propsd,
                int nsteps, uint64_t seed, int chain, int prior_mode, const double* prior_mu,
                const double* prior_var, double* samples, double* acc);
int or_typical_readings(int ns, int nr, int dim, const double* Cu, const double* mean, const double* xs,
                        double eps, double rho, double sd, double ld, uint64_t seed, double* Y);
double or_matern_tau(double sigma, double kappa, double nu, int dim);
int or_precision_apply(int n, const int* rp, const int* ci, const double* va, const double* mass,
                       double sigma, double kappa, double nu, int dim, int k, const double* X, double* Y);
int or_mcmc_gauss(double mu, double sd, double x0, double propsd, int nsteps, uint64_t seed,
                  int chain, double* samples, double* acc);

#ifdef __cplusplus
}
#endif
