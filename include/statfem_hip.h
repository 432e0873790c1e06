/* statfem_hip.h — C ABI of the MI355X statistical-FEM hot path.
 *
 * Plain C types only (pointers, sizes, doubles); no torch or C++ types cross
 * this boundary. Every entry point returns an int status (SFH_OK = 0) and never
 * aborts or throws; sfh_last_error() gives the message of the last failure on
 * the calling thread. Arrays named *_dev are device pointers on the context's
 * GPU; all others are host pointers.
 *
 * Multi-RHS blocks are row-major N x k ("DOF-major, right-hand sides
 * contiguous"): entry (dof i, rhs c) is at i*k + c.
 *
 * Each declaration cites the reference interface it replaces (paths relative to
 * the statFEM tree).
 *
 * Threading: the work of a context, and of the systems, likelihoods and
 * networks made on it, is ordered on the context's stream; call one context
 * from one host thread at a time (a context keeps scratch between calls, e.g.
 * for the batched products). Separate contexts may be used from separate
 * threads.
 */
#ifndef STATFEM_HIP_H
#define STATFEM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
enum {
    SFH_OK = 0,
    SFH_EARG = -1,        /* bad argument / size */
    SFH_ENEGJAC = -2,     /* "Negative Jacobian." corlib/ElementBasic.ipp:199 */
    SFH_ESINGULAR = -3,   /* singular system, corlib/SystemSolveEigenSparse.hpp:303-305: a non-finite
                             residual, or an iterative solve whose FP64 residual stopped falling (an
                             operator without the constraints that make it invertible, or a right-hand
                             side outside its range) */
    SFH_ENOTSPD = -4,     /* LLT failure, statX/infer/LikelihoodKOH.ipp:93 */
    SFH_EQUADFORM = -5,   /* non-negative quadratic form, statX/infer/LikelihoodKOH.ipp:113 */
    SFH_ENOCONV = -6,     /* iterative solve hit its iteration cap while its residual still fell */
    SFH_EDEVICE = -7,     /* HIP runtime error */
    SFH_ENOMEM = -8,      /* device allocation failed */
    SFH_ESTATE = -9       /* call out of order (e.g. solve before assembly) */
};

/* element kinds: corlib/Shapefun.ipp LINE2 :39-55, TRI3 :181-197, QUAD4 :460-476, HEX8 :926-975 */
enum { SFH_LINE2 = 1, SFH_TRI3 = 2, SFH_QUAD4 = 3, SFH_HEX8 = 4 };
/* element integrands */
enum {
    SFH_POISSON = 0,     /* del2/fem/ElementPotential.ipp:56-74, p1 = conductivity */
    SFH_ELASTICITY = 1,  /* solid/fem/ElementStatic.ipp:284-321 at F=I, p1 = E, p2 = nu */
    SFH_MATERN = 2       /* del2/fem/ElementMatern.ipp:57-82, p1 = kappa, p2 = diffusivity */
};

typedef struct sfh_ctx sfh_ctx;
typedef struct sfh_system sfh_system;
typedef struct sfh_lik sfh_lik;
typedef struct sfh_mlp sfh_mlp;
typedef struct sfh_lhs sfh_lhs;
typedef struct sfh_gpq sfh_gpq;
typedef struct sfh_fpert sfh_fpert;

const char* sfh_status_string(int status);
const char* sfh_last_error(void);
const char* sfh_version(void);

/* ----------------------------------------------------------------- context */
/* Opaque device + stream. No reference counterpart (the reference is CPU only). */
int sfh_ctx_create(int device, sfh_ctx** out);
int sfh_ctx_destroy(sfh_ctx* ctx);
int sfh_ctx_sync(sfh_ctx* ctx);
/* The context's device: ordinal, compute units and architecture name (e.g. "gfx950:sramecc+:xnack-",
 * truncated to len - 1 characters); keys of persisted launch-shape choices. */
int sfh_ctx_info(sfh_ctx* ctx, int* device, int* num_cu, char* arch, int len);
/* Device memory helpers for callers without a GPU array library. */
int sfh_malloc(sfh_ctx* ctx, size_t bytes, void** dptr);
int sfh_free(sfh_ctx* ctx, void* dptr);
int sfh_memcpy_h2d(sfh_ctx* ctx, void* dst_dev, const void* src, size_t bytes);
int sfh_memcpy_d2h(sfh_ctx* ctx, void* dst, const void* src_dev, size_t bytes);
int sfh_memcpy_d2d(sfh_ctx* ctx, void* dst_dev, const void* src_dev, size_t bytes);
int sfh_memset(sfh_ctx* ctx, void* dst_dev, int value, size_t bytes);
/* Load operator of a nodal scalar field acting on one displacement component,
 * with per-node weights (the lumped form of BodyForce + VectorAssembler,
 * corlib/BodyForce.hpp:66-71, corlib/VectorAssembler.hpp:101-114), and its
 * transpose. Blocks are row-major with k columns.
 * gather:  dst[j][c] = w[j] src[j*stride + comp][c]          (n x k from (n*stride) x k)
 * scatter: dst[j*stride + q][c] = (q == comp) w[j] src[j][c]  ((n*stride) x k from n x k) */
int sfh_dof_gather(sfh_ctx* ctx, int n, int k, int stride, int comp, const double* w_dev,
                   const double* src_dev, double* dst_dev);
int sfh_dof_scatter(sfh_ctx* ctx, int n, int k, int stride, int comp, const double* w_dev,
                    const double* src_dev, double* dst_dev);
/* Y = diag(w) X (n x k, may alias): the lumped-mass weights of the SPDE
 * covariance products (statX/inferQ/PrecisionMatrixRational.ipp:278-298). */
int sfh_row_scale(sfh_ctx* ctx, int n, int k, const double* w_dev, const double* X_dev, double* Y_dev);
/* Elapsed device time between two points on the context stream, in ms. */
int sfh_timer_start(sfh_ctx* ctx);
int sfh_timer_stop(sfh_ctx* ctx, double* ms);

/* ------------------------------------------------------------ FE system */
/* Replaces SystemSolveEigenSparse(numDofs) (corlib/SystemSolveEigenSparse.hpp:54-60)
 * together with Mesh<ELEMENT>(istream&) (corlib/Mesh.ipp:62-102) and the DOF
 * numbering of NodePotential::numberDOFs (del2/fem/NodePotential.hpp:65) /
 * NodeStatic::numberDOFs (solid/fem/NodeStatic.ipp:15-20): uploads the mesh and
 * builds, on the GPU, the CSR pattern that Triplet::finishAssembly produces
 * (corlib/Triplet.hpp:126-170,264-304) plus the element-to-CSR scatter map.
 * dofpn = 1 (scalar) or 3 (interleaved vector DOFs). */
int sfh_system_create(sfh_ctx* ctx, int kind, int dofpn, int nnodes, const double* coords,
                      int nel, const int* conn, sfh_system** out);
int sfh_system_destroy(sfh_system* sys);
/* A system from (row, col, value) triplets in insertion order, for callers that
 * insert element matrices one at a time through the SystemSolve concept
 * (SystemSolveEigenSparse::insertToMatrix, corlib/SystemSolveEigenSparse.hpp:171-182).
 * The CSR pattern and values are Triplet::finishAssembly's
 * (corlib/Triplet.hpp:126-170,264-339): rows in order, columns ascending,
 * equal keys merged by summation (in insertion order), explicit zeros kept.
 * Every row must hold a diagonal entry. Host arrays; no mesh, so the
 * mesh-based assembly and covariance entry points return SFH_ESTATE. */
int sfh_system_create_triplets(sfh_ctx* ctx, int ndof, long long n, const int* rows, const int* cols,
                               const double* vals, sfh_system** out);
/* New values for the same insertion sequence (clearMatrix + re-assembly,
 * corlib/Triplet.hpp:351-363): merged without sorting again. Clears the
 * constrained-DOF marks; apply constraints again afterwards. */
int sfh_system_set_triplet_values(sfh_system* sys, long long n, const double* vals);
/* The right-hand side F (host, ndof), replacing the assembled one. */
int sfh_system_set_rhs(sfh_system* sys, const double* F);
/* numNonZeros (corlib/SystemSolveEigenSparse.hpp:121) and size */
int sfh_system_info(sfh_system* sys, int* ndof, long long* nnz);
/* Triplet::setCompressedRowArrays pattern (corlib/Triplet.hpp:264-304), host copies */
int sfh_system_pattern(sfh_system* sys, int* rowptr, int* colind);
/* Declares the mesh a tensor grid (x fastest) with the given axis node
 * coordinates, enabling the separable forcing-covariance product. Verified
 * against the node coordinates. */
int sfh_system_set_grid(sfh_system* sys, int nx, int ny, int nz, const double* gx,
                        const double* gy, const double* gz);

/* Whole-mesh MatrixComputeAndAssembleFun + mesh.iterateOverElements(maf)
 * (corlib/MatrixAssembler.hpp:42-100,198-232 over ComputeElementMatrix,
 * corlib/ComputeElementMatrix.hpp:62-77). Overwrites the matrix values.
 * p1_dev/p2_dev: per-element parameters (nel), or NULL to use s1/s2 for all. */
int sfh_assemble_matrix(sfh_system* sys, int physics, int nq1d, const double* p1_dev,
                        const double* p2_dev, double s1, double s2, double factor);
/* BodyForce + Integrator + VectorAssembler into the RHS (corlib/BodyForce.hpp:66-71,
 * corlib/Integrator.hpp:62-74, corlib/VectorAssembler.hpp:101-114, integrand
 * del2/fem/ElementPotential.ipp:146-175). f(x) = fconst[c] + sum_K fnodal[K*dofpn+c] phi_K(x).
 * accumulate = 0 overwrites F. */
int sfh_assemble_rhs(sfh_system* sys, int nq1d, const double* fconst, const double* fnodal_dev,
                     double factor, int accumulate);
/* Unit lumped mass int phi_i (ElementMatern::lumpedMassIntegrand,
 * del2/fem/ElementMatern.ipp:96-119; CovarianceForcing's covVec,
 * del2/apps/reference/laplaceRHSinfer/inference.hpp:103-114) into out_dev (ndof). */
int sfh_lumped_mass(sfh_system* sys, int nq1d, double* out_dev);
/* SystemSolveEigenSparse::applyConstraints (corlib/SystemSolveEigenSparse.hpp:197-215):
 * row-zero, diagonal = scalar, F[c] = factor*value*scalar. values may be NULL (zeros). */
int sfh_apply_constraints(sfh_system* sys, int nc, const int* dofs, const double* values,
                          double factor, double scalar);
/* clearMatrix / clearRhs (corlib/SystemSolveEigenSparse.hpp:104-107) */
int sfh_clear_matrix(sfh_system* sys);
int sfh_clear_rhs(sfh_system* sys);
/* Host copies of the assembled values (row-zero form) and of F; device views. */
int sfh_matrix_values(sfh_system* sys, double* vals);
int sfh_rhs(sfh_system* sys, double* F);
int sfh_system_device_ptrs(sfh_system* sys, const int** rowptr_dev, const int** colind_dev,
                           const double** vals_dev, const double** F_dev);

/* ------------------------------------------------------------------ solve */
typedef struct {
    double tol;          /* FP64 relative residual target per RHS (default 1e-10) */
    double inner_tol;    /* FP64 residual replacement once the FP32 residual fell by this factor
                            (default 0: 0.01 with the Jacobi preconditioner; 1e-14 with multigrid, whose
                            FP64 recurrences need no replacement before the target) */
    int max_outer;       /* cap on residual replacements (default 500) */
    int max_inner;       /* cap on PCG iterations (default 50000) */
    int check_every;     /* inner iterations between host convergence checks
                            (default 0: 32 with the Jacobi preconditioner, 2 with multigrid) */
    int precision;       /* 0: FP32 recurrences + FP64 replacement; 1: FP64 throughout */
    int rhs_tile;        /* >0: solve the right-hand sides in column tiles of this width (0: all at once) */
    int layout;          /* 0 (default): auto, 1: CSR, 2: 9-coefficient stencil with LDS-tiled kernels on
                            declared 2D tensor grids (FP32, at most 64 right-hand sides per tile);
                            3: 3 x 3 node blocks for vector-DOF meshes (FP32, one problem);
                            4: 27-point block stencil on a declared 3D grid (1 or 3 DOFs per node, one
                            problem), with the multigrid preconditioner;
                            auto uses 2, 4 or 3 where they apply, else CSR */
    int precond;         /* 0 (default): auto, 1: Jacobi, 2: geometric multigrid V(1,1)-cycle (bi- or
                            trilinear prolongation, Galerkin coarse stencils, l1-Jacobi smoothing (nodal
                            block Jacobi for 3D vector DOFs), dense
                            coarsest solve) with FP64 recurrences, on declared 2D grids (one or a batch)
                            and declared 3D grids; auto uses 2 wherever it applies, else 1 */
    int graph;           /* 1 (default): on declared 2D grids with multigrid, replay each pair of iterations
                            between host convergence checks as a captured HIP graph (captured once per
                            workspace and launch shape, replayed across solves); 0: launch kernel by kernel */
    int lanes;           /* lanes per node of the iteration kernels, 4 right-hand sides per lane: 0
                            (default) automatic (up to 16 on the stencil layouts, 64 on CSR), else 1, 2, 4,
                            8 or 16 (a power of two at most this). Narrow column tiles let the work-groups
                            of a batch skip the padding and converged columns of each problem */
    int mg_split;        /* 2D multigrid V-cycle: 0 (default) runs the coarse levels from the first one
                            below which every grid's remaining hierarchy fits one work-group's LDS as a
                            single launch (one work-group per column and grid); 1 launches a restriction
                            and a prolongation per level and the coarse solve separately */
} sfh_solve_opts;

typedef struct {
    int outer;            /* FP64 residual replacements */
    int inner_iters;      /* PCG iterations */
    double max_relres;    /* max over RHS of ||b - A x|| / ||b|| in FP64 */
    double ms;            /* device time */
    int converged;
    int layout;           /* operator layout the iteration kernels used: 1 CSR, 2 2D stencil, 3 3x3 blocks,
                             4 3D block stencil */
    int precond;          /* preconditioner used: 1 Jacobi, 2 multigrid */
    int levels;           /* multigrid levels including the fine grid (1 for Jacobi) */
    int graph_iters;      /* iterations run as HIP graph replays */
} sfh_solve_stats;

void sfh_solve_opts_default(sfh_solve_opts* o);
/* Multi-RHS solveSystem / solveSystemTranspose
 * (corlib/SystemSolveEigenSparse.hpp:318-360 and :364-404): X = A^-1 B (trans = 0)
 * or A^-T B (trans = 1) for the row-zero operator, k right-hand sides at once.
 * The batched form of ForwardGaussRHS's loop of transposed solves
 * (statX/infer/ForwardGaussRHS.hpp:105-112). B_dev, X_dev: N x k. */
int sfh_solve(sfh_system* sys, int trans, int k, const double* B_dev, double* X_dev,
              const sfh_solve_opts* opts, sfh_solve_stats* stats);
/* A batch of independent problems in one solve: the system holds their
 * block-diagonal union (disjoint meshes, rows seg_rows[q] .. seg_rows[q+1]-1 of
 * problem q, seg_rows host, nseg + 1 entries), and every (problem, column) pair
 * runs its own PCG scalars and convergence test in the same kernel launches.
 * The per-problem form of solveSystemTranspose / solveSystem
 * (corlib/SystemSolveEigenSparse.hpp:318-404) for many uneven meshes at once.
 * A problem with fewer right-hand sides leaves its extra columns of B zero. */
/* Declares that the batch's problems are 2D tensor grids with nodes numbered x
 * fastest: problem q has nx[q] x ny[q] nodes on rows seg_rows[q] ..
 * seg_rows[q+1]-1 (host arrays), so the batched solve can use the stencil
 * layout (sfh_solve_opts.layout) on every problem. */
int sfh_system_set_segment_grids(sfh_system* sys, int nseg, const int* seg_rows, const int* nx, const int* ny);
int sfh_solve_batched(sfh_system* sys, int nseg, const int* seg_rows, int trans, int k,
                      const double* B_dev, double* X_dev, const sfh_solve_opts* opts,
                      sfh_solve_stats* stats);
/* solveSystem() on the assembled F (corlib/SystemSolveEigenSparse.hpp:273-311). */
int sfh_solve_rhs(sfh_system* sys, double* X_dev, const sfh_solve_opts* opts,
                  sfh_solve_stats* stats);

/* Instrumentation, no reference counterpart: sampled device time of the
 * inner-PCG launches inside real solves. With every > 0, the inner iterations
 * numbered 0 and 1 modulo every (a pair, so that iterations with and without
 * the multigrid update's x step are sampled alike) of the following solves on
 * this context (at most max_samples per
 * solve; only solves of at least min_rhs right-hand sides per column tile) is
 * bracketed by HIP events on the context stream; the accumulators are reset.
 * every = 0 turns sampling off. sfh_kernel_times returns the average device
 * time (ms per iteration) of ms[0] direction + SpMV + p.Ap, ms[1] SpMV +
 * x/r update + r.z, r.r (the two launches of an iteration) and ms[2] the
 * multigrid V-cycle (0 with the Jacobi preconditioner, which is fused into
 * ms[1]), and the number of samples since the last sfh_kernel_sampling call. */
int sfh_kernel_sampling(sfh_ctx* ctx, int every, int max_samples, int min_rhs);
int sfh_kernel_times(sfh_ctx* ctx, double* ms /* 3 */, long long* samples);
/* Instrumentation, no reference counterpart: while sampling is on (every > 0
 * above), the device time of the matrix assembly kernels of each
 * sfh_assemble_matrix call (HIP events around them, the status read-back
 * excluded); returns their average in ms and their count since the last
 * sfh_kernel_sampling call. */
int sfh_assembly_times(sfh_ctx* ctx, double* ms, long long* calls);

/* ------------------------------------------------------------- statistics */
/* CovarianceForcing::giveProjectedCovar building block (statX/infer/CovarianceForcing.ipp:37-109):
 * V = C_f W with C_f(i,j) = covvec_i sigma^2 exp(-|x_i-x_j|^2/(2 ell^2)) covvec_j
 * over the system's node coordinates (scalar DOFs). mode 0: separable if a grid
 * is set, else dense; 1: dense matrix-free; 2: separable. W_dev, V_dev: N x k. */
int sfh_covforcing_apply(sfh_system* sys, const double* covvec_dev, double sigma, double ell,
                         int k, const double* W_dev, double* V_dev, int mode);
/* Gram product C = X^T V (N x m and N x n -> m x n), FP64 MFMA: the dense
 * R C_f R^T step (CovarianceForcing.ipp:106). */
int sfh_gram(sfh_ctx* ctx, int N, int m, int n, const double* X_dev, int ldx, const double* V_dev,
             int ldv, double* C_dev);

/* Per-problem Gram products of a batch: C[q] = X_q^T V_q (m x n each, device,
 * nseg blocks) over the row ranges seg_rows[q] .. seg_rows[q+1]-1 (host). */
int sfh_gram_segmented(sfh_ctx* ctx, int nseg, const int* seg_rows, int m, int n, const double* X_dev,
                       int ldx, const double* V_dev, int ldv, double* C_dev);
/* The separable CovarianceForcing product (statX/infer/CovarianceForcing.ipp:37-109)
 * on a tensor grid given by its axes (host, x fastest; ny = nz = 1 for lower
 * dimensions), independent of any system: V = C_f W, W_dev, V_dev: n x k. */
int sfh_covforcing_grid(sfh_ctx* ctx, int nx, int ny, int nz, const double* gx, const double* gy,
                        const double* gz, const double* covvec_dev, double sigma, double ell, int k,
                        const double* W_dev, double* V_dev);
/* The same product for every problem of a batch of 2D grids in one call (the
 * per-problem loop of CovarianceForcing over the problems of a union system,
 * statX/infer/CovarianceForcing.ipp:37-109): problem q owns rows
 * seg_rows[q] .. seg_rows[q+1]-1 of W_dev / V_dev / covvec_dev (row-major,
 * k columns), an nx[q] x ny[q] grid whose axes follow each other in axes
 * (host: nx[q] x coordinates, then ny[q] y coordinates, problem after
 * problem), and its own sigma[q], ell[q] (host). */
int sfh_covforcing_batched(sfh_ctx* ctx, int nseg, const int* seg_rows, const int* nx, const int* ny,
                           const double* axes, const double* covvec_dev, const double* sigma,
                           const double* ell, int k, const double* W_dev, double* V_dev);
/* ForwardGaussRHS::updateSolution (statX/infer/ForwardGaussRHS.hpp:95-139) for a
 * sparse projection P given row-wise (ns rows of pn (col, val) pairs):
 *   R^T = A^-T P^T, mean = R F, Cu = R C_f R^T + (sigma_f reg_eps)^2 P P^T.
 * C_f's weights are the lumped mass zeroed on constrained DOFs
 * (CovarianceForcing.hpp:101-114). When field_dev is non-NULL it also returns the
 * field-by-sensor prior covariance C_u P^T = A^-1 C_f R^T + (sigma_f reg_eps)^2 P^T
 * (N x ns), and ubar_dev (N) = A^-1 F, the inputs of the Woodbury posterior.
 * Outputs mean (ns) and Cu (ns x ns) are host arrays. */
/* The weights covVec of CovarianceForcing (statX/infer/CovarianceForcing.hpp:70-114:
 * insertToRhs of the unit body force, zeroed by applyConstraints) that the
 * following sfh_forward_gauss calls use in C_f, host (ndof); NULL returns to
 * the default, the system's lumped mass zeroed on its constrained DOFs. */
int sfh_system_set_covvec(sfh_system* sys, const double* covvec);
int sfh_forward_gauss(sfh_system* sys, int ns, int pn, const int* pcols, const double* pvals,
                      double sigma_f, double ell_f, double reg_eps, const sfh_solve_opts* opts,
                      double* mean, double* Cu, double* field_dev, double* ubar_dev,
                      sfh_solve_stats* stats);

/* LikelihoodKOH + MismatchCovariance state (statX/infer/LikelihoodKOH.hpp:55-75,
 * statX/infer/MismatchCovariance.hpp:62-99). xs: ns x dim sensor coordinates,
 * Cu: ns x ns, mean: ns, Y: no x ns (snapshot-major = column-major ns x no),
 * eps = mismatch.sensorEps. */
int sfh_lik_create(sfh_ctx* ctx, int ns, int no, int dim, const double* xs, const double* Cu,
                   const double* mean, const double* Y, double eps, sfh_lik** out);
int sfh_lik_destroy(sfh_lik* lik);
/* giveLogLikelihood (statX/infer/LikelihoodKOH.ipp:74-120) for M hyperparameter
 * sets theta = (rho, mismatch.sigma, mismatch.length), M x 3 host; out, status: M. */
int sfh_lik_loglik_batch(sfh_lik* lik, int M, const double* theta, double* out, int* status);
/* log det(K_s + rho^2 Cu) for M hyperparameter sets: the log-determinant term of
 * giveLogLikelihood (statX/infer/LikelihoodKOH.ipp:93-100). Host arrays. */
int sfh_lik_logdet_batch(sfh_lik* lik, int M, const double* theta, double* logdet, int* status);
/* MismatchCovariance::giveCovariance (MismatchCovariance.hpp:76-99), ns x ns host */
int sfh_lik_mismatch(sfh_lik* lik, double sigma_d, double ell_d, double* Ks);
/* giveProjFEconditionedData (statX/infer/LikelihoodKOH.ipp:310-361) in the
 * equivalent Cholesky form; mean: ns, cov: ns x ns host. */
int sfh_lik_post_proj(sfh_lik* lik, double rho, double sigma_d, double ell_d, double* mean,
                      double* cov);
/* giveFEconditionedDataWoodbury mean (statX/infer/LikelihoodKOH.ipp:441-506) on
 * the whole field from field_dev = C_u P^T (N x ns) and ubar_dev (N); the
 * posterior variance at the sensors comes from sfh_lik_post_proj. */
int sfh_lik_post_field(sfh_lik* lik, double rho, double sigma_d, double ell_d, int N,
                       const double* field_dev, const double* ubar_dev, double* mean_dev);
/* giveFEconditionedDataWoodbury covariance (statX/infer/LikelihoodKOH.ipp:459-488):
 * C_{u|y} = D C_u = C_u - (C_u P^T) S^-1 (P C_u), S = C_u^proj + K_s/(rho^2 n_o),
 * from field_dev = C_u P^T (N x ns, sfh_forward_gauss) and the prior field
 * covariance. full = 0: prior_dev is the prior variance diag(C_u) (N) and
 * out_dev receives the posterior variance (N); full = 1: prior_dev is the
 * prior covariance C_u (N x N, the reference's updateSolution with P = I) and
 * out_dev the posterior covariance (N x N; may alias prior_dev). Device arrays. */
int sfh_lik_post_field_cov(sfh_lik* lik, double rho, double sigma_d, double ell_d, int N,
                           const double* field_dev, const double* prior_dev, int full, double* out_dev);

/* MCMC<3>::sample (statX/infer/MCMC.hpp:123-158) with the target of
 * hyperParamLearning (del2/apps/reference/laplaceRHSinfer/hyperParamLearning.hpp:172-190),
 * for `chains` independent chains in lockstep. theta0: chains x 3 (log space),
 * propsd: 3, prior_mode 0 = flat in x (+sum theta), 1 = reference enrichWithPrior,
 * 2 = log-normal prior N(theta; mu, var). samples (host, may be NULL):
 * chains x steps x 4 = (theta, log target); acc: chains. Streams: Philox4x32-10,
 * counter {step, chain, purpose, 0}, key {seed}. */
int sfh_mcmc(sfh_lik* lik, int chains, int steps, const double* theta0, const double* propsd,
             uint64_t seed, int prior_mode, const double* prior_mu, const double* prior_var,
             double* samples, double* acc);

/* LikelihoodKOH::giveTypicalReadings + NormalSampler::giveSample
 * (statX/infer/LikelihoodKOH.ipp:42-71, statX/infer/NormalSampler.hpp:46-72):
 * Y (nr x ns host) = rho mean + L z, L L^T = K_s + rho^2 Cu. z ~ N(0,1) from
 * Philox4x32-10, counter {reading, index, 3, 0}, key {seed} (Box-Muller, cosine). */
int sfh_lik_typical_readings(sfh_lik* lik, double rho, double sigma_d, double ell_d, int nr,
                             uint64_t seed, double* Y);

/* Y = A X for k right-hand sides with the assembled (row-zero) operator;
 * the sparse products of statX/inferQ/PrecisionMatrixRational.ipp:278-298. */
int sfh_spmv(sfh_system* sys, int k, const double* X_dev, double* Y_dev);
/* PrecisionMatrixRational::giveTau_ (statX/inferQ/PrecisionMatrixRational.ipp:258-275):
 * tau^2 = Gamma(nu) / (sigma^2 Gamma(nu + d/2) (4 pi)^{d/2} kappa^{2 nu}). */
int sfh_matern_tau(double sigma, double kappa, double nu, int dim, double* tau);
/* givePrecisionMatrix for integer beta = (nu + d/2)/2 applied to X
 * (statX/inferQ/PrecisionMatrixRational.ipp:24-39,177-209,278-298):
 * Y = tau^2 kappa^{4 beta} (A M^-1)^{2 alpha - 1} A X, alpha = max(1, floor(beta)),
 * A the system's assembled Matern matrix (M + kappa^-2 G: SFH_MATERN with
 * factor 1/kappa^2, DriverMaxLikelihood.hpp:202-230), mass_dev the lumped mass.
 * Fractional beta returns SFH_EARG here; sfh_precision_apply_rational applies
 * the fractional precision from its rational factors. */
int sfh_precision_apply(sfh_system* sys, const double* mass_dev, double sigma, double kappa,
                        double nu, int k, const double* X_dev, double* Y_dev);
/* giveLeftFracMatrix / giveRightFracMatrix (statX/inferQ/PrecisionMatrixRational.ipp:41-161)
 * applied to X: Y = lead prod_j (I - r_j M^-1 A) X, or with transpose != 0 its
 * transpose lead prod_j (I - r_j A M^-1) X; roots = (lead, r_0 .. r_{degree-1})
 * (host; top.out for P_r, bot.out for P_l), A the assembled Matern matrix. */
int sfh_frac_apply(sfh_system* sys, const double* mass_dev, int degree, const double* roots, int transpose, int k,
                   const double* X_dev, double* Y_dev);
/* givePrecisionMatrix for a fractional beta (PrecisionMatrixRational.ipp:24-39)
 * applied to X: Q_v X = P_l^T Q P_l X, Q the integer-part precision of
 * sfh_precision_apply with alpha = max(1, floor(beta)) and the scale
 * tau^2 kappa^{4 beta}, P_l from bot = (b, d_0 .. d_{degree-1}) (host). */
int sfh_precision_apply_rational(sfh_system* sys, const double* mass_dev, double sigma, double kappa, double nu,
                                 int degree, const double* bot, int k, const double* X_dev, double* Y_dev);

/* ---------------------------------------------------- learned surrogate
 * No reference counterpart. An MLP 3 -> 64 -> 64 -> 1 (tanh) of the
 * log-likelihood over theta = log(rho, sigma_d, ell_d); BF16 weights, FP32
 * accumulation, FP32 Adam masters; hand-written forward/backward kernels.
 * Parameter layout (4481): W1[64][3], b1[64], W2[64][64], b2[64], W3[64], b3.
 * norm8 = (mu_theta[3], sd_theta[3], mu_logL, sd_logL). */
int sfh_mlp_create(sfh_ctx* ctx, uint64_t seed, sfh_mlp** out);
int sfh_mlp_destroy(sfh_mlp* mlp);
int sfh_mlp_set_normalization(sfh_mlp* mlp, const double* norm8);
int sfh_mlp_params(sfh_mlp* mlp, float* master, uint16_t* bf16, double* norm8);
int sfh_mlp_set_params(sfh_mlp* mlp, const float* master);
/* gradient of (1/n) sum (net(x_s) - t_s)^2 in standardised units; grad: 4481 */
int sfh_mlp_grad(sfh_mlp* mlp, int n, const double* theta, const double* logl, float* grad, double* loss);
/* full-batch Adam with cosine-decayed learning rate; standardisation from the data */
int sfh_mlp_train(sfh_mlp* mlp, int n, const double* theta, const double* logl, int steps, double lr,
                  double* loss);
int sfh_mlp_predict(sfh_mlp* mlp, int n, const double* theta, double* logl);
/* Delayed-acceptance (surrogate-transition) MCMC: per outer step every chain
 * runs `sub` random-walk steps on the surrogate target (NN + prior terms of
 * sfh_mcmc), then the end point is accepted with the exact likelihood
 * (sfh_lik_loglik_batch) by min(1, pi(y) pi~(x) / (pi(x) pi~(y))), which keeps
 * the exact posterior invariant. Arguments as sfh_mcmc; sub_acc: surrogate
 * acceptance rate per chain; sub <= 256, else SFH_EARG. */
int sfh_mcmc_da(sfh_lik* lik, sfh_mlp* mlp, int chains, int steps, int sub, const double* theta0,
                const double* propsd, uint64_t seed, int prior_mode, const double* prior_mu,
                const double* prior_var, double* samples, double* acc, double* sub_acc);

/* GaussianProcess<KERNEL,N> fit / giveLogLikelihood / predictMean / predictVariance
 * (statX/infer/GaussianProcess.hpp:83-242) on the GPU. Host arrays. */
int sfh_gp(sfh_ctx* ctx, int n, int dim, const double* X, const double* y, double sigma,
           double ell, double eps, int np, const double* Xp, double* logl, double* mean,
           double* var);
/* Allocation-free form of sfh_gp, with no counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc). Inputs and outputs are
 * device arrays (X n x dim, y n, Xp np x dim; logl 1, mean and var np,
 * status 1 = SFH_OK or SFH_ENOTSPD), the work is enqueued on the context
 * stream and the call returns without waiting; the scratch (Cholesky factor,
 * weights, prediction block) is a fixed array of this GPU (about 4.7 MB, held
 * until the process ends), so n <= 256 and np <= 2048, else SFH_EARG. Calls
 * that overlap on one GPU race on that array. */
int sfh_gp_noalloc(sfh_ctx* ctx, int n, int dim, const double* X_dev, const double* y_dev, double sigma, double ell,
                   double eps, int np, const double* Xp_dev, double* logl_dev, double* mean_dev, double* var_dev,
                   int* status_dev);
/* GaussianProcess::predictCovariance (statX/infer/GaussianProcess.hpp:160-190):
 * cov (np x np, host) = k(X*, X*) - k(X, X*)^T (K + eps^2 I)^-1 k(X, X*) of the
 * process fitted as in sfh_gp. */
int sfh_gp_predict_cov(sfh_ctx* ctx, int n, int dim, const double* X, const double* y, double sigma,
                       double ell, double eps, int np, const double* Xp, double* cov);

/* ------------------------------------------- LHS (conductivity) uncertainty */
/* ForwardGaussLHS (statX/infer/ForwardGaussLHS.hpp:41-105) with the solve
 * function of laplaceInverse (del2/apps/reference/laplaceInverse/inference.hpp:72-120)
 * for nch conductivity fields at once. sys is the disjoint union of nch copies
 * of one scalar mesh (copy c: rows seg_rows[c] .. seg_rows[c+1]-1, elements
 * c nel .. (c+1) nel - 1), with its right-hand side assembled (the unit body
 * force, fem.hpp:19-27) and constrained; bc (host, nbc union DOFs) are the
 * homogeneous Dirichlet DOFs re-applied after every assembly (fem.hpp:31-43).
 * A hyperparameter set theta (log space) has na + 1 entries: the anchor values
 * w = exp(theta[0 .. na-1]) of the GaussianProcess<ExpKernel, na> whose mean at
 * each element centroid is the element's conductivity (inference.hpp:72-88;
 * GaussianProcess.hpp:117-135 with anchors Xa (na x dim, host), gp_sigma,
 * gp_ell, gp_eps), and the sensor noise mismatch.sensorEps = exp(theta[na]).
 * The projection P has ns rows of pn (column, value) pairs in one copy's DOF
 * numbering (readMatrixPhi, ioHelpers.hpp:88-123). na + 1 <= 16. */
int sfh_lhs_create(sfh_system* sys, int nch, const int* seg_rows, int nel, int na, const double* Xa,
                   double gp_sigma, double gp_ell, double gp_eps, int nq1d, int nbc, const int* bc, int ns,
                   int pn, const int* pcols, const double* pvals, sfh_lhs** out);
int sfh_lhs_destroy(sfh_lhs* lhs);
/* LikelihoodKOH + MismatchCovariance data and the prior of hyperParamLearning
 * (del2/apps/reference/laplaceInverse/hyperParamLearning.hpp:126-150): readings
 * Y (no x ns, host, reading-major), sensor coordinates xs (ns x sdim, host),
 * rho, mismatch.sigma, mismatch.length; prior != 0 adds the normal prior
 * density of (theta[0 .. na-1], exp(theta[na])) with means prior_mu and
 * variances prior_var (na + 1 each, host; enrichWithPrior). */
int sfh_lhs_set_data(sfh_lhs* lhs, int no, const double* Y, int sdim, const double* xs, double rho,
                     double mis_sigma, double mis_ell, int prior, const double* prior_mu,
                     const double* prior_var);
/* ForwardGaussLHS::updateSolution (ForwardGaussLHS.hpp:83-92) for the nch sets
 * theta (nch x (na + 1), host): conductivity, assembly, constraints, one
 * batched solve, projection. mean (host, nch x ns) may be NULL. SFH_ENOTSPD
 * when a set interpolates a non-positive conductivity. */
int sfh_lhs_update(sfh_lhs* lhs, const double* theta, const sfh_solve_opts* opts, double* mean,
                   sfh_solve_stats* stats);
/* keep copy `copy`'s projected mean of the last update as the mean of every
 * later evaluation with resolve = 0: the reference's laplaceInverse computes
 * the FE solution once, at the hyperparameters it starts from
 * (inference.hpp:139), and its likelihood reuses it at every MCMC step
 * (LikelihoodKOH.ipp:82-84). */
int sfh_lhs_fix_mean(sfh_lhs* lhs, int copy);
/* the log target posteriorGivenData (hyperParamLearning.hpp:172-190) of M sets
 * theta (M x (na + 1), host): the KOH log-likelihood with C_u = 0, plus
 * theta[na], plus the log prior. resolve = 1 re-solves each set's field
 * (M must equal nch); resolve = 0 uses the fixed mean (any M). out, status:
 * M (host); -inf where the likelihood or the conductivity failed. */
int sfh_lhs_log_target(sfh_lhs* lhs, int M, const double* theta, int resolve, const sfh_solve_opts* opts,
                       double* out, int* status, sfh_solve_stats* stats);
/* MCMC<na + 1>::sample (statX/infer/MCMC.hpp:123-158) of that target for nch
 * chains in lockstep from theta0 (nch x (na + 1), host), random-walk proposal
 * theta + propsd .* z; chain c's draws come from Philox counters
 * {step, c, 32 + i, 0} (normals, Box-Muller pairs) and {step, c, 31, 0}
 * (acceptance uniform) under the key `seed`. resolve = 1 re-interpolates,
 * re-assembles and re-solves every proposal (all chains in one batched solve);
 * resolve = 0 keeps the fixed mean as the reference does. samples (host,
 * nch x steps x (na + 2): the state after each step and its log target) may
 * be NULL; acc (host, nch) acceptance ratios; stats sums the solves. */
int sfh_lhs_mcmc(sfh_lhs* lhs, int steps, const double* theta0, const double* propsd, uint64_t seed,
                 int resolve, const sfh_solve_opts* opts, double* samples, double* acc,
                 sfh_solve_stats* stats);

/* ------------------------------------------ sparse-precision GP (SPDE) */
/* GaussianProcessQ (statX/inferQ/GaussianProcessQ.hpp:65-159) on the scalar
 * mesh of sys with the precision of DriverMaxLikelihood
 * (statX/inferQ/DriverMaxLikelihood.hpp:202-312):
 *   Q = tau^2 kappa^(4 beta) (A M^-1)^(2 alpha - 1) A, beta = (nu + d/2) / 2,
 *   alpha = max(1, floor beta),
 * A the Matern matrix assembled by the library (SFH_MATERN, factor 1/kappa^2,
 * no constraints), M its lumped mass (nq1d-point rule), P the projection
 * (ns rows of pn (column, value) pairs, host), readings Y (no x ns, host,
 * reading-major). Integer beta: P_r = P_l = I; a fractional beta needs its
 * rational factors (sfh_gpq_set_rational). Declare a tensor grid
 * (sfh_system_set_grid) to solve with multigrid. */
int sfh_gpq_create(sfh_system* sys, int nq1d, int ns, int pn, const int* pcols, const double* pvals, int no,
                   const double* Y, sfh_gpq** out);
int sfh_gpq_destroy(sfh_gpq* gpq);
/* The rational approximation x^exponent ~ a prod (x - c_i) / (b prod (x - d_j))
 * for exponent = beta - alpha, as PrecisionMatrixRational::readRoots_ reads it
 * from top.out and bot.out (statX/inferQ/PrecisionMatrixRational.ipp:300-372):
 * top = (a, c_0 .. c_{degree-1}), bot = (b, d_0 .. d_{degree-1}), host. It
 * defines P_l = b prod_j (I - d_j M^-1 A) (giveLeftFracMatrix, :41-98),
 * P_r = a prod_i (I - c_i M^-1 A) (giveRightFracMatrix, :100-161),
 * Q_v = P_l^T Q P_l (givePrecisionMatrix, :24-39) and u = P_r v
 * (GaussianProcessQ.ipp:54,140,182). The poles d_j must be negative (each
 * costs one solve with the SPD matrix M - d_j A). One set per exponent; a
 * second call with the same exponent replaces it. */
int sfh_gpq_set_rational(sfh_gpq* gpq, double exponent, int degree, const double* top, const double* bot);
/* giveLogLikelihood (GaussianProcessQ.ipp:37-99) for M hyperparameter sets
 * hps (M x 4, host: gaussianProcess.sigma, kappa, nu, eps), computed through
 * G = P Q^-1 P^T (determinant lemma, push-through identity) from multi-RHS
 * solves with A instead of two sparse Cholesky factorisations; sets sharing
 * (sigma, kappa, nu) share one set of solves. out, status: M (host). */
int sfh_gpq_loglik(sfh_gpq* gpq, int M, const double* hps, const sfh_solve_opts* opts, double* out, int* status,
                   sfh_solve_stats* stats);
/* predictMean / predictCovariance (GaussianProcessQ.ipp:101-186) for one set
 * hp (4, host): mean_dev (N) and/or cov_dev (N x N, dense; N + 1 solves with
 * N columns, for moderate N); logl (host, nullable) the log-likelihood. */
int sfh_gpq_predict(sfh_gpq* gpq, const double* hp, const sfh_solve_opts* opts, double* mean_dev, double* cov_dev,
                    double* logl, sfh_solve_stats* stats);

/* ------------------------------------ first-order perturbation (LHS + RHS) */
/* ForwardPerturbation (statX/infer/ForwardPerturbation.hpp:43-215) on the
 * assembled and constrained system sys (its right-hand side is f): the
 * projection P (ns rows of pn (column, value) pairs, host), nk material
 * parameters with their stiffness derivatives D_e = dA/dk_e as dense
 * nb x nb blocks on nb DOFs each (dofs: nk x nb, blk: nk x nb x nb, host;
 * setStiffnessDerivatives, :84-89) and their coordinates for the
 * CovarianceMaterial kernel (pcoords: nk x pdim, CovarianceMaterial.hpp:85-101),
 * reg_factor of the constructor (:66-70). nk nb <= 65536; dense N x N and
 * (nk nb)^2 work matrices, so for moderate meshes. */
int sfh_fpert_create(sfh_system* sys, int nq1d, int ns, int pn, const int* pcols, const double* pvals, int nk, int nb,
                     const int* dofs, const double* blk, int pdim, const double* pcoords, double reg_factor,
                     sfh_fpert** out);
int sfh_fpert_destroy(sfh_fpert* fp);
/* updateSolution (:108-190) with forcing.sigma/length and material.sigma/
 * length: mean = P u (ns, host), cov = R C_f R^T + sum_ed C_k[e,d] R D_e
 * (A^-1 C_f A^-T + u u^T) D_d^T R^T + (reg sigma_k)^2 P P^T (ns x ns, host),
 * with the full A^-1 (the reference fills n_s rows of it; equal when
 * n_s = N). u_dev (N, nullable) receives the FE solution. */
int sfh_fpert_update(sfh_fpert* fp, double sigma_f, double ell_f, double sigma_k, double ell_k,
                     const sfh_solve_opts* opts, double* mean, double* cov, double* u_dev, sfh_solve_stats* stats);

#ifdef __cplusplus
}
#endif
#endif /* STATFEM_HIP_H */
