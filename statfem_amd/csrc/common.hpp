// Internal declarations shared by the HIP translation units of libstatfem_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/statfem_hip.h"

namespace sfh {

void set_error(const char* fmt, ...);

#define SFH_CHECK_HIP(expr)                                                              \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            ::sfh::set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr,                  \
                             hipGetErrorString(e_));                                     \
            return SFH_EDEVICE;                                                          \
        }                                                                                \
    } while (0)

#define SFH_TRY(expr)                                                                    \
    do {                                                                                 \
        int s_ = (expr);                                                                 \
        if (s_ != SFH_OK) return s_;                                                     \
    } while (0)

#define SFH_LAUNCH_CHECK() SFH_CHECK_HIP(hipGetLastError())

// Element traits (corlib/Shapefun.ipp node orders, [0,1]^d reference cells).
__host__ __device__ inline int kind_dim(int kind) {
    return kind == SFH_LINE2 ? 1 : (kind == SFH_HEX8 ? 3 : (kind == SFH_TRI3 || kind == SFH_QUAD4 ? 2 : 0));
}
__host__ __device__ inline int kind_nen(int kind) {
    return kind == SFH_LINE2 ? 2 : kind == SFH_TRI3 ? 3 : kind == SFH_QUAD4 ? 4 : kind == SFH_HEX8 ? 8 : 0;
}
inline int num_qp(int kind, int nq1d) {
    if (kind == SFH_TRI3) return 1;
    int d = kind_dim(kind), n = 1;
    for (int i = 0; i < d; ++i) n *= nq1d;
    return n;
}

// Device buffer owned by a handle; grows on demand, never shrinks.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    int ensure(size_t b);
    void release();
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Page-locked host buffer (hipHostMalloc), grown on demand: device-to-host
// copies into it are plain DMA transfers, where copies into pageable memory are
// staged by the runtime and cost tens of microseconds more each.
struct PinnedBuf {
    void* p = nullptr;
    size_t bytes = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() { release(); }
    int ensure(size_t b);
    void release();
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Captured HIP graphs of the multigrid iteration pairs of one workspace (the
// host checks convergence every two iterations): the first pair of a cycle
// and the pairs starting at j = 2, 4, 6 mod 6 (the accumulator slots rotate
// with j mod 2 and j mod 3). `key` lists every launch argument and shape the
// graphs were captured with; a solve with another key captures them again.
struct GraphCache {
    std::vector<unsigned long long> key;
    hipGraphExec_t exec[4] = {nullptr, nullptr, nullptr, nullptr};
    GraphCache() = default;
    GraphCache(const GraphCache&) = delete;
    GraphCache& operator=(const GraphCache&) = delete;
    ~GraphCache() { clear(); }
    void clear() {
        for (auto& e : exec)
            if (e) { (void)hipGraphExecDestroy(e); e = nullptr; }
        key.clear();
    }
};

// Multi-RHS mixed-precision PCG workspace (per system, sized for N x kp).
struct SolveWork {
    int kp = 0;
    DevBuf b64, x64, r64;            // FP64 refinement state
    DevBuf x32, r32, p32, q32;       // FP32 inner PCG (or FP64 when precision == 1)
    DevBuf part0, part1;             // per-block partial dot products
    DevBuf scal;                     // per-RHS scalars
    DevBuf acc;                      // per-column FP64 accumulators of the iteration kernels
    DevBuf tab;                      // row-partition tables of the current solve
    DevBuf mgv;                      // multigrid: restricted residuals and corrections of the coarse levels
    DevBuf cnb;                      // rows with a constrained neighbour (forward solves)
    GraphCache graphs;               // multigrid iteration pairs as HIP graphs
    PinnedBuf host;                  // the per-column state read back at each host check
    std::vector<int> tab_host;       // the partition tables last uploaded to `tab`
};

// Geometric multigrid hierarchies of a batch of 2D tensor grids (one grid for
// a single problem): level 0 of grid q is its W x H node grid in the operator
// (stencil in sfh_system::dia32 / dia64), level l + 1 keeps every other node of
// level l. The levels of all grids are stored side by side: grid q's level-l
// nodes start at node gbase[l * nseg + q] of level l's arrays (level 0: the
// grid's first row of the operator). Per level the FP32 stencil (levels >= 1)
// and the smoother weights; per grid the dense FP32 inverse of its coarsest
// level.
constexpr int kMGMaxLev = 16;
constexpr int kMGCoarse = 160;       // coarsest level: at most this many nodes
// Coarsening of one grid axis of n nodes: the coarse axis keeps every even
// fine node and, when n is even, the last fine node as well (n / 2 + 1 coarse
// nodes, the last coarse interval half as long), so that both ends of the axis
// are coarse nodes and the interpolation is linear up to the edge. (Dropping
// the last node and extending the last coarse value as a constant makes the
// coarse correction weak near that edge; on grids whose levels have even node
// counts it roughly doubles the PCG iterations.)
__host__ __device__ inline int mg_coarse_n(int n) { return (n & 1) ? (n + 1) / 2 : n / 2 + 1; }
// weight of coarse index X in the interpolated value at fine index x (n fine
// nodes): 1 for the coinciding coarse node (even x, and the last node of an
// even axis), 1/2 for the two neighbours of an odd interior x
__host__ __device__ inline float mg_axis_w(int x, int X, int n) {
    if ((x & 1) == 0) return x == 2 * X ? 1.f : 0.f;
    if (x == n - 1) return (x + 1) / 2 == X ? 1.f : 0.f;
    return ((x - 1) / 2 == X || (x + 1) / 2 == X) ? 0.5f : 0.f;
}
// coarse parents of fine index x: the first parent p0, their count (1 or 2:
// p0 and p0 + 1) and the weight of each
__host__ __device__ inline int mg_axis_parents(int x, int n, int& p0, float& w) {
    if ((x & 1) == 0) { p0 = x / 2; w = 1.f; return 1; }
    if (x == n - 1) { p0 = (x + 1) / 2; w = 1.f; return 1; }
    p0 = (x - 1) / 2; w = 0.5f;
    return 2;
}
// smallest pivot of the coarsest level's factorisation, relative to its row's
// diagonal, that counts as positive (a singular operator leaves pivots at the
// rounding level of the diagonal)
constexpr double kPivotRel = 1e-12;
// allocation-free dense GP (sfh_gp_noalloc): at most kGpFixedN observations
// and kGpFixedNP prediction points, scratch of L, alpha and L^-1 k*
constexpr int kGpFixedN = 256, kGpFixedNP = 2048;
constexpr size_t kGpFixedScratch = (size_t)kGpFixedN * kGpFixedN + kGpFixedN + (size_t)kGpFixedN * kGpFixedNP;
constexpr int kMaxDevices = 64;
struct MGHier {
    bool built = false;              // current for the stencil copy (reset when that is rebuilt)
    int L = 0;                       // coarse levels of the deepest grid (0: no hierarchy)
    int nseg = 0;
    std::vector<int> segL;           // coarse levels of every grid (at least 1 when L > 0)
    std::vector<int> gW, gH, gbase;  // per (level, grid), index l * nseg + q
    long long N[kMGMaxLev + 1] = {}; // nodes of every level (all grids)
    DevBuf coef64[kMGMaxLev + 1], coef32[kMGMaxLev + 1], dsm[kMGMaxLev + 1];
    // levels >= 1: the FP32 stencil coefficient-major (9 rows of N[l]), read
    // by the single-launch coarse levels with one coalesced load per coefficient
    DevBuf coefT32[kMGMaxLev + 1];
    DevBuf cinv;                     // coarsest-level inverses, grid after grid
    std::vector<int> cinv_off;
    int cmax = 0;                    // nodes of the largest coarsest level
    // set-up tables (per level: the grids coarsened from it; per coarsest
    // level: the grids whose inverse it holds), FP64 work and SPD flags
    DevBuf gtab, cwork, cbad;
};

// Multigrid hierarchy of a declared 3D grid (1 or 3 DOFs per node): the
// operator as a 27-point block stencil (FP64 and FP32), the coarse levels'
// Galerkin stencils, l1-Jacobi weights, and the coarsest level's inverse.
struct MG3Hier {
    bool built = false;              // current for the assembled operator
    int L = 0, dpn = 0;              // coarse levels (0: no hierarchy), DOFs per node
    int W[kMGMaxLev + 1] = {}, H[kMGMaxLev + 1] = {}, D[kMGMaxLev + 1] = {};
    DevBuf dia64, dia32;
    DevBuf coef64[kMGMaxLev + 1], coef32[kMGMaxLev + 1], dsm[kMGMaxLev + 1];
    DevBuf cinv;
    int nc = 0;                      // DOFs of the coarsest level
    double lam[kMGMaxLev + 1] = {};  // DPN 3: lambda_max(B^-1 A) estimate per smoothed level
};

// per-solve layout of the 3D V-cycle: coarse-level vectors in the
// workspace's multigrid buffer, coarse-solve table, launch shapes
struct G3Plan {
    int kp = 0, lpr = 0, ntile = 0, nsmooth = 0;
    size_t off[kMGMaxLev + 1] = {}, total = 0;
    int ctab[4] = {};
    const int* d_ctab = nullptr;
};

}  // namespace sfh

struct sfh_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;               // concurrent second solves
    hipEvent_t ev0 = nullptr, ev1 = nullptr;    // internal (solve timing)
    hipEvent_t tev0 = nullptr, tev1 = nullptr;  // public timer
    int num_cu = 0;
    size_t lds_per_block = 0;
    char arch[64] = {0};
    std::vector<void*> allocs;
    // sampled in-solve timing of the inner-PCG launches (direction + SpMV,
    // SpMV + update, multigrid V-cycle): every ksample-th iteration is bracketed by events
    // (0 = off)
    int ksample = 0;
    int kmin_rhs = 0;
    std::vector<hipEvent_t> kev;     // 4 events per sample
    double ksum[3] = {0.0, 0.0, 0.0};
    long long kcount = 0;
    // while sampling is on: device time of the matrix assembly kernels of each
    // sfh_assemble_matrix call (events created on first use)
    hipEvent_t aev0 = nullptr, aev1 = nullptr;
    double asum = 0.0;
    long long acount = 0;
    // device address of this GPU's fixed GP scratch (sfh_gp_noalloc)
    double* gp_scratch = nullptr;
    // kept between calls of the batched dense products: tables, axis factors
    // and row factors; the intermediate block; split-K partial sums; their tables
    sfh::DevBuf bat_tab, bat_blk, bat_part, bat_ptab;
};

struct sfh_system {
    sfh_ctx* ctx = nullptr;
    int kind = 0, dofpn = 1, nnodes = 0, nel = 0, nen = 0, dim = 0, ndof = 0;
    long long nnz = 0;
    long long nnz_node = 0;
    int maxrow = 0;              // longest DOF row
    sfh::DevBuf coords, conn;
    sfh::DevBuf n2e_ptr, n2e;    // node -> (element<<3 | local node) adjacency, element order
    sfh::DevBuf emap;            // per adjacency slot x nen: position of the element node in the node row
    sfh::DevBuf rowptr_node, colind_node;
    // two-phase hex8 elasticity: element matrices (36 blocks of 3 x 3 each) and
    // the per-(node, neighbour) contribution lists, built at the first assembly
    sfh::DevBuf el_ke, el_cptr, el_cent;
    bool el_ready = false;
    // declared 2D or 3D grid whose quad4 / hex8 elements are its cells
    // (checked once per grid declaration): the element of every cell, for the
    // closed-form assembly (empty when the elements are in cell order); state
    // 0 unchecked, 1 usable, -1 not
    sfh::DevBuf gcell;
    int gcell_state = 0;
    bool gcell_identity = false;
    // status word of the assembly kernels and its page-locked read-back
    // (kept across calls: no allocation per assembly)
    sfh::DevBuf asm_err;
    sfh::PinnedBuf asm_host;
    sfh::DevBuf rowptr, colind, diagpos;
    sfh::DevBuf vals, vals32, F, dinv32, dinv64;
    sfh::DevBuf diag32, diag64;  // A_ii (the PCG keeps z = D^-1 r and needs both)
    sfh::DevBuf cv32;            // FP32 operator as packed (column, value bits) pairs
    sfh::DevBuf bsr_val;         // vector DOFs: 9 FP32 values per node pair (node-level pattern)
    // systems built from triplets: sorted-to-insertion permutation and the
    // first sorted position of every merged entry (nnz + 1)
    sfh::DevBuf coo_perm, coo_start;
    long long ncoo = 0;
    sfh::DevBuf cmask;           // 1 on constrained DOFs
    sfh::DevBuf cdofs;           // constrained DOF list
    int nc = 0;
    bool assembled = false;
    bool dirty = true;           // preconditioner needs rebuilding
    // tensor grid (separable covariance)
    int gn[3] = {0, 0, 0};
    sfh::DevBuf gaxis[3];
    int nq1d = 2;                // quadrature of the last matrix assembly
    // stencil (DIA-9) copy of the operator for 2D tensor grids: coefficient k of
    // node (x,y) couples (x + k%3 - 1, y + k/3 - 1); built lazily per assembly
    sfh::DevBuf dia32, dia64;
    int dia_state = 0;           // 0 stale, 1 valid, -1 mesh does not fit the stencil
    std::vector<int> sgrid;      // batch of 2D grids: (first row, W, H) per problem
    sfh::DevBuf dia_tab;         // the same table on the device (stencil-copy build)
    sfh::MGHier mg;              // built with the stencil copy of declared 2D grids
    sfh::MG3Hier mg3;            // declared 3D grid
    sfh::SolveWork work;
    sfh::SolveWork work2;        // workspace of a concurrent second solve
    sfh::DevBuf scratch0, scratch1, scratch2, scratch3;
    sfh::DevBuf fg_pt, fg_rt, fg_v, fg_cv, fg_cu, fg_gram, fg_cols, fg_vals;
    sfh::PinnedBuf fg_host;          // C_u and the mean on their way to the caller
    // forcing-covariance weights given by the caller (sfh_system_set_covvec)
    sfh::DevBuf fg_user_cv;
    bool fg_user = false;
};

struct sfh_lik {
    sfh_ctx* ctx = nullptr;
    int ns = 0, no = 0, dim = 0;
    double eps = 0.0;
    sfh::DevBuf xs, Cu, mean, Y;
    sfh::DevBuf work, theta, out, status, ybar;
    sfh::DevBuf mc_theta, mc_y, mc_acc, mc_samples, mc_prop, mc_yprop, mc_stat;
    std::vector<double> Cu_h, mean_h, Y_h, xs_h;
};

// ---- entry points between translation units (host side)
namespace sfh {
int build_pattern(sfh_system* s);
int assemble_matrix(sfh_system* s, int physics, int nq1d, const double* p1, const double* p2,
                    double s1, double s2, double factor);
int assemble_rhs(sfh_system* s, int nq1d, const double* fconst, const double* fnodal,
                 double factor, int accumulate, double* out);
int apply_constraints(sfh_system* s, int nc, const int* dofs, const double* values, double factor,
                      double scalar);
int apply_constraints_dev(sfh_system* s, int nc, const int* dofs_dev, double factor, double scalar);
// nseg > 1: a block-diagonal batch of independent problems whose rows are
// seg_rows[q] .. seg_rows[q+1]-1; every (problem, column) pair runs its own
// PCG scalars and convergence test
int solve(sfh_system* s, int trans, int k, const double* B, int ldb, double* X, int ldx,
          const sfh_solve_opts* o, sfh_solve_stats* st, int nseg = 1, const int* seg_rows = nullptr);
// the same solve on another stream and workspace (a concurrent solve on the
// system from another host thread); prepare_operator first, on the context
// stream, builds what both read
int solve_on(sfh_system* s, hipStream_t st, SolveWork& w, int trans, int k, const double* B, int ldb, double* X,
             int ldx, const sfh_solve_opts* o, sfh_solve_stats* stats);
int prepare_operator(sfh_system* s, int nseg = 1, const int* seg_rows = nullptr);
// multigrid hierarchies of the system's grids (first row, W, H per grid) from
// its FP64 stencil copy (s->mg.L stays 0 when every grid is small enough to be
// preconditioned by Jacobi alone, or when a coarsest level is not SPD)
int mg_build(sfh_system* s, hipStream_t st, const std::vector<int>& grids);
// 3D grids: whether the system is a declared 3D grid the hierarchy applies to;
// build it (s->mg3.L stays 0 when the operator is not 27-point on the grid or
// the coarsest level is not SPD); plan and launch the iteration and V-cycle
bool g3_applies(const sfh_system* s);
int g3_prepare(sfh_system* s, hipStream_t st);
void g3_plan(G3Plan& P, const MG3Hier& m, int kp, int lpr);
void g3_dir(const G3Plan& P, const MG3Hier& m, hipStream_t st, int it, double* S, const float* z, const float* pold,
            float* pnew, double* acc);
void g3_update(const G3Plan& P, const MG3Hier& m, hipStream_t st, int it, const double* S, const float* p, double* x,
               double* r, float* rf, double* acc);
void g3_vcycle(const G3Plan& P, const MG3Hier& m, hipStream_t st, int it, int mode, const unsigned char* cmask,
               const float* r, float* z, float* mgv, double* acc, double* p0, double* p1);
int triplets_build(sfh_system* s, long long n, const int* rows, const int* cols, const double* vals);
int triplets_values(sfh_system* s, long long n, const double* vals);
int covforcing_grid(sfh_ctx* c, int d, const int gn[3], const double* const ax[3], const double* cv, double sigma,
                    double ell, int k, const double* W, double* V, sfh::DevBuf& sc0, sfh::DevBuf& sc1,
                    sfh::DevBuf& sc2);
int covforcing_apply(sfh_system* s, const double* covvec, double sigma, double ell, int k,
                     const double* W, double* V, int mode);
int gram(sfh_ctx* c, int N, int m, int n, const double* X, int ldx, const double* V, int ldv,
         double* C, DevBuf& work);
// the per-problem products of a batch in one launch each (rows seg[q]..seg[q+1]-1
// belong to problem q): C_q = X_q^T V_q (m x n, split-K, fixed-order sums), and
// V_q = sigma_q^2 D (Ky_q (x) Kx_q) D W_q on each problem's nx x ny grid
int gram_segmented(sfh_ctx* c, int nseg, const int* seg, int m, int n, const double* X, int ldx, const double* V,
                   int ldv, double* C);
int covforcing_batched(sfh_ctx* c, int nseg, const int* seg, const int* nx, const int* ny, const double* axes,
                       const double* cv, const double* sigma, const double* ell, int k, const double* W, double* V);
int dgemm_shared_a(sfh_ctx* c, int M, int N, int K, const double* A, int lda, const double* B,
                   int ldb, long long strideB, double* C, int ldc, long long strideC, int batch);
int loglik_batch(sfh_lik* L, int M, const double* theta_dev, double* out_dev, int* status_dev,
                 double* logdet_dev = nullptr);
// the log-likelihood of M parameter sets on the device: set b uses
// theta[3 b ..] = (rho, mismatch sigma, mismatch length), mean + b mstride and
// eps_v[b] (eps when eps_v is null); K_s + rho^2 Cu with Cu shared
int loglik_sets(sfh_ctx* c, int ns, int no, int dim, const double* xs, const double* Cu, const double* mean,
                int mstride, const double* Y, double eps, const double* eps_v, int M, const double* theta,
                double* out, int* status, double* logdet, DevBuf& work);
// the sparse-precision GP's sensor-space log-likelihood for M noise levels
// sharing G = g_scale * Graw (ns x ns, device; Y n_o x ns; eps_dev, out,
// status: M); wout (M x ns, nullable) receives B^-1 s
int gpq_sensor_loglik(sfh_ctx* c, int ns, int no, const double* Graw, double g_scale, int M, const double* eps_dev,
                      const double* Y, double* out, int* status, double* wout, DevBuf& work);
int gpq_hmat(sfh_ctx* c, int ns, int no, const double* Graw, double g_scale, double eps, double* h);
int dense_spd_solve(sfh_ctx* c, int n, double* A_dev, int m, double* B_dev, int* status_dev);
int mcmc_run(sfh_lik* L, int C, int steps, const double* theta0, const double* propsd, uint64_t seed,
             int mode, const double* pm, const double* pv, double* samples, double* acc);
int typical_readings(sfh_lik* L, double rho, double sd, double ld, int nr, uint64_t seed, double* Y);
int spmv(sfh_system* s, int k, const double* X, double* Y);
int precision_apply(sfh_system* s, const double* mass, double tau2k4b, int alpha, int k, const double* X, double* Y);
int frac_apply(sfh_system* s, const double* mass, int degree, const double* roots, int transpose, int k,
               const double* X, double* Y);
int gp_scratch_init(int device, double** ptr);
int gp_run_fixed(sfh_ctx* c, int n, int dim, const double* X, const double* y, double sigma, double ell, double eps,
                 int np, const double* Xp, double* logl, double* mean, double* var, int* status);
int gp_run(sfh_ctx* c, int n, int dim, const double* X, const double* y, double sigma, double ell,
           double eps, int np, const double* Xp, double* logl, double* mean, double* var, double* cov);
}  // namespace sfh
