// Geometric multigrid hierarchy for the preconditioner of a single 2D tensor
// grid. No reference counterpart: the reference factorises the assembled
// operator directly (corlib/SystemSolveEigenSparse.hpp:273-311); on a declared
// grid this library solves iteratively instead, and this hierarchy is what
// makes the iteration count independent of the grid size.
//
// Levels: level 0 is the W x H node grid of the operator (its 9-coefficient
// stencil); a batch of grids gets one hierarchy per grid, levels side by side.
// Level l+1 keeps every other node of level l in each direction and the last
// node of an axis with an even node count (mg_coarse_n: (W+1)/2 nodes for odd
// W, W/2+1 for even W). Prolongation P is bilinear: a fine node that coincides
// with a coarse node copies it, one at an odd interior index averages its two
// neighbouring parents. On level 0, P is restricted to
// the free (unconstrained) nodes, so the coarse spaces see the operator's free
// block. Coarse operators are Galerkin products A_{l+1} = P^T A_l P, formed in
// FP64; for a 9-point stencil and bilinear P they are 9-point again.
//
// Smoother: l1-Jacobi (Baker, Falgout, Kolev, Yang, SIAM J. Sci. Comput. 33,
// 2011), S = diag(omega_i / D_l1,i) with D_l1,i = sum_j |a_ij|. A <= D_l1 for
// any symmetric A, so with omega_i <= 16/9 < 2 the sweep contracts in the
// A-norm on every level whatever the coefficients, and the symmetric V-cycle
// is SPD. omega_i comes from a local Fourier analysis of node i's stencil
// taken as constant: its symbol S(t) = sum_k a_k cos(t . d_k) on a 5 x 5 grid
// of frequencies t in [0, pi]^2 gives lambda_i = max S / D_l1,i (<= 1), the top
// of the smoothed spectrum. omega_i = 4 / (3 lambda_i) centres the
// high-frequency half [lambda_i / 2, lambda_i] on 1; it is used for stencils of
// the 5-point family (lambda_i > 0.9, weights 4/3 .. 1.48), every other row
// takes the Q1 interior's 16/9. (Rows near a boundary and on the coarse levels
// read lambda between 3/4 and 0.9 from the constant-stencil analysis, which
// does not hold there; on them 16/9 measured fewer iterations than
// 4 / (3 lambda_i), and no row's weight exceeds 16/9.) On the Q1
// stencil (lambda = 3/4) that is 16/9, Jacobi damped by 8/9, on the 5-point
// stencil of P1 triangles (lambda = 1) 4/3: the smoothing factor is 1/3 on
// both, where one fixed weight gives 1/2 on one of them and costs a fifth more
// PCG iterations. The coarsest level (at most kMGCoarse nodes) is solved with
// a dense inverse, computed in FP64 on the device.
#include <cmath>

#include "common.hpp"

namespace sfh {
namespace {

constexpr double kOmegaMax = 16.0 / 9.0;
constexpr double kLambda5pt = 0.9;   // smallest lambda_i of the 5-point family

// weight of coarse index X in fine index x (grid of n fine nodes)
__host__ __device__ inline double pweight(int x, int X, int n) { return (double)mg_axis_w(x, X, n); }

// One coarsening step of every grid that has level l + 1, in one launch:
// work-group row blockIdx.y is grid tab[6 q ..] = {W, H, Wc, Hc, fine base,
// coarse base} (bases: first node of the grid in its level's arrays).
// A_c = P^T A P, one coefficient per lane: lane K of a group of nine forms
// coefficient K of coarse node I (seven nodes per 64-lane wave, lane 63 idle)
// as the sum over the fine nodes i of I's support and their stencil
// neighbours j of P[i][I] a_ij P[j][J], J = I + offset(K). cmask (level 0
// only) drops constrained fine nodes from P; a coarse node without free
// support gets a decoupled unit row. Each lane sums in a fixed (i, stencil
// entry) order into one register, with no nine-entry array indexed at run time.
// The nine lanes of a node first stage its support's fine stencils (lane K
// loads the 9 coefficients of support node K, one contiguous row) and the
// constraint flags of the 5 x 5 fine nodes around it in LDS, so the sums read
// LDS instead of issuing dependent global loads behind their branches.
constexpr int kGalNodesPerWave = 7;
__global__ __launch_bounds__(256) void k_mg_galerkin9(const int* __restrict__ tab, const double* __restrict__ A,
                                                      const unsigned char* __restrict__ cmask,
                                                      double* __restrict__ Ac) {
    constexpr int kNodes = 4 * kGalNodesPerWave;   // 256 threads
    __shared__ double sa[kNodes][9][9];
    __shared__ unsigned char sc[kNodes][25];
    const int* t = tab + 6 * blockIdx.y;
    const int W = t[0], H = t[1], Wc = t[2], Hc = t[3], fb = t[4], cb = t[5];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int node = lane / 9, K = lane - 9 * node;
    const int I = (blockIdx.x * (blockDim.x >> 6) + wave) * kGalNodesPerWave + node;
    const bool live = node < kGalNodesPerWave && I < Wc * Hc;
    const int bn = wave * kGalNodesPerWave + node;   // the node's slot in this work-group
    const double* Af = A + (size_t)fb * 9;
    const unsigned char* cm = cmask ? cmask + fb : nullptr;
    const int Ix = live ? I % Wc : 0, Iy = live ? I / Wc : 0;
    if (live) {
        const int ix = 2 * Ix + K % 3 - 1, iy = 2 * Iy + K / 3 - 1;   // support node K
        const bool in = ix >= 0 && ix < W && iy >= 0 && iy < H;
        const size_t i = in ? (size_t)iy * W + ix : 0;
        double row[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) row[k] = in ? Af[i * 9 + k] : 0.0;
#pragma unroll
        for (int k = 0; k < 9; ++k) sa[bn][K][k] = row[k];
        for (int e = K; e < 25; e += 9) {
            const int fx = 2 * Ix - 2 + e % 5, fy = 2 * Iy - 2 + e / 5;
            sc[bn][e] = (cm && fx >= 0 && fx < W && fy >= 0 && fy < H) ? cm[fy * W + fx] : 0;
        }
    }
    __syncthreads();
    double v = 0.0;
    if (live) {
        const int Jx = Ix + K % 3 - 1, Jy = Iy + K / 3 - 1;
        if (Jx >= 0 && Jx < Wc && Jy >= 0 && Jy < Hc) {
            for (int dy = -1; dy <= 1; ++dy) {
                const int iy = 2 * Iy + dy;
                if (iy < 0 || iy >= H) continue;
                const double wy = pweight(iy, Iy, H);
                if (wy == 0.0) continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    const int ix = 2 * Ix + dx;
                    if (ix < 0 || ix >= W) continue;
                    const double wi = wy * pweight(ix, Ix, W);
                    if (wi == 0.0 || sc[bn][(dy + 2) * 5 + (dx + 2)]) continue;
                    const double* ai = sa[bn][(dy + 1) * 3 + (dx + 1)];
                    // neighbours j of i inside J's support (|j - 2J| <= 1 per axis),
                    // in ascending stencil-entry order
                    const int jy0 = max(max(iy - 1, 2 * Jy - 1), 0), jy1 = min(min(iy + 1, 2 * Jy + 1), H - 1);
                    const int jx0 = max(max(ix - 1, 2 * Jx - 1), 0), jx1 = min(min(ix + 1, 2 * Jx + 1), W - 1);
                    for (int jy = jy0; jy <= jy1; ++jy) {
                        const double wyj = pweight(jy, Jy, H);
                        if (wyj == 0.0) continue;
                        for (int jx = jx0; jx <= jx1; ++jx) {
                            const double wxj = pweight(jx, Jx, W);
                            if (wxj == 0.0) continue;
                            const double a = ai[(jy - iy + 1) * 3 + (jx - ix + 1)];
                            if (a == 0.0 || sc[bn][(jy - 2 * Iy + 2) * 5 + (jx - 2 * Ix + 2)]) continue;
                            v += wi * a * wyj * wxj;
                        }
                    }
                }
            }
        }
    }
    // a coarse node without free support gets a decoupled unit row
    const double centre = __shfl(v, node * 9 + 4, 64);
    if (live) Ac[((size_t)cb + I) * 9 + K] = centre > 0.0 ? v : (K == 4 ? 1.0 : 0.0);
}

// smoother weights omega_i / sum_j |a_ij| (omega_i from the stencil's symbol,
// frequencies m pi / 4, m = 0..4 per axis) and (levels >= 1) the FP32 stencil,
// node-major (A32) and coefficient-major (AT32, 9 rows of n)
__global__ void k_mg_level(int n, const double* __restrict__ A, float* __restrict__ dsm, float* __restrict__ A32,
                           float* __restrict__ AT32) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a[9], s = 0.0;
    for (int k = 0; k < 9; ++k) {
        a[k] = A[(size_t)i * 9 + k];
        s += fabs(a[k]);
        if (A32) A32[(size_t)i * 9 + k] = (float)a[k];
        if (AT32) AT32[(size_t)k * n + i] = (float)a[k];
    }
    if (!(s > 0.0)) {
        dsm[i] = 0.f;
        return;
    }
    // cos(m pi / 4) for m = -8..8
    const double h = 0.70710678118654752440;
    const double cs[17] = {1, h, 0, -h, -1, -h, 0, h, 1, h, 0, -h, -1, -h, 0, h, 1};
    double smax = 0.0;
    for (int ty = 0; ty <= 4; ++ty)
        for (int tx = 0; tx <= 4; ++tx) {
            double v = 0.0;
            for (int k = 0; k < 9; ++k) v += a[k] * cs[8 + (k % 3 - 1) * tx + (k / 3 - 1) * ty];
            smax = fmax(smax, v);
        }
    const double lam = smax / s;   // <= 1
    const double om = lam > kLambda5pt ? 4.0 / (3.0 * lam) : kOmegaMax;
    dsm[i] = (float)(om / s);
}

// Dense inverse of each grid's coarsest level, one work-group per grid (grid
// tab[6 r ..] = {W, H, first node in the level's arrays, offset of its n x n
// inverse in inv, offset of its n^2 FP64 work area, grid index q}): the
// symmetrised FP64 matrix of the 9-point stencil is inverted in place by
// Gauss-Jordan elimination without pivoting (the pivots of an SPD matrix are
// its positive Schur-complement diagonals), every pivot step spread over the
// work-group; the matrix lives in LDS (LDS = true, n^2 doubles of dynamic
// shared memory) or in the global work area. The inverse is written in FP32.
// bad[q] = 1 when a pivot is not above kPivotRel times its row's diagonal
// (the operator is not SPD, or singular to rounding: an operator without
// Dirichlet rows), else 0.
template <bool LDS>
__global__ __launch_bounds__(256) void k_mg_coarse_inverse(const int* __restrict__ tab,
                                                           const double* __restrict__ a9,
                                                           double* __restrict__ work, float* __restrict__ inv,
                                                           int* __restrict__ bad) {
    extern __shared__ double smem[];
    __shared__ double col[kMGCoarse];
    const int* t = tab + 6 * blockIdx.x;
    const int W = t[0], n = W * t[1], q = t[5];
    const double* s9 = a9 + (size_t)t[2] * 9;
    float* out = inv + t[3];
    double* A = LDS ? smem : work + t[4];
    const int tid = threadIdx.x, bs = blockDim.x;
    for (int e = tid; e < n * n; e += bs) {
        const int i = e / n, j = e % n;
        const int dx = j % W - i % W, dy = j / W - i / W;
        double v = 0.0;
        if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1) {
            const int k = (dy + 1) * 3 + (dx + 1);
            v = i == j ? s9[(size_t)i * 9 + 4] : 0.5 * (s9[(size_t)i * 9 + k] + s9[(size_t)j * 9 + (8 - k)]);
        }
        A[e] = v;
    }
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        const double p = A[(size_t)k * n + k];
        if (!(p > kPivotRel * s9[(size_t)k * 9 + 4])) {   // the same value in every thread
            if (tid == 0) bad[q] = 1;
            return;
        }
        __syncthreads();   // every thread has read the pivot
        for (int i = tid; i < n; i += bs) col[i] = A[(size_t)i * n + k];
        __syncthreads();
        for (int j = tid; j < n; j += bs) A[(size_t)k * n + j] = j == k ? 1.0 / p : A[(size_t)k * n + j] / p;
        __syncthreads();
        const double pk = 1.0 / p;
        // rows i from tid / 32 in steps of bs / 32, columns j from tid % 32 in steps of 32
        for (int i = tid >> 5; i < n; i += bs >> 5) {
            if (i == k) continue;
            const double ci = col[i];
            double* Ai = A + (size_t)i * n;
            for (int j = tid & 31; j < n; j += 32) Ai[j] = j == k ? -ci * pk : Ai[j] - ci * A[(size_t)k * n + j];
        }
        __syncthreads();
    }
    if (tid == 0) bad[q] = 0;
    for (int e = tid; e < n * n; e += bs) out[e] = (float)A[e];
}

// The same elimination with the matrix in registers, for n <= 32 NR:
// 1024 threads, thread (r, c) = (tid / 32, tid % 32) holds the entries of rows
// r + 32 a and columns c + 32 b (a, b < NR; NR = 3 up to 96 nodes, so that a
// 9 x 9 coarsest grid does not carry the 5 x 5 blocks of 160 nodes). Each step publishes the pivot column and the
// pivot row through LDS (two buffers, so one barrier per step) and every
// entry is updated with the operations of k_mg_coarse_inverse, so the inverse
// is the same. The LDS kernel's step is a chain of dependent LDS loads and
// stores; here the loads of a step are independent.
constexpr int kCinvR = (kMGCoarse + 31) / 32;
template <int NR>
__global__ __launch_bounds__(1024) void k_mg_coarse_inverse_reg(const int* __restrict__ tab,
                                                                const double* __restrict__ a9,
                                                                float* __restrict__ inv, int* __restrict__ bad) {
    __shared__ double colb[2][NR * 32], rowb[2][NR * 32];
    const int* t = tab + 6 * blockIdx.x;
    const int W = t[0], n = W * t[1], q = t[5];
    const double* s9 = a9 + (size_t)t[2] * 9;
    float* out = inv + t[3];
    const int c = threadIdx.x & 31, r = threadIdx.x >> 5;
    double v[NR][NR];
#pragma unroll
    for (int a = 0; a < NR; ++a)
#pragma unroll
        for (int b = 0; b < NR; ++b) {
            const int i = r + 32 * a, j = c + 32 * b;
            double x = 0.0;
            if (i < n && j < n) {
                const int dx = j % W - i % W, dy = j / W - i / W;
                if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1) {
                    const int k = (dy + 1) * 3 + (dx + 1);
                    x = i == j ? s9[(size_t)i * 9 + 4] : 0.5 * (s9[(size_t)i * 9 + k] + s9[(size_t)j * 9 + (8 - k)]);
                }
            }
            v[a][b] = x;
        }
    for (int k = 0; k < n; ++k) {
        double* cb = colb[k & 1];
        double* rb = rowb[k & 1];
        const int kh = k >> 5, kl = k & 31;
        if (c == kl)   // owners of column k
#pragma unroll
            for (int a = 0; a < NR; ++a) {
                double x = v[a][0];
#pragma unroll
                for (int b = 1; b < NR; ++b) x = b == kh ? v[a][b] : x;
                if (r + 32 * a < n) cb[r + 32 * a] = x;
            }
        if (r == kl)   // owners of row k
#pragma unroll
            for (int b = 0; b < NR; ++b) {
                double x = v[0][b];
#pragma unroll
                for (int a = 1; a < NR; ++a) x = a == kh ? v[a][b] : x;
                if (c + 32 * b < n) rb[c + 32 * b] = x;
            }
        __syncthreads();
        const double p = rb[k];
        if (!(p > kPivotRel * s9[(size_t)k * 9 + 4])) {   // the same value in every thread
            if (threadIdx.x == 0) bad[q] = 1;
            return;
        }
        const double pk = 1.0 / p;
        double sj[NR], ci[NR];
#pragma unroll
        for (int b = 0; b < NR; ++b) {
            const int j = c + 32 * b;
            sj[b] = j < n ? rb[j] / p : 0.0;   // the scaled pivot row
        }
#pragma unroll
        for (int a = 0; a < NR; ++a) {
            const int i = r + 32 * a;
            ci[a] = i < n ? cb[i] : 0.0;
        }
#pragma unroll
        for (int a = 0; a < NR; ++a)
#pragma unroll
            for (int b = 0; b < NR; ++b) {
                const int i = r + 32 * a, j = c + 32 * b;
                if (i < n && j < n) {
                    if (i == k)
                        v[a][b] = j == k ? pk : sj[b];
                    else
                        v[a][b] = j == k ? -ci[a] * pk : v[a][b] - ci[a] * sj[b];
                }
            }
    }
    if (threadIdx.x == 0) bad[q] = 0;
#pragma unroll
    for (int a = 0; a < NR; ++a)
#pragma unroll
        for (int b = 0; b < NR; ++b) {
            const int i = r + 32 * a, j = c + 32 * b;
            if (i < n && j < n) out[(size_t)i * n + j] = (float)v[a][b];
        }
}

}  // namespace

int mg_build(sfh_system* s, hipStream_t st, const std::vector<int>& grids) {
    MGHier& m = s->mg;
    const int ng = (int)grids.size() / 3;
    m.L = 0;
    m.nseg = ng;
    m.segL.assign(ng, 0);
    bool big = false;
    for (int q = 0; q < ng; ++q) big = big || (long long)grids[3 * q + 1] * grids[3 * q + 2] > kMGCoarse;
    if (!big || ng == 0) return SFH_OK;   // Jacobi alone for small grids
    // level sizes: coarsen every grid until it has at most kMGCoarse nodes
    // (at least once, so that every grid of the batch has a coarse level)
    std::vector<int> W(grids.size() / 3), H(W.size());
    for (int q = 0; q < ng; ++q) {
        int w = grids[3 * q + 1], h = grids[3 * q + 2], l = 0;
        while (l < kMGMaxLev && (l == 0 || (long long)w * h > kMGCoarse) && (w > 2 || h > 2)) {
            w = mg_coarse_n(w);
            h = mg_coarse_n(h);
            ++l;
        }
        m.segL[q] = l;
        m.L = std::max(m.L, l);
    }
    m.gW.assign((size_t)(m.L + 1) * ng, 0);
    m.gH.assign(m.gW.size(), 0);
    m.gbase.assign(m.gW.size(), 0);
    for (int l = 0; l <= m.L; ++l) {
        long long base = 0;
        for (int q = 0; q < ng; ++q) {
            if (l > m.segL[q]) continue;
            int w = grids[3 * q + 1], h = grids[3 * q + 2];
            for (int u = 0; u < l; ++u) { w = mg_coarse_n(w); h = mg_coarse_n(h); }
            const size_t o = (size_t)l * ng + q;
            m.gW[o] = w;
            m.gH[o] = h;
            m.gbase[o] = l == 0 ? grids[3 * q] : (int)base;
            base += (long long)w * h;
        }
        m.N[l] = l == 0 ? s->ndof : base;
    }
    SFH_TRY(m.dsm[0].ensure(sizeof(float) * (size_t)s->ndof));
    k_mg_level<<<(s->ndof + 255) / 256, 256, 0, st>>>(s->ndof, s->dia64.as<double>(), m.dsm[0].as<float>(), nullptr,
                                                       nullptr);
    SFH_LAUNCH_CHECK();
    // set-up tables for every level, uploaded once: the grids coarsened from
    // level l ({W, H, Wc, Hc, fine base, coarse base}) and the grids whose
    // coarsest level is l ({W, H, base, inverse offset, work offset, grid})
    m.cinv_off.assign(ng, 0);
    m.cmax = 0;
    size_t ctot = 0, wtot = 0;
    std::vector<size_t> woff(ng);
    for (int q = 0; q < ng; ++q) {
        const size_t o = (size_t)m.segL[q] * ng + q;
        const int n = m.gW[o] * m.gH[o];
        m.cinv_off[q] = (int)ctot;
        woff[q] = wtot;
        ctot += (size_t)n * n;
        wtot += (size_t)n * n;
        m.cmax = std::max(m.cmax, n);
    }
    if (ctot > (size_t)INT32_MAX || wtot > (size_t)INT32_MAX) { set_error("coarse levels too large"); return SFH_EARG; }
    std::vector<int> tab;
    std::vector<int> grow(m.L + 1, 0), gofs(m.L + 1, 0), gmax(m.L + 1, 0), crow(m.L + 1, 0), cofs(m.L + 1, 0);
    for (int l = 0; l < m.L; ++l) {
        gofs[l] = (int)tab.size();
        for (int q = 0; q < ng; ++q) {
            if (m.segL[q] <= l) continue;
            const size_t o = (size_t)l * ng + q, oc = (size_t)(l + 1) * ng + q;
            tab.insert(tab.end(), {m.gW[o], m.gH[o], m.gW[oc], m.gH[oc], m.gbase[o], m.gbase[oc]});
            gmax[l] = std::max(gmax[l], m.gW[oc] * m.gH[oc]);
            ++grow[l];
        }
    }
    for (int l = 1; l <= m.L; ++l) {
        cofs[l] = (int)tab.size();
        for (int q = 0; q < ng; ++q) {
            if (m.segL[q] != l) continue;
            const size_t o = (size_t)l * ng + q;
            tab.insert(tab.end(), {m.gW[o], m.gH[o], m.gbase[o], m.cinv_off[q], (int)woff[q], q});
            ++crow[l];
        }
    }
    SFH_TRY(m.gtab.ensure(sizeof(int) * tab.size()));
    SFH_CHECK_HIP(hipMemcpyAsync(m.gtab.p, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice, st));
    for (int l = 0; l < m.L; ++l) {
        const long long nc = m.N[l + 1];
        SFH_TRY(m.coef64[l + 1].ensure(sizeof(double) * 9 * (size_t)nc));
        SFH_TRY(m.coef32[l + 1].ensure(sizeof(float) * 9 * (size_t)nc));
        SFH_TRY(m.coefT32[l + 1].ensure(sizeof(float) * 9 * (size_t)nc));
        SFH_TRY(m.dsm[l + 1].ensure(sizeof(float) * (size_t)nc));
        const double* Af = l == 0 ? s->dia64.as<double>() : m.coef64[l].as<double>();
        for (int r0 = 0; r0 < grow[l]; r0 += 65535) {
            const int rn = std::min(65535, grow[l] - r0);
            constexpr int per_block = kGalNodesPerWave * 4;   // 256 threads
            k_mg_galerkin9<<<dim3((gmax[l] + per_block - 1) / per_block, rn), 256, 0, st>>>(
                m.gtab.as<int>() + gofs[l] + 6 * (size_t)r0, Af, l == 0 ? s->cmask.as<unsigned char>() : nullptr,
                m.coef64[l + 1].as<double>());
            SFH_LAUNCH_CHECK();
        }
        k_mg_level<<<(unsigned)((nc + 255) / 256), 256, 0, st>>>((int)nc, m.coef64[l + 1].as<double>(),
                                                                 m.dsm[l + 1].as<float>(),
                                                                 m.coef32[l + 1].as<float>(),
                                                                 m.coefT32[l + 1].as<float>());
        SFH_LAUNCH_CHECK();
    }
    // dense inverses of the coarsest levels, on the device: one launch per
    // distinct coarsest level, one work-group per grid
    SFH_TRY(m.cinv.ensure(sizeof(float) * ctot));
    SFH_TRY(m.cwork.ensure(sizeof(double) * wtot));
    SFH_TRY(m.cbad.ensure(sizeof(int) * (size_t)ng));
    for (int l = 1; l <= m.L; ++l) {
        if (crow[l] == 0) continue;
        int nmax = 0;
        for (int r = 0; r < crow[l]; ++r) nmax = std::max(nmax, tab[cofs[l] + 6 * r] * tab[cofs[l] + 6 * r + 1]);
        const size_t lds = sizeof(double) * (size_t)nmax * nmax;
        if (nmax <= 3 * 32)
            k_mg_coarse_inverse_reg<3><<<crow[l], 1024, 0, st>>>(m.gtab.as<int>() + cofs[l],
                                                                m.coef64[l].as<double>(), m.cinv.as<float>(),
                                                                m.cbad.as<int>());
        else if (nmax <= kCinvR * 32)
            k_mg_coarse_inverse_reg<kCinvR><<<crow[l], 1024, 0, st>>>(m.gtab.as<int>() + cofs[l],
                                                                     m.coef64[l].as<double>(), m.cinv.as<float>(),
                                                                     m.cbad.as<int>());
        else if (lds + sizeof(double) * kMGCoarse <= s->ctx->lds_per_block)
            k_mg_coarse_inverse<true><<<crow[l], 256, lds, st>>>(m.gtab.as<int>() + cofs[l], m.coef64[l].as<double>(),
                                                                 m.cwork.as<double>(), m.cinv.as<float>(),
                                                                 m.cbad.as<int>());
        else
            k_mg_coarse_inverse<false><<<crow[l], 256, 0, st>>>(m.gtab.as<int>() + cofs[l], m.coef64[l].as<double>(),
                                                                m.cwork.as<double>(), m.cinv.as<float>(),
                                                                m.cbad.as<int>());
        SFH_LAUNCH_CHECK();
    }
    std::vector<int> bad(ng);
    SFH_CHECK_HIP(hipMemcpyAsync(bad.data(), m.cbad.p, sizeof(int) * (size_t)ng, hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    for (int q = 0; q < ng; ++q)
        if (bad[q]) { m.L = 0; return SFH_OK; }   // not SPD: Jacobi only
    return SFH_OK;
}

}  // namespace sfh
