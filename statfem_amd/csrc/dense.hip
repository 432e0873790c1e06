// Dense / statistics kernels:
//  * FP64 MFMA GEMM (v_mfma_f64_16x16x4f64): shared-A batched form for the
//    separable forcing-covariance product and split-K transposed form for the
//    sensor-space Gram product R C_f R^T (statX/infer/CovarianceForcing.ipp:106);
//  * matrix-free dense C_f W for unstructured node sets
//    (statX/infer/CovarianceForcing.ipp:37-100, ExpKernel.hpp:74-92);
//  * one-work-group-per-matrix Cholesky, log-determinant and triangular solves:
//    the batched KOH log-likelihood (statX/infer/LikelihoodKOH.ipp:74-120), the
//    conditioned moments and the dense GP (statX/infer/GaussianProcess.hpp:117-242);
//  * random-walk Metropolis in lockstep across chains (statX/infer/MCMC.hpp:123-158)
//    with Philox4x32-10 streams.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "common.hpp"
#include "mcmc.hpp"

namespace sfh {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int TM = 64, TN = 128, TK = 16;

// C[z] (=|+=) op(A) B[z]. op(A) = A (M x K, lda) or, with TRANSA, X^T for X K x M (lda).
// z selects a batch (B, C strides) and/or a K chunk [z*kstep, z*kstep + kz).
// One-dimensional grid of nwg (a multiple of 8) work-groups over nM x nN x nZ
// tiles, M tiles fastest: consecutive logical tiles share their B panel (and,
// for a K-chunked Gram product, their A panel), and the XCD remap below puts
// consecutive logical tiles on the same XCD, so a panel is fetched from HBM
// into that XCD's L2 once rather than by every XCD.
//
// A 64 x 128 tile per work-group: each of the 4 waves owns 32 x 64 of C as
// 2 x 4 blocks of v_mfma_f64_16x16x4f64. The K loop is double-buffered: the
// next 16-deep panels of A and B are loaded into registers while the current
// ones are multiplied from LDS, then stored to the other LDS buffer (one
// barrier per step). Interior tiles load 16-byte pairs.
typedef double GemmAs[2][TM][TK + 1];
typedef double GemmBs[2][TK][TN + 2];

// One 64 x 128 tile (rows m0.., columns n0..) of C = op(A) B over K range
// [kbeg, kend), by the 256 threads of a work-group, through the LDS panels.
template <bool TRANSA>
__device__ __forceinline__ void dgemm_tile(int M, int N, const double* __restrict__ A, int lda,
                                           const double* __restrict__ Bz, int ldb, double* __restrict__ Cz, int ldc,
                                           int kbeg, int kend, int m0, int n0, GemmAs& As, GemmBs& Bs) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w >> 1, wn = w & 1;
    // 16-byte loads when every row of the tile's panels is in range and aligned
    const bool vec = m0 + TM <= M && n0 + TN <= N && ((lda | ldb) & 1) == 0 &&
                     (((uintptr_t)A | (uintptr_t)Bz) & 15) == 0;
    v4d acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
    // register staging of one K panel: A 64 x 16 (4 per thread), B 16 x 128 (8 per thread)
    double ra[4], rb[8];
    auto load = [&](int k0) {
        if (vec && k0 + TK <= kend) {
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int idx = tid + q * 256;   // 512 pairs
                double2 v;
                if (TRANSA) {   // A^T panel: X rows k, columns m (pairs along m)
                    const int c = idx / (TM / 2), r2 = idx % (TM / 2);
                    v = *reinterpret_cast<const double2*>(A + (size_t)(k0 + c) * lda + m0 + 2 * r2);
                } else {        // rows m, pairs along k
                    const int r = idx / (TK / 2), c2 = idx % (TK / 2);
                    v = *reinterpret_cast<const double2*>(A + (size_t)(m0 + r) * lda + k0 + 2 * c2);
                }
                ra[2 * q] = v.x;
                ra[2 * q + 1] = v.y;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int idx = tid + q * 256;   // 1024 pairs
                const int r = idx / (TN / 2), c2 = idx % (TN / 2);
                const double2 v = *reinterpret_cast<const double2*>(Bz + (size_t)(k0 + r) * ldb + n0 + 2 * c2);
                rb[2 * q] = v.x;
                rb[2 * q + 1] = v.y;
            }
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int idx = tid + q * 256;
                int r, c;
                if (TRANSA) { c = idx / TM; r = idx % TM; }   // coalesced along M
                else { r = idx / TK; c = idx % TK; }
                const int gm = m0 + r, gk = k0 + c;
                double v = 0.0;
                if (gm < M && gk < kend) v = TRANSA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
                ra[q] = v;
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int idx = tid + q * 256;
                const int r = idx / TN, c = idx % TN;
                const int gk = k0 + r, gn = n0 + c;
                rb[q] = (gk < kend && gn < N) ? Bz[(size_t)gk * ldb + gn] : 0.0;
            }
        }
    };
    auto store = [&](int buf, int k0) {
        if (vec && k0 + TK <= kend) {
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int idx = tid + q * 256;
                if (TRANSA) {
                    const int c = idx / (TM / 2), r2 = idx % (TM / 2);
                    As[buf][2 * r2][c] = ra[2 * q];
                    As[buf][2 * r2 + 1][c] = ra[2 * q + 1];
                } else {
                    const int r = idx / (TK / 2), c2 = idx % (TK / 2);
                    As[buf][r][2 * c2] = ra[2 * q];
                    As[buf][r][2 * c2 + 1] = ra[2 * q + 1];
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int idx = tid + q * 256;
                const int r = idx / (TN / 2), c2 = idx % (TN / 2);
                Bs[buf][r][2 * c2] = rb[2 * q];
                Bs[buf][r][2 * c2 + 1] = rb[2 * q + 1];
            }
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int idx = tid + q * 256;
                int r, c;
                if (TRANSA) { c = idx / TM; r = idx % TM; }
                else { r = idx / TK; c = idx % TK; }
                As[buf][r][c] = ra[q];
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int idx = tid + q * 256;
                Bs[buf][idx / TN][idx % TN] = rb[q];
            }
        }
    };
    int buf = 0;
    if (kbeg < kend) {
        load(kbeg);
        store(0, kbeg);
    }
    __syncthreads();
    for (int k0 = kbeg; k0 < kend; k0 += TK) {
        const bool more = k0 + TK < kend;
        if (more) load(k0 + TK);   // in flight while the current panel is multiplied
#pragma unroll
        for (int kk = 0; kk < TK / 4; ++kk) {
            const int kr = kk * 4 + (lane >> 4);
            double a[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = As[buf][wm * 32 + i * 16 + (lane & 15)][kr];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const double b = Bs[buf][kr][wn * 64 + j * 16 + (lane & 15)];
#pragma unroll
                for (int i = 0; i < 2; ++i) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b, acc[i][j], 0, 0, 0);
            }
        }
        if (more) store(buf ^ 1, k0 + TK);
        __syncthreads();
        buf ^= 1;
    }
    // v_mfma_f64_16x16x4f64 result: register r of lane l holds D[l/16 + 4r][l%16]
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm * 32 + i * 16 + (lane >> 4) + 4 * r;
                const int gn = n0 + wn * 64 + j * 16 + (lane & 15);
                if (gm < M && gn < N) Cz[(size_t)gm * ldc + gn] = acc[i][j][r];
            }
}

// XCD-aware logical work-group index: consecutive logical tiles run on one XCD
__device__ inline int xcd_lid() {
    const int nwg = gridDim.x;
    return (nwg % 8) ? (int)blockIdx.x : ((int)blockIdx.x % 8) * (nwg / 8) + (int)blockIdx.x / 8;
}

template <bool TRANSA>
__global__ __launch_bounds__(256) void k_dgemm(int M, int N, int K, const double* __restrict__ A, int lda,
                                               const double* __restrict__ B, int ldb, long long strideB,
                                               double* __restrict__ C, int ldc, long long strideC,
                                               int kz, int kstep, int nM, int nN, int nZ) {
    __shared__ GemmAs As;
    __shared__ GemmBs Bs;
    const int lid = xcd_lid();
    if (lid >= nM * nN * nZ) return;
    const int mt = lid % nM, nt = (lid / nM) % nN, z = lid / (nM * nN);
    const int kbeg = z * kstep;
    dgemm_tile<TRANSA>(M, N, A, lda, B + (size_t)z * strideB, ldb, C + (size_t)z * strideC, ldc, kbeg,
                       min(K, kbeg + kz), mt * TM, nt * TN, As, Bs);
}

// Many independent products in one launch (problems of a batch): group g
// covers logical tiles [tile0, tile0 + nM nN nZ) with the layout of k_dgemm;
// a work-group finds its group by binary search over the tile0 column.
struct GemmGroup {
    const double* A;
    const double* B;
    double* C;
    long long strideB, strideC;
    int M, N, K, lda, ldb, ldc, kz, kstep, nM, nN, nZ, tile0;
};

template <bool TRANSA>
__global__ __launch_bounds__(256) void k_dgemm_grouped(int ngroups, int ntiles, const GemmGroup* __restrict__ gg) {
    __shared__ GemmAs As;
    __shared__ GemmBs Bs;
    const int lid = xcd_lid();
    if (lid >= ntiles) return;
    int lo = 0, hi = ngroups - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (gg[mid].tile0 <= lid) lo = mid;
        else hi = mid - 1;
    }
    const GemmGroup& g = gg[lo];
    const int l = lid - g.tile0;
    const int mt = l % g.nM, nt = (l / g.nM) % g.nN, z = l / (g.nM * g.nN);
    const int kbeg = z * g.kstep;
    dgemm_tile<TRANSA>(g.M, g.N, g.A, g.lda, g.B + (size_t)z * g.strideB, g.ldb, g.C + (size_t)z * g.strideC, g.ldc,
                       kbeg, min(g.K, kbeg + g.kz), mt * TM, nt * TN, As, Bs);
}

// Matrix-vector form of the grouped split-K product for n = 1 (a batch's
// projected means R_q F_q): one work-group per (problem, K chunk); wave w of
// the four walks rows kbeg + w, kbeg + w + 4, ... in order, its 64 lanes over
// the m columns (coalesced 512-byte row segments), and the four row phases
// are added in a fixed order. Slice z of group g is written at g.C + z strideC.
__device__ inline void gemv_t_chunk(int M, const double* __restrict__ A, int lda, const double* __restrict__ B,
                                    int ldb, int kbeg, int kend, double* __restrict__ Cz) {
    __shared__ double red[4][64];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    for (int j0 = 0; j0 < M; j0 += 64) {
        const int j = j0 + l;
        double a0 = 0.0, a1 = 0.0;
        if (j < M) {
            int i = kbeg + w;
            for (; i + 4 < kend; i += 8) {
                a0 += A[(size_t)i * lda + j] * B[(size_t)i * ldb];
                a1 += A[(size_t)(i + 4) * lda + j] * B[(size_t)(i + 4) * ldb];
            }
            if (i < kend) a0 += A[(size_t)i * lda + j] * B[(size_t)i * ldb];
        }
        red[w][l] = a0 + a1;
        __syncthreads();
        if (w == 0 && j < M) Cz[j] = (red[0][l] + red[1][l]) + (red[2][l] + red[3][l]);
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_gemv_t_grouped(int ngroups, int ntiles, const GemmGroup* __restrict__ gg) {
    const int lid = blockIdx.x;
    if (lid >= ntiles) return;
    int lo = 0, hi = ngroups - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (gg[mid].tile0 <= lid) lo = mid;
        else hi = mid - 1;
    }
    const GemmGroup& g = gg[lo];
    const int z = lid - g.tile0;
    const int kbeg = z * g.kstep;
    gemv_t_chunk(g.M, g.A, g.lda, g.B, g.ldb, kbeg, min(g.K, kbeg + g.kz), g.C + (size_t)z * g.strideC);
}

// one product X^T f (K x M by K x 1): chunk z of kz rows per work-group, slice
// z of the partial sums at part + z M
__global__ __launch_bounds__(256) void k_gemv_t(int M, int K, const double* __restrict__ A, int lda,
                                                const double* __restrict__ B, int ldb, int kz,
                                                double* __restrict__ part) {
    const int z = blockIdx.x, kbeg = z * kz;
    gemv_t_chunk(M, A, lda, B, ldb, kbeg, min(K, kbeg + kz), part + (size_t)z * M);
}

// C = sum over nz partial products (deterministic split-K reduction): a
// work-group of 256 threads takes TX outputs, 256 / TX threads per output
// each add the slices z = y, y + 256 / TX, ... in four interleaved chains (so
// four loads are in flight), and the threads' sums are added in a fixed order
template <int TX>
__device__ inline void sum_partials(long long mn, int nz, const double* __restrict__ part, double* __restrict__ C);

template <int TX>
__global__ __launch_bounds__(256) void k_sum_partials(long long mn, int nz, const double* __restrict__ part,
                                                      double* __restrict__ C) {
    sum_partials<TX>(mn, nz, part, C);
}

// the same for the problems of a batch: problem q = blockIdx.y sums its
// nzq[q] slices at part + offq[q] into C + q mn
template <int TX>
__global__ __launch_bounds__(256) void k_sum_partials_grouped(long long mn, const int* __restrict__ nzq,
                                                              const long long* __restrict__ offq,
                                                              const double* __restrict__ part,
                                                              double* __restrict__ C) {
    const int q = blockIdx.y;
    sum_partials<TX>(mn, nzq[q], part + offq[q], C + (size_t)q * mn);
}

template <int TX>
__device__ inline void sum_partials(long long mn, int nz, const double* __restrict__ part, double* __restrict__ C) {
    constexpr int TZ = 256 / TX;
    __shared__ double red[TZ][TX];
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const long long t = (long long)blockIdx.x * TX + tx;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    if (t < mn) {
        int z = ty;
        for (; z + 3 * TZ < nz; z += 4 * TZ) {
#pragma unroll
            for (int j = 0; j < 4; ++j) a[j] += part[(size_t)(z + j * TZ) * mn + t];
        }
        for (int j = 0; z < nz; z += TZ, ++j) a[j] += part[(size_t)z * mn + t];
    }
    red[ty][tx] = (a[0] + a[1]) + (a[2] + a[3]);
    __syncthreads();
    if (ty == 0 && t < mn) {
        double s = 0.0;
#pragma unroll
        for (int y = 0; y < TZ; ++y) s += red[y][tx];
        C[t] = s;
    }
}

// ----------------------------------------------------- forcing covariance
// V = C_f W, matrix-free: 16 rows x 64 columns per block, kernel tiles in LDS
__global__ __launch_bounds__(256) void k_cf_dense(int n, int dim, int k, const double* __restrict__ X,
                                                  const double* __restrict__ cv, double sigma, double ell,
                                                  const double* __restrict__ W, double* __restrict__ V) {
    __shared__ double Kt[16][65];
    __shared__ double Wt[64][65];
    const int tid = threadIdx.x, ty = tid / 16, tx = tid % 16;
    const int i0 = blockIdx.x * 16, c0 = blockIdx.y * 64;
    double acc[4] = {0, 0, 0, 0};
    const double inv2l2 = 1.0 / (2. * ell * ell);
    for (int j0 = 0; j0 < n; j0 += 64) {
        for (int q = 0; q < 4; ++q) {
            const int idx = tid + q * 256;
            const int r = idx / 64, c = idx % 64;
            const int gi = i0 + r, gj = j0 + c;
            double v = 0.0;
            if (gi < n && gj < n) {
                double sq = 0.0;
                for (int d = 0; d < dim; ++d) {
                    const double dd = X[(size_t)gi * dim + d] - X[(size_t)gj * dim + d];
                    sq += dd * dd;
                }
                v = sigma * sigma * exp(-sq * inv2l2);
            }
            Kt[r][c] = v;
        }
        for (int q = 0; q < 16; ++q) {
            const int idx = tid + q * 256;
            const int r = idx / 64, c = idx % 64;
            const int gj = j0 + r, gc = c0 + c;
            Wt[r][c] = (gj < n && gc < k) ? cv[gj] * W[(size_t)gj * k + gc] : 0.0;
        }
        __syncthreads();
        for (int jj = 0; jj < 64; ++jj) {
            const double kv = Kt[ty][jj];
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] += kv * Wt[jj][tx * 4 + q];
        }
        __syncthreads();
    }
    const int gi = i0 + ty;
    if (gi < n)
        for (int q = 0; q < 4; ++q) {
            const int gc = c0 + tx * 4 + q;
            if (gc < k) V[(size_t)gi * k + gc] = cv[gi] * acc[q];
        }
}

// 1D factor of the separable SE kernel on a grid axis: K1[a][b] = exp(-(g_a-g_b)^2 / (2 ell^2))
__global__ void k_axis_kernel(int n, const double* g, double ell, double* K1) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * n) return;
    const int a = t / n, b = t % n;
    const double d = g[a] - g[b];
    K1[t] = exp(-(d * d) / (2. * ell * ell));
}

// out[i][c] = s * w[i] * in[i][c]
__global__ void k_row_scale(long long total, int k, double s, const double* __restrict__ w,
                            const double* __restrict__ in, double* __restrict__ out) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    out[t] = s * w[t / k] * in[t];
}

// the 1D factors of a batch's separable kernels in one launch: matrix j
// (blockIdx.y) has order an[j], its axis points at axes + aoff[j], its values
// at K + koff[j], length scale ell[j]
__global__ void k_axis_kernels(const int* __restrict__ an, const long long* __restrict__ aoff,
                               const long long* __restrict__ koff, const double* __restrict__ ell,
                               const double* __restrict__ axes, double* __restrict__ K) {
    const int j = blockIdx.y, n = an[j];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * n) return;
    const double* g = axes + aoff[j];
    const double d = g[t / n] - g[t % n];
    const double l = ell[j];
    K[koff[j] + t] = exp(-(d * d) / (2. * l * l));
}

// rs[i] = sigma_q^2 cv[i] for the problem q whose rows seg[q]..seg[q+1]-1 hold i
__global__ void k_seg_row_factor(int N, int nseg, const int* __restrict__ seg, const double* __restrict__ sig,
                                 const double* __restrict__ cv, double* __restrict__ rs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (seg[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    const double s = sig[lo];
    rs[i] = (s * s) * cv[i];
}

// ------------------------------------------------------ dense SPD kernels
// Right-looking Cholesky of the lower triangle of a (n x n, row stride lda)
// by one work-group; col is an n-double LDS scratch. Returns false when a
// pivot is not positive (Eigen::LLT failure).
__device__ bool chol_block(int n, double* a, int lda, double* col, int* sflag) {
    const int tid = threadIdx.x, bs = blockDim.x;
    const int lane = tid & 63, wv = tid >> 6, nw = bs >> 6;
    if (tid == 0) *sflag = 0;
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        if (tid == 0) {
            const double d = a[(size_t)j * lda + j];
            if (!(d > 0.0)) {
                *sflag = 1;
            } else {
                const double L = sqrt(d);
                a[(size_t)j * lda + j] = L;
                col[j] = L;
            }
        }
        __syncthreads();
        if (*sflag) return false;
        const double L = col[j];
        for (int i = j + 1 + tid; i < n; i += bs) {
            const double v = a[(size_t)i * lda + j] / L;
            a[(size_t)i * lda + j] = v;
            col[i] = v;
        }
        __syncthreads();
        for (int i = j + 1 + wv; i < n; i += nw) {
            const double cij = col[i];
            double* ai = a + (size_t)i * lda;
            for (int k = j + 1 + lane; k <= i; k += 64) ai[k] -= cij * col[k];
        }
        __syncthreads();
    }
    return true;
}

// forward substitution L Y = D for m columns, D row-major n x m (in place)
__device__ void forward_block(int n, const double* a, int lda, double* D, int m) {
    const int tid = threadIdx.x, bs = blockDim.x;
    for (int j = 0; j < n; ++j) {
        const double L = a[(size_t)j * lda + j];
        for (int t = tid; t < m; t += bs) D[(size_t)j * m + t] /= L;
        __syncthreads();
        const int rows = n - j - 1;
        for (int q = tid; q < rows * m; q += bs) {
            const int i = j + 1 + q / m, t = q % m;
            D[(size_t)i * m + t] -= a[(size_t)i * lda + j] * D[(size_t)j * m + t];
        }
        __syncthreads();
    }
}

// backward substitution L^T X = Y
__device__ void backward_block(int n, const double* a, int lda, double* D, int m) {
    const int tid = threadIdx.x, bs = blockDim.x;
    for (int j = n - 1; j >= 0; --j) {
        const double L = a[(size_t)j * lda + j];
        for (int t = tid; t < m; t += bs) D[(size_t)j * m + t] /= L;
        __syncthreads();
        for (int q = tid; q < j * m; q += bs) {
            const int i = q / m, t = q % m;
            D[(size_t)i * m + t] -= a[(size_t)j * lda + i] * D[(size_t)j * m + t];
        }
        __syncthreads();
    }
}

__device__ double block_sum(double v, double* red) {
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// log-likelihood per hyperparameter set (one work-group per set):
// C = K_d(sigma_d, ell_d) + eps^2 I + rho^2 Cu; LLT; logL = no(-ns/2 log 2pi - sum log L_ii)
// - 1/2 sum_k |L^-1 (y_k - rho mean)|^2   (statX/infer/LikelihoodKOH.ipp:74-120)
__global__ __launch_bounds__(256) void k_loglik(int ns, int no, int dim, const double* __restrict__ xs,
                                                const double* __restrict__ Cu, const double* __restrict__ mean0,
                                                int mstride, const double* __restrict__ Y, double eps0,
                                                const double* __restrict__ eps_v,
                                                const double* __restrict__ theta, double* __restrict__ work,
                                                double* __restrict__ out, int* __restrict__ status,
                                                double* __restrict__ logdet_out) {
    extern __shared__ double sh[];
    double* col = sh;               // ns
    double* D = sh + ns;            // ns * no
    double* red = D + (size_t)ns * no;   // blockDim
    __shared__ int sflag;
    const int b = blockIdx.x, tid = threadIdx.x, bs = blockDim.x;
    const double rho = theta[3 * b], sd = theta[3 * b + 1], ld = theta[3 * b + 2];
    const double* mean = mean0 + (size_t)b * mstride;
    const double eps = eps_v ? eps_v[b] : eps0;
    double* a = work + (size_t)b * ns * ns;
    for (int idx = tid; idx < ns * ns; idx += bs) {
        const int i = idx / ns, k = idx % ns;
        if (k > i) continue;
        double v = se_kernel(dim, xs + (size_t)i * dim, xs + (size_t)k * dim, sd, ld);
        if (i == k) v += eps * eps;
        a[idx] = v + rho * rho * Cu[idx];
    }
    for (int idx = tid; idx < ns * no; idx += bs) {
        const int i = idx / no, k = idx % no;
        D[idx] = Y[(size_t)k * ns + i] - rho * mean[i];
    }
    __syncthreads();
    if (!chol_block(ns, a, ns, col, &sflag)) {
        if (tid == 0) { out[b] = -INFINITY; status[b] = SFH_ENOTSPD; }
        return;
    }
    double lsum = 0.0;
    for (int i = tid; i < ns; i += bs) lsum += log(a[(size_t)i * ns + i]);
    const double logdet = block_sum(lsum, red);
    if (logdet_out && tid == 0) logdet_out[b] = 2.0 * logdet;
    forward_block(ns, a, ns, D, no);
    int bad = 0;
    double tmp3 = 0.0;
    for (int k = 0; k < no; ++k) {
        double q = 0.0;
        for (int i = tid; i < ns; i += bs) q += D[(size_t)i * no + k] * D[(size_t)i * no + k];
        q = block_sum(q, red);
        const double c = -0.5 * q;
        if (!(c < 0.0)) bad = 1;
        tmp3 += c;
    }
    if (tid == 0) {
        const double tmp1 = -0.5 * ns * log(2. * M_PI);
        out[b] = no * (tmp1 - logdet) + tmp3;
        status[b] = bad ? SFH_EQUADFORM : SFH_OK;
    }
}

// Register-resident form of the same log-likelihood for n_s + n_o <= 32 NB.
// One 1024-thread work-group per hyperparameter set; thread (tr, tc) of the
// 32 x 32 grid holds the lower-triangle entries (i, j), i = tr + 32 ii,
// j = tc + 32 jj, jj <= ii, of the augmented matrix
//     [ K_s + rho^2 Cu      .  ]    (n_s x n_s)
//     [ (y_m - rho mean)^T  0  ]    (n_o data rows)
// built in place. A right-looking Cholesky over the first n_s columns leaves
// -|L^-1 d_m|^2 on the data rows' diagonal (the Schur complement), so the
// factor, the log-determinant and all quadratic forms come from one pass with
// no global-memory work matrix. Step k reads the current column k from an LDS
// buffer (every thread forms the pivot's square root itself), updates its
// entries, and the owners of column k+1 publish it into the other buffer: one
// barrier per step.
template <int NB>
__global__ __launch_bounds__(1024) void k_loglik_reg(int ns, int no, int dim, const double* __restrict__ xs,
                                                     const double* __restrict__ Cu, const double* __restrict__ mean0,
                                                     int mstride, const double* __restrict__ Y, double eps0,
                                                     const double* __restrict__ eps_v,
                                                     const double* __restrict__ theta, double* __restrict__ out,
                                                     int* __restrict__ status, double* __restrict__ logdet_out) {
    __shared__ double colb[2][32 * NB];
    __shared__ double piv[32 * NB];
    __shared__ double red[1024 / 64];
    const int b = blockIdx.x;
    const int tr = threadIdx.x >> 5, tc = threadIdx.x & 31;
    const int na = ns + no;
    const double rho = theta[3 * b], sd = theta[3 * b + 1], ld = theta[3 * b + 2];
    const double* mean = mean0 + (size_t)b * mstride;
    const double eps = eps_v ? eps_v[b] : eps0;
    double a[NB][NB];
#pragma unroll
    for (int ii = 0; ii < NB; ++ii)
#pragma unroll
        for (int jj = 0; jj <= ii; ++jj) {
            const int i = tr + 32 * ii, j = tc + 32 * jj;
            double v = 0.0;
            if (i < na && j <= i) {
                if (i < ns) {
                    v = se_kernel(dim, xs + (size_t)i * dim, xs + (size_t)j * dim, sd, ld) + rho * rho * Cu[(size_t)i * ns + j];
                    if (i == j) v += eps * eps;
                } else if (j < ns) {
                    v = Y[(size_t)(i - ns) * ns + j] - rho * mean[j];
                }
            }
            a[ii][jj] = v;
        }
    // publish column 0
    if (tc == 0) {
#pragma unroll
        for (int ii = 0; ii < NB; ++ii) {
            const int i = tr + 32 * ii;
            if (i < na) colb[0][i] = a[ii][0];
        }
    }
    __syncthreads();
    bool bad = false;
    for (int k = 0; k < ns; ++k) {
        const double* cb = colb[k & 1];
        const double d = cb[k];
        if (!(d > 0.0)) { bad = true; break; }   // the same value in every thread
        const double l = sqrt(d);
        const double inv = 1.0 / l;
        if (threadIdx.x == 0) piv[k] = l;
        double cj[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const int j = tc + 32 * q;
            cj[q] = (j > k && j < na) ? cb[j] * inv : 0.0;
        }
#pragma unroll
        for (int ii = 0; ii < NB; ++ii) {
            const int i = tr + 32 * ii;
            const double ci = (i > k && i < na) ? cb[i] * inv : 0.0;
#pragma unroll
            for (int jj = 0; jj <= ii; ++jj) a[ii][jj] -= ci * cj[jj];
        }
        // owners of column k + 1 publish it (entries below and on the diagonal)
        const int kn = k + 1;
        if (kn < ns && tc == (kn & 31)) {
            double* nb = colb[kn & 1];
#pragma unroll
            for (int ii = 0; ii < NB; ++ii)
#pragma unroll
                for (int jj = 0; jj <= ii; ++jj)
                    if (jj == (kn >> 5)) {
                        const int i = tr + 32 * ii;
                        if (i >= kn && i < na) nb[i] = a[ii][jj];
                    }
        }
        __syncthreads();
    }
    if (bad) {
        if (threadIdx.x == 0) {
            out[b] = -INFINITY;
            status[b] = SFH_ENOTSPD;
            if (logdet_out) logdet_out[b] = NAN;
        }
        return;
    }
    // sum of log pivots (fixed order: per-thread strided sums, then per wave)
    double ls = 0.0;
    for (int k = threadIdx.x; k < ns; k += 1024) ls += log(piv[k]);
    for (int off = 32; off > 0; off >>= 1) ls += __shfl_xor(ls, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ls;
    // data rows: diagonal entries hold -|L^-1 d_m|^2
#pragma unroll
    for (int ii = 0; ii < NB; ++ii) {
        const int i = tr + 32 * ii;
        if (tr == tc && i >= ns && i < na) colb[0][i - ns] = a[ii][ii];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double slog = 0.0;
        for (int w = 0; w < 1024 / 64; ++w) slog += red[w];
        double tmp3 = 0.0;
        int badq = 0;
        for (int m = 0; m < no; ++m) {
            const double c = 0.5 * colb[0][m];
            if (!(c < 0.0)) badq = 1;
            tmp3 += c;
        }
        const double tmp1 = -0.5 * ns * log(2. * M_PI);
        out[b] = no * (tmp1 - slog) + tmp3;
        status[b] = badq ? SFH_EQUADFORM : SFH_OK;
        if (logdet_out) logdet_out[b] = 2.0 * slog;
    }
}

// ------------------------------------- MFMA blocked Cholesky log-likelihood
// The log-likelihood of k_loglik_reg by a blocked right-looking Cholesky of
// the same augmented matrix on 16 x 16 tiles, with v_mfma_f64_16x16x4f64 for
// the panel and trailing updates. Rows and columns, in whole tiles:
//   0 .. ns-1        K_s + rho^2 C_u + eps^2 I (the observation covariance)
//   ns .. nsp-1      identity (pads the sensor block to whole tiles: pivots 1,
//                    no coupling, so neither the factor nor log det changes)
//   nsp .. nsp+no-1  data rows (y_m - rho mean)^T, zero in the data block
// One 1024-thread work-group (16 waves) per hyperparameter set. Wave 0
// factors the diagonal tiles; waves 1..15 own the lower triangle's tiles
// (tile t = ib (ib + 1) / 2 + jb to wave 1 + t % 15, slot t / 15), held as
// MFMA accumulators, register r of lane l holding tile entry
// ((l >> 4) + 4 r, l & 15). Pivot step kb of nsp / 16:
//   (1) the owners of tile column kb store its tiles in LDS in the MFMA operand
//       layout Q[kk][l] = X[l & 15][4 kk + (l >> 4)], which serves both the
//       A operand of X and the B operand of X^T;
//   (2) wave 0 factors the diagonal tile in registers (lane r holds row r; each
//       step's pivot and column are read with readlane) and forms L^-1 by rows
//       the same way, stored in the operand layout;
//   (3) the owners turn the tiles below the diagonal into
//       L_ib,kb = A_ib,kb L_kk^-T (four MFMAs each) and store them back;
//   (4) the owners update their tiles right of column kb,
//       A_ib,jb -= L_ib,kb L_jb,kb^T (four MFMAs each, operands from LDS).
// Look-ahead: in step (4) the owners first update and store tile column
// kb + 1, then wave 0 factors its diagonal tile while they finish the rest of
// (4). Three barriers per step, passed by both kinds of wave; consecutive tile
// columns use alternate LDS buffers. The factoring wave holds no tiles, so its
// unrolled registers do not add to the owners'. The data block ends as the Schur
// complement -D L^-T L^-1 D^T: its diagonal holds -|L^-1 d_m|^2, as in the
// register kernel.
constexpr int kLLT = 14;                                // tiles per side at most (224 rows)
constexpr int kLLMfmaMin = 48;                          // sensors from which the blocked kernel is used
constexpr int kLLOwners = 15;                           // tile-owning waves
constexpr int kLLS = (kLLT * (kLLT + 1) / 2 + kLLOwners - 1) / kLLOwners;   // accumulator tiles per owner

__device__ inline double readlane_f64(double x, int src) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u & 0xffffffffull), src);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), src);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// sqrt(d) and 1 / sqrt(d) for d > 0: the hardware reciprocal square root and
// two Newton steps y += y (1 - d y^2) / 2 (each squares the relative error),
// then one correction of the square root, l += y (d - l^2) / 2
__device__ inline void sqrt_rsqrt(double d, double& l, double& y) {
    y = __builtin_amdgcn_rsq(d);
    y = fma(0.5 * y, fma(-d * y, y, 1.0), y);
    y = fma(0.5 * y, fma(-d * y, y, 1.0), y);
    l = d * y;
    l = fma(0.5 * y, fma(-l, l, d), l);
}

// lower-triangle tile t -> (ib, jb), t = ib (ib + 1) / 2 + jb
__device__ inline void tri_tile(int t, int& ib, int& jb) {
    int i = 0;
    while ((i + 1) * (i + 2) / 2 <= t) ++i;
    ib = i;
    jb = t - i * (i + 1) / 2;
}

// accumulator tile (MFMA result layout) -> operand layout in LDS
__device__ inline void store_operand(double* q, const v4d& v, int lane) {
    const int col = lane & 15;
#pragma unroll
    for (int r = 0; r < 4; ++r) q[(col >> 2) * 64 + (col & 3) * 16 + (lane >> 4) + 4 * r] = v[r];
}

// A -= L_i L_j^T for operand tiles Li, Lj in LDS
__device__ inline void tile_update(v4d& acc, const double* __restrict__ Li, const double* __restrict__ Lj, int lane) {
    double x[4], y[4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        x[kk] = Li[kk * 64 + lane];
        y[kk] = Lj[kk * 64 + lane];
    }
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(-x[kk], y[kk], acc, 0, 0, 0);
}

__global__ __launch_bounds__(1024) void k_loglik_mfma(int ns, int no, int dim, const double* __restrict__ xs,
                                                     const double* __restrict__ Cu, const double* __restrict__ mean0,
                                                     int mstride, const double* __restrict__ Y, double eps0,
                                                     const double* __restrict__ eps_v,
                                                     const double* __restrict__ theta, double* __restrict__ out,
                                                     int* __restrict__ status, double* __restrict__ logdet_out) {
    __shared__ double Pb[2][kLLT][256];
    __shared__ double Li[256];
    __shared__ double qd[kLLT * 16];
    __shared__ double pivprod[kLLT];
    __shared__ int sh_bad;
    const int b = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nsp = (ns + 15) & ~15, ts = nsp >> 4;
    const int T = ts + ((no + 15) >> 4);
    if (threadIdx.x == 0) sh_bad = 0;
    if (wave == 0) {
        // ---- the factoring wave: diagonal tile kb while the owners finish the
        // trailing update of step kb - 1
        const int r = lane & 15;
        __syncthreads();   // (A) column 0 is in LDS
        for (int kb = 0; kb < ts; ++kb) {
            const double* D = &Pb[kb & 1][kb][0];
            double a[16], x[16], iv[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                a[c] = D[(c >> 2) * 64 + (c & 3) * 16 + r];
                x[c] = c == r ? 1.0 : 0.0;
            }
            bool bad = false;
            double pr = 1.0;   // product of the pivots (their logs are summed at the end)
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                // factor step c: pivot, column c of L (zero above the diagonal), trailing columns
                const double d = readlane_f64(a[c], c);
                bad = bad || !(d > 0.0);
                double inv, l;
                sqrt_rsqrt(d, l, inv);
                pr *= l;
                iv[c] = inv;
                a[c] = r > c ? a[c] * inv : (r == c ? l : 0.0);
#pragma unroll
                for (int c2 = c + 1; c2 < 16; ++c2) a[c2] -= a[c] * readlane_f64(a[c], c2);
            }
            // rows of X = L^-1: row k is final once divided by L[k][k], then it is
            // eliminated from the rows below (a[k] = L[r][k])
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                if (r == k) {
#pragma unroll
                    for (int j = 0; j <= k; ++j) x[j] *= iv[k];
                }
#pragma unroll
                for (int j = 0; j <= k; ++j) {
                    const double xk = readlane_f64(x[j], k);
                    if (r > k) x[j] -= a[k] * xk;
                }
            }
            if (lane < 16) {   // L^-1 (r, j) in the operand layout
#pragma unroll
                for (int j = 0; j < 16; ++j) Li[(j >> 2) * 64 + (j & 3) * 16 + r] = x[j];
            }
            if (lane == 0) {
                pivprod[kb] = pr;
                if (bad) sh_bad = 1;
            }
            __syncthreads();   // (B) L^-1 is in LDS
            if (sh_bad) {
                if (threadIdx.x == 0) {
                    out[b] = -INFINITY;
                    status[b] = SFH_ENOTSPD;
                    if (logdet_out) logdet_out[b] = NAN;
                }
                return;
            }
            __syncthreads();   // (C) the panel is in LDS
            __syncthreads();   // (A) column kb + 1 is in LDS
        }
        __syncthreads();   // (D) the data diagonal is in LDS
        if (threadIdx.x == 0) {
            double slog = 0.0;
            for (int k = 0; k < ts; ++k) slog += log(pivprod[k]);
            double tmp3 = 0.0;
            int badq = 0;
            for (int m = 0; m < no; ++m) {
                const double c = 0.5 * qd[m];
                if (!(c < 0.0)) badq = 1;
                tmp3 += c;
            }
            const double tmp1 = -0.5 * ns * log(2. * M_PI);
            out[b] = no * (tmp1 - slog) + tmp3;
            status[b] = badq ? SFH_EQUADFORM : SFH_OK;
            if (logdet_out) logdet_out[b] = 2.0 * slog;
        }
        return;
    }
    // ---- the tile-owning waves
    const int ow = wave - 1;
    const int ntile = T * (T + 1) / 2;
    const double rho = theta[3 * b], sd = theta[3 * b + 1], ld = theta[3 * b + 2];
    const double* mean = mean0 + (size_t)b * mstride;
    const double eps = eps_v ? eps_v[b] : eps0;
    int tib[kLLS], tjb[kLLS];
    v4d acc[kLLS];
#pragma unroll
    for (int s = 0; s < kLLS; ++s) {
        const int t = ow + kLLOwners * s;
        int ib = -1, jb = -1;
        if (t < ntile) tri_tile(t, ib, jb);
        tib[s] = ib;
        tjb[s] = jb;
        v4d v = {0.0, 0.0, 0.0, 0.0};
        if (ib >= 0) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 16 * ib + (lane >> 4) + 4 * r, j = 16 * jb + (lane & 15);
                double e = 0.0;
                if (i < ns && j < ns) {
                    e = se_kernel(dim, xs + (size_t)i * dim, xs + (size_t)j * dim, sd, ld) + rho * rho * Cu[(size_t)i * ns + j];
                    if (i == j) e += eps * eps;
                } else if (i < nsp) {
                    e = i == j ? 1.0 : 0.0;
                } else if (j < ns && i - nsp < no) {
                    e = Y[(size_t)(i - nsp) * ns + j] - rho * mean[j];
                }
                v[r] = e;
            }
        }
        acc[s] = v;
    }
    // (1) tile column 0 into LDS
#pragma unroll
    for (int s = 0; s < kLLS; ++s)
        if (tjb[s] == 0) store_operand(&Pb[0][0][0] + 256 * tib[s], acc[s], lane);
    __syncthreads();   // (A)
    for (int kb = 0; kb < ts; ++kb) {
        double* P = &Pb[kb & 1][0][0];
        __syncthreads();   // (B) wave 0 has stored L_kk^-1
        if (sh_bad) return;
        // (3) panel: L_ib,kb = A_ib,kb L_kk^-T
#pragma unroll
        for (int s = 0; s < kLLS; ++s)
            if (tjb[s] == kb && tib[s] > kb) {
                double* X = P + 256 * tib[s];
                double xa[4], lb[4];
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    xa[kk] = X[kk * 64 + lane];
                    lb[kk] = Li[kk * 64 + lane];
                }
                v4d z = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) z = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[kk], lb[kk], z, 0, 0, 0);
                store_operand(X, z, lane);
            }
        __syncthreads();   // (C)
        // (4a) tile column kb + 1 first, stored for the next diagonal step (after
        // the last pivot step: the first data column, stored but unused)
        double* Pn = &Pb[(kb + 1) & 1][0][0];
#pragma unroll
        for (int s = 0; s < kLLS; ++s)
            if (tjb[s] == kb + 1) {
                tile_update(acc[s], P + 256 * tib[s], P + 256 * tjb[s], lane);
                store_operand(Pn + 256 * tib[s], acc[s], lane);
            }
        __syncthreads();   // (A)
        // (4b) the rest of the trailing update, beside the next factorisation
#pragma unroll
        for (int s = 0; s < kLLS; ++s)
            if (tjb[s] > kb + 1) tile_update(acc[s], P + 256 * tib[s], P + 256 * tjb[s], lane);
    }
    // data rows: diagonal of the Schur complement
#pragma unroll
    for (int s = 0; s < kLLS; ++s)
        if (tib[s] >= ts && tib[s] == tjb[s]) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = (lane >> 4) + 4 * r;
                const int m = 16 * tib[s] + row - nsp;
                if (row == (lane & 15) && m < no) qd[m] = acc[s][r];
            }
        }
    __syncthreads();   // (D)
}

// in-place SPD solve A X = B (A n x n lower used, B n x m), one work-group
__global__ __launch_bounds__(256) void k_spd_solve(int n, double* A, int m, double* B, int* status) {
    extern __shared__ double col[];
    __shared__ int sflag;
    if (!chol_block(n, A, n, col, &sflag)) {
        if (threadIdx.x == 0) *status = SFH_ENOTSPD;
        return;
    }
    forward_block(n, A, n, B, m);
    backward_block(n, A, n, B, m);
    if (threadIdx.x == 0) *status = SFH_OK;
}

// log target of hyperParamLearning.hpp:172-190 in log space
__device__ inline double target(double ll, int st, const double* th, int mode, const double* pm, const double* pv) {
    if (st != SFH_OK) return -INFINITY;
    return ll + log_prior(th, mode, pm, pv);
}

__global__ void k_mc_natural(int C, const double* __restrict__ th, double* __restrict__ nat) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 3 * C) nat[t] = exp(th[t]);
}

__global__ void k_mc_propose(int C, int step, uint64_t seed, const double* __restrict__ cur,
                             double p0, double p1, double p2, double* __restrict__ prop,
                             double* __restrict__ nat) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    double z[3], u;
    mcmc_draws(seed, c, step, z, &u);
    const double sd[3] = {p0, p1, p2};
    for (int i = 0; i < 3; ++i) {
        const double v = cur[3 * c + i] + sd[i] * z[i];
        prop[3 * c + i] = v;
        nat[3 * c + i] = exp(v);
    }
}

__global__ void k_mc_init(int C, const double* __restrict__ cur, const double* __restrict__ ll,
                          const int* __restrict__ st, int mode, double m0, double m1, double m2,
                          double v0, double v1, double v2, double* __restrict__ ycur) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const double pm[3] = {m0, m1, m2}, pv[3] = {v0, v1, v2};
    ycur[c] = target(ll[c], st[c], cur + 3 * c, mode, pm, pv);
}

__global__ void k_mc_accept(int C, int step, int steps, uint64_t seed, double* __restrict__ cur,
                            double* __restrict__ ycur, const double* __restrict__ prop,
                            const double* __restrict__ ll, const int* __restrict__ st, int mode, double m0,
                            double m1, double m2, double v0, double v1, double v2, int* __restrict__ nacc,
                            double* __restrict__ samples) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    double z[3], u;
    mcmc_draws(seed, c, step, z, &u);
    const double pm[3] = {m0, m1, m2}, pv[3] = {v0, v1, v2};
    const double yp = target(ll[c], st[c], prop + 3 * c, mode, pm, pv);
    const double ratio = fmin(0.0, yp - ycur[c]);
    if (log(u) <= ratio) {
        for (int i = 0; i < 3; ++i) cur[3 * c + i] = prop[3 * c + i];
        ycur[c] = yp;
        nacc[c] += 1;
    }
    if (samples) {
        double* s = samples + ((size_t)c * steps + step) * 4;
        s[0] = cur[3 * c]; s[1] = cur[3 * c + 1]; s[2] = cur[3 * c + 2]; s[3] = ycur[c];
    }
}

// ---------------------------------------------------------- typical readings
// LikelihoodKOH::giveTypicalReadings (statX/infer/LikelihoodKOH.ipp:42-71) with
// NormalSampler::giveSample (statX/infer/NormalSampler.hpp:46-72):
// y_r = rho mean + L z_r, L L^T = K_s + rho^2 Cu; z from Philox counter
// {reading, index, 3, 0} (Box-Muller, cosine branch). One work-group.
__device__ inline double reading_normal(uint64_t seed, int reading, int idx) {
    uint32_t a[4];
    philox10((uint32_t)reading, (uint32_t)idx, 3u, 0u, (uint32_t)seed, (uint32_t)(seed >> 32), a);
    const double u1 = to_unit(a[0], a[1]), u2 = to_unit(a[2], a[3]);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__global__ __launch_bounds__(256) void k_typical(int ns, int nr, int dim, const double* __restrict__ xs,
                                                 const double* __restrict__ Cu, const double* __restrict__ mean,
                                                 double eps, double rho, double sd, double ld, uint64_t seed,
                                                 double* __restrict__ a, double* __restrict__ Y, int* status) {
    extern __shared__ double col[];
    __shared__ int sflag;
    const int tid = threadIdx.x, bs = blockDim.x;
    for (int idx = tid; idx < ns * ns; idx += bs) {
        const int i = idx / ns, k = idx % ns;
        double v = 0.0;
        if (k <= i) {
            v = se_kernel(dim, xs + (size_t)i * dim, xs + (size_t)k * dim, sd, ld);
            if (i == k) v += eps * eps;
            v += rho * rho * Cu[idx];
        }
        a[idx] = v;
    }
    __syncthreads();
    if (!chol_block(ns, a, ns, col, &sflag)) {
        if (tid == 0) *status = SFH_ENOTSPD;
        return;
    }
    for (int q = tid; q < nr * ns; q += bs) {
        const int r = q / ns, i = q % ns;
        double acc = 0.0;
        for (int j = 0; j <= i; ++j) acc += a[(size_t)i * ns + j] * reading_normal(seed, r, j);
        Y[(size_t)r * ns + i] = rho * mean[i] + acc;
    }
    if (tid == 0) *status = SFH_OK;
}

// ------------------------------------------------------------------- GP
__global__ __launch_bounds__(256) void k_gp_fit(int n, int dim, const double* __restrict__ X,
                                                const double* __restrict__ y, double sigma, double ell,
                                                double eps, double* __restrict__ L, double* __restrict__ alpha,
                                                double* __restrict__ logl, int* __restrict__ status) {
    extern __shared__ double sh[];
    double* col = sh;
    double* red = sh + n;
    __shared__ int sflag;
    const int tid = threadIdx.x, bs = blockDim.x;
    for (int idx = tid; idx < n * n; idx += bs) {
        const int i = idx / n, j = idx % n;
        double v = se_kernel(dim, X + (size_t)i * dim, X + (size_t)j * dim, sigma, ell);
        if (i == j) v += eps * eps;
        L[idx] = v;
    }
    for (int i = tid; i < n; i += bs) alpha[i] = y[i];
    __syncthreads();
    if (!chol_block(n, L, n, col, &sflag)) {
        if (tid == 0) *status = SFH_ENOTSPD;
        return;
    }
    forward_block(n, L, n, alpha, 1);
    double q = 0.0, ls = 0.0;
    for (int i = tid; i < n; i += bs) {
        q += alpha[i] * alpha[i];
        ls += log(L[(size_t)i * n + i]);
    }
    q = block_sum(q, red);
    ls = block_sum(ls, red);
    backward_block(n, L, n, alpha, 1);
    if (tid == 0) {
        *logl = -0.5 * q - ls - 0.5 * n * log(2.0 * M_PI);
        *status = SFH_OK;
    }
}

// mean k*^T alpha, variance k** - |L^-1 k*|^2, one thread per prediction point;
// v = L^-1 k* is kept in work as column p of an n x np row-major block (the
// threads of a wave touch consecutive entries of each row)
__global__ void k_gp_predict(int n, int dim, int np, const double* __restrict__ X, const double* __restrict__ Xp,
                             double sigma, double ell, const double* __restrict__ L,
                             const double* __restrict__ alpha, const int* __restrict__ status,
                             double* __restrict__ work, double* __restrict__ mean, double* __restrict__ var) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np || *status != SFH_OK) return;   // (no factor when the fit failed)
    const double* xp = Xp + (size_t)p * dim;
    double* v = work + p;
    double m = 0.0;
    for (int i = 0; i < n; ++i) {
        const double ks = se_kernel(dim, X + (size_t)i * dim, xp, sigma, ell);
        v[(size_t)i * np] = ks;
        m += ks * alpha[i];
    }
    double q = 0.0;
    for (int i = 0; i < n; ++i) {
        double s = v[(size_t)i * np];
        for (int k = 0; k < i; ++k) s -= L[(size_t)i * n + k] * v[(size_t)k * np];
        s /= L[(size_t)i * n + i];
        v[(size_t)i * np] = s;
        q += s * s;
    }
    mean[p] = m;
    var[p] = se_kernel(dim, xp, xp, sigma, ell) - q;
}

// predictCovariance: cov = k(x*, x*) - V^T V with V = L^-1 K* (C = V^T V given)
__global__ void k_gp_cov(int np, int dim, const double* __restrict__ Xp, double sigma, double ell,
                         const double* __restrict__ C, double* __restrict__ cov) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)np * np) return;
    const int i = (int)(t / np), j = (int)(t % np);
    cov[t] = se_kernel(dim, Xp + (size_t)i * dim, Xp + (size_t)j * dim, sigma, ell) - C[t];
}

// ------------------------------------------------- sparse-precision GP (sensor space)
// GaussianProcessQ::giveLogLikelihood (statX/inferQ/GaussianProcessQ.ipp:37-99)
// for integer beta, written through G = P Q^-1 P^T (ns x ns). With
// B = sigma_e^2 I + n_o G the matrix determinant lemma gives
//   log|sigma_e^2 Q + n_o P^T P| - log|Q| = N log sigma_e^2 + log|B| - ns log sigma_e^2
// and the push-through identity P (sigma_e^2 Q + n_o P^T P)^-1 P^T = G B^-1, so
//   log p(Y) = -(|Y|^2 - s^T G B^-1 s) / (2 sigma_e^2)
//              - (ns (n_o - 1) log sigma_e^2 + log|B|) / 2 - n_o ns log(2 pi) / 2,
// s the sum of the readings. One work-group per set b: G = g_scale * Graw
// (shared by the sets of one precision), sigma_e = eps[b]; Y is n_o x ns.
__global__ __launch_bounds__(256) void k_gpq_loglik(int ns, int no, const double* __restrict__ Graw, double g_scale,
                                                    const double* __restrict__ eps, const double* __restrict__ Y,
                                                    double* __restrict__ work, double* __restrict__ out,
                                                    int* __restrict__ status, double* __restrict__ wout) {
    extern __shared__ double sh[];
    double* col = sh;              // ns
    double* w = sh + ns;           // ns: s, then B^-1 s
    double* red = w + ns;          // blockDim
    __shared__ int sflag;
    const int b = blockIdx.x, tid = threadIdx.x, bs = blockDim.x;
    const double e2 = eps[b] * eps[b];
    double* a = work + (size_t)b * ns * ns;
    for (int idx = tid; idx < ns * ns; idx += bs) {
        const int i = idx / ns, j = idx % ns;
        a[idx] = j <= i ? no * g_scale * Graw[idx] + (i == j ? e2 : 0.0) : 0.0;
    }
    double yy = 0.0;
    for (int i = tid; i < ns; i += bs) {
        double si = 0.0;
        for (int k = 0; k < no; ++k) {
            const double y = Y[(size_t)k * ns + i];
            si += y;
            yy += y * y;
        }
        w[i] = si;
    }
    yy = block_sum(yy, red);
    if (!chol_block(ns, a, ns, col, &sflag)) {
        if (tid == 0) { out[b] = -INFINITY; status[b] = SFH_ENOTSPD; }
        return;
    }
    double ls = 0.0;
    for (int i = tid; i < ns; i += bs) ls += log(a[(size_t)i * ns + i]);
    const double logdet = 2.0 * block_sum(ls, red);
    // s^T G B^-1 s with the unchanged s kept in col
    for (int i = tid; i < ns; i += bs) col[i] = w[i];
    __syncthreads();
    forward_block(ns, a, ns, w, 1);
    backward_block(ns, a, ns, w, 1);
    if (wout)
        for (int i = tid; i < ns; i += bs) wout[(size_t)b * ns + i] = w[i];
    double q = 0.0;
    for (int i = tid; i < ns; i += bs) {
        double gi = 0.0;
        for (int j = 0; j < ns; ++j) gi += Graw[(size_t)i * ns + j] * w[j];
        q += col[i] * g_scale * gi;
    }
    q = block_sum(q, red);
    if (tid == 0) {
        const double tmp1 = -0.5 * (yy - q) / e2;
        const double tmp2 = -0.5 * (ns * (no - 1) * log(e2) + logdet);
        const double tmp3 = -0.5 * no * ns * log(2.0 * M_PI);
        out[b] = tmp1 + tmp2 + tmp3;
        status[b] = SFH_OK;
    }
}

// h = (sigma_e^2 / n_o) I + g_scale * Graw (lower and upper), for the
// posterior covariance's Woodbury solve
__global__ void k_gpq_hmat(int ns, int no, const double* __restrict__ Graw, double g_scale, double eps,
                           double* __restrict__ h) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ns * ns) return;
    h[t] = g_scale * Graw[t] + (t / ns == t % ns ? eps * eps / no : 0.0);
}

inline unsigned nblocks(long long n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// ---------------------------------------------------------------- host
// work-groups for nt tiles, padded to whole XCD groups
inline unsigned dgemm_grid(long long nt) { return (unsigned)((nt + 7) / 8 * 8); }

int dgemm_shared_a(sfh_ctx* c, int M, int N, int K, const double* A, int lda, const double* B, int ldb,
                   long long strideB, double* C, int ldc, long long strideC, int batch) {
    const int nM = (M + TM - 1) / TM, nN = (N + TN - 1) / TN;
    if ((long long)nM * nN * batch > (1LL << 30)) { set_error("product too large"); return SFH_EARG; }
    k_dgemm<false><<<dgemm_grid((long long)nM * nN * batch), 256, 0, c->stream>>>(
        M, N, K, A, lda, B, ldb, strideB, C, ldc, strideC, K, 0, nM, nN, batch);
    SFH_LAUNCH_CHECK();
    return SFH_OK;
}

int gram(sfh_ctx* c, int N, int m, int n, const double* X, int ldx, const double* V, int ldv, double* C,
         DevBuf& work) {
    if (n == 1) {   // a matrix-vector product: one pass over X, about four work-groups per CU
        int kz = std::max(TK, (N + 4 * c->num_cu - 1) / (4 * c->num_cu));
        kz = (kz + TK - 1) / TK * TK;
        const int nz = (N + kz - 1) / kz;
        SFH_TRY(work.ensure((size_t)nz * m * 8));
        k_gemv_t<<<nz, 256, 0, c->stream>>>(m, N, X, ldx, V, ldv, kz, work.as<double>());
        SFH_LAUNCH_CHECK();
        k_sum_partials<16><<<nblocks(m, 16), 256, 0, c->stream>>>(m, nz, work.as<double>(), C);
        SFH_LAUNCH_CHECK();
        return SFH_OK;
    }
    const int tiles = ((m + TM - 1) / TM) * ((n + TN - 1) / TN);
    int nz = std::max(1, std::min(256, (4 * c->num_cu + tiles - 1) / tiles));
    int kz = (N + nz - 1) / nz;
    kz = (kz + TK - 1) / TK * TK;
    nz = (N + kz - 1) / kz;
    const long long mn = (long long)m * n;
    SFH_TRY(work.ensure((size_t)nz * mn * 8));
    const int nM = (m + TM - 1) / TM, nN = (n + TN - 1) / TN;
    k_dgemm<true><<<dgemm_grid((long long)nM * nN * nz), 256, 0, c->stream>>>(
        m, n, N, X, ldx, V, ldv, 0, work.as<double>(), n, mn, kz, kz, nM, nN, nz);
    SFH_LAUNCH_CHECK();
    if (mn >= 64LL * c->num_cu)   // enough outputs for the GPU: 64 per work-group, 4 threads each
        k_sum_partials<64><<<nblocks(mn, 64), 256, 0, c->stream>>>(mn, nz, work.as<double>(), C);
    else
        k_sum_partials<16><<<nblocks(mn, 16), 256, 0, c->stream>>>(mn, nz, work.as<double>(), C);
    SFH_LAUNCH_CHECK();
    return SFH_OK;
}

int covforcing_apply(sfh_system* s, const double* cv, double sigma, double ell, int k, const double* W,
                     double* V, int mode) {
    hipStream_t st = s->ctx->stream;
    if (s->dofpn != 1) { set_error("forcing covariance needs scalar DOFs"); return SFH_EARG; }
    const bool have_grid = s->gn[0] > 0;
    if (mode == 2 && !have_grid) { set_error("separable mode needs sfh_system_set_grid"); return SFH_EARG; }
    const int n = s->ndof;
    if (mode == 1 || (mode == 0 && !have_grid)) {
        dim3 g((n + 15) / 16, (k + 63) / 64);
        k_cf_dense<<<g, 256, 0, st>>>(n, s->dim, k, s->coords.as<double>(), cv, sigma, ell, W, V);
        SFH_LAUNCH_CHECK();
        return SFH_OK;
    }
    const int gn[3] = {s->gn[0], s->dim > 1 ? s->gn[1] : 1, s->dim > 2 ? s->gn[2] : 1};
    const double* ax[3] = {s->gaxis[0].as<double>(), s->gaxis[1].as<double>(), s->gaxis[2].as<double>()};
    return covforcing_grid(s->ctx, s->dim, gn, ax, cv, sigma, ell, k, W, V, s->scratch0, s->scratch1, s->scratch2);
}

// separable: V = sigma^2 D (Kz (x) Ky (x) Kx) D W on a tensor grid, axes x
// fastest; W, V: (nx ny nz) x k row-major
int covforcing_grid(sfh_ctx* c, int d, const int gn_[3], const double* const ax[3], const double* cv, double sigma,
                    double ell, int k, const double* W, double* V, DevBuf& sc0, DevBuf& sc1, DevBuf& sc2) {
    hipStream_t st = c->stream;
    const int nx = gn_[0], ny = d > 1 ? gn_[1] : 1, nz = d > 2 ? gn_[2] : 1;
    const size_t n = (size_t)nx * ny * nz;
    const size_t total = n * k;
    SFH_TRY(sc0.ensure(total * 8));
    SFH_TRY(sc1.ensure(total * 8));
    size_t kb = (size_t)nx * nx + (size_t)ny * ny + (size_t)nz * nz;
    SFH_TRY(sc2.ensure(kb * 8));
    double* Kx = sc2.as<double>();
    double* Ky = Kx + (size_t)nx * nx;
    double* Kz = Ky + (size_t)ny * ny;
    const int gn[3] = {nx, ny, nz};
    double* Ks[3] = {Kx, Ky, Kz};
    for (int a = 0; a < d; ++a) {
        k_axis_kernel<<<nblocks((long long)gn[a] * gn[a], 256), 256, 0, st>>>(gn[a], ax[a], ell, Ks[a]);
        SFH_LAUNCH_CHECK();
    }
    double* t0 = sc0.as<double>();
    double* t1 = sc1.as<double>();
    k_row_scale<<<nblocks((long long)total, 256), 256, 0, st>>>((long long)total, k, 1.0, cv, W, t0);
    SFH_LAUNCH_CHECK();
    // x axis: batch over (iy, iz): Kx (nx x nx) * t0[b] (nx x k)
    {
        int batch = ny * nz;
        for (int b0 = 0; b0 < batch; b0 += 65535) {
            const int bb = std::min(65535, batch - b0);
            SFH_TRY(dgemm_shared_a(c, nx, k, nx, Kx, nx, t0 + (size_t)b0 * nx * k, k, (long long)nx * k,
                                   t1 + (size_t)b0 * nx * k, k, (long long)nx * k, bb));
        }
        std::swap(t0, t1);
    }
    if (d > 1) {   // y axis: batch over iz: Ky (ny x ny) * t0[iz] (ny x nx k)
        const long long st2 = (long long)ny * nx * k;
        SFH_TRY(dgemm_shared_a(c, ny, nx * k, ny, Ky, ny, t0, nx * k, st2, t1, nx * k, st2, nz));
        std::swap(t0, t1);
    }
    if (d > 2) {   // z axis: Kz (nz x nz) * t0 (nz x ny nx k)
        SFH_TRY(dgemm_shared_a(c, nz, ny * nx * k, nz, Kz, nz, t0, ny * nx * k, 0, t1, ny * nx * k, 0, 1));
        std::swap(t0, t1);
    }
    k_row_scale<<<nblocks((long long)total, 256), 256, 0, st>>>((long long)total, k, sigma * sigma, cv, t0, V);
    SFH_LAUNCH_CHECK();
    return SFH_OK;
}

namespace {
// host tables of one batched call, packed into one upload: sections aligned to 16 bytes
struct Packer {
    std::vector<unsigned char> h;
    template <class T>
    size_t add(const T* v, size_t n) {
        const size_t off = (h.size() + 15) / 16 * 16;
        h.resize(off + n * sizeof(T));
        if (n) std::memcpy(h.data() + off, v, n * sizeof(T));
        return off;
    }
};

int grouped_launch(sfh_ctx* c, bool transa, const GemmGroup* dg, int ngroups, long long ntiles) {
    if (ngroups <= 0 || ntiles <= 0) return SFH_OK;
    if (ntiles > (1LL << 30)) { set_error("product too large"); return SFH_EARG; }
    if (transa)
        k_dgemm_grouped<true><<<dgemm_grid(ntiles), 256, 0, c->stream>>>(ngroups, (int)ntiles, dg);
    else
        k_dgemm_grouped<false><<<dgemm_grid(ntiles), 256, 0, c->stream>>>(ngroups, (int)ntiles, dg);
    SFH_LAUNCH_CHECK();
    return SFH_OK;
}
}  // namespace

// C_q = X_q^T V_q for every problem of a batch: one split-K launch over all
// problems' tiles and one fixed-order reduction. The K chunk is common to all
// problems and sized so that the launch holds about four work-groups per CU;
// a problem's products are those of gram() with its own chunking.
int gram_segmented(sfh_ctx* c, int nseg, const int* seg, int m, int n, const double* X, int ldx, const double* V,
                   int ldv, double* C) {
    const int nM = (m + TM - 1) / TM, nN = (n + TN - 1) / TN;
    const long long rows = (long long)seg[nseg] - seg[0];
    const long long mn = (long long)m * n;
    const int tM = n == 1 ? 1 : nM, tN = n == 1 ? 1 : nN;   // tiles per chunk
    long long kz = (long long)tM * tN * std::max(1LL, rows) / std::max(1, 4 * c->num_cu);
    kz = std::max<long long>(TK, (kz + TK - 1) / TK * TK);
    std::vector<GemmGroup> g;
    std::vector<int> nzq(nseg, 0);
    std::vector<long long> offq(nseg, 0);
    long long tiles = 0, slices = 0;
    for (int q = 0; q < nseg; ++q) {
        const int r0 = seg[q], nr = seg[q + 1] - r0;
        offq[q] = slices * mn;
        if (nr <= 0) continue;
        const long long zq = (nr + kz - 1) / kz;
        nzq[q] = (int)zq;
        GemmGroup e;
        e.A = X + (size_t)r0 * ldx;
        e.B = V + (size_t)r0 * ldv;
        e.C = nullptr;   // set once the partial buffer exists
        e.strideB = 0;
        e.strideC = mn;
        e.M = m; e.N = n; e.K = nr; e.lda = ldx; e.ldb = ldv; e.ldc = n;
        e.kz = (int)kz; e.kstep = (int)kz; e.nM = tM; e.nN = tN; e.nZ = (int)zq;
        e.tile0 = (int)tiles;
        g.push_back(e);
        tiles += (long long)tM * tN * zq;
        slices += zq;
    }
    SFH_TRY(c->bat_part.ensure((size_t)std::max(1LL, slices) * mn * 8));
    double* part = c->bat_part.as<double>();
    size_t gi = 0;
    for (int q = 0; q < nseg; ++q)
        if (nzq[q] > 0) g[gi++].C = part + offq[q];
    Packer pk;
    const size_t o_g = pk.add(g.data(), g.size());
    const size_t o_nz = pk.add(nzq.data(), nzq.size());
    const size_t o_off = pk.add(offq.data(), offq.size());
    SFH_TRY(c->bat_ptab.ensure(pk.h.size()));
    unsigned char* dt = c->bat_ptab.as<unsigned char>();
    SFH_CHECK_HIP(hipMemcpyAsync(dt, pk.h.data(), pk.h.size(), hipMemcpyHostToDevice, c->stream));
    if (n == 1) {   // matrix-vector products: one work-group per K chunk
        if (!g.empty()) {
            k_gemv_t_grouped<<<(unsigned)tiles, 256, 0, c->stream>>>((int)g.size(), (int)tiles,
                                                                    reinterpret_cast<const GemmGroup*>(dt + o_g));
            SFH_LAUNCH_CHECK();
        }
    } else {
        SFH_TRY(grouped_launch(c, true, reinterpret_cast<const GemmGroup*>(dt + o_g), (int)g.size(), tiles));
    }
    const int* dnz = reinterpret_cast<const int*>(dt + o_nz);
    const long long* doff = reinterpret_cast<const long long*>(dt + o_off);
    if (mn * nseg >= 64LL * c->num_cu)
        k_sum_partials_grouped<64><<<dim3(nblocks(mn, 64), nseg), 256, 0, c->stream>>>(mn, dnz, doff, part, C);
    else
        k_sum_partials_grouped<16><<<dim3(nblocks(mn, 16), nseg), 256, 0, c->stream>>>(mn, dnz, doff, part, C);
    SFH_LAUNCH_CHECK();
    // the upload reads pageable host memory: keep the tables until it is done
    SFH_CHECK_HIP(hipStreamSynchronize(c->stream));
    return SFH_OK;
}

// V_q = sigma_q^2 D (Ky_q (x) Kx_q) D W_q for every problem of a batch, as
// covforcing_grid does for one grid, in five launches for the whole batch:
// the axis factors; t = D W; V = (I (x) Kx) t (problem q: ny_q products
// Kx_q (nx_q x nx_q) times an nx_q x k block); t = (Ky (x) I) V (one product
// Ky_q times the ny_q x (nx_q k) block); V = sigma^2 D t. A problem with one
// row of nodes (ny_q = 1) has Ky_q = [1], whose product copies the block.
int covforcing_batched(sfh_ctx* c, int nseg, const int* seg, const int* nx, const int* ny, const double* axes,
                       const double* cv, const double* sigma, const double* ell, int k, const double* W, double* V) {
    // rows of W, V and cv are numbered from seg[0]
    const int r00 = seg[0];
    const int N = seg[nseg] - r00;
    W += (size_t)r00 * k;
    V += (size_t)r00 * k;
    cv += r00;
    std::vector<int> segr(nseg + 1);
    for (int q = 0; q <= nseg; ++q) segr[q] = seg[q] - r00;
    std::vector<int> an(2 * nseg);
    std::vector<long long> aoff(2 * nseg), koff(2 * nseg);
    std::vector<double> el(2 * nseg);
    long long na = 0, nk = 0;
    int maxn = 1;
    for (int q = 0; q < nseg; ++q)
        for (int a = 0; a < 2; ++a) {
            const int j = 2 * q + a, n = a ? ny[q] : nx[q];
            an[j] = n; aoff[j] = na; koff[j] = nk; el[j] = ell[q];
            na += n;
            nk += (long long)n * n;
            maxn = std::max(maxn, n);
        }
    // device layout of bat_tab: axis factors, row factors, then the tables
    const size_t kbytes = (size_t)nk * 8, rbytes = (size_t)N * 8;
    Packer pk;
    const size_t o_ax = pk.add(axes, (size_t)na);
    const size_t o_an = pk.add(an.data(), an.size());
    const size_t o_ao = pk.add(aoff.data(), aoff.size());
    const size_t o_ko = pk.add(koff.data(), koff.size());
    const size_t o_el = pk.add(el.data(), el.size());
    const size_t o_sg = pk.add(segr.data(), segr.size());
    const size_t o_si = pk.add(sigma, (size_t)nseg);
    const size_t o_gx = pk.add((const GemmGroup*)nullptr, 0);
    // the group tables need the device addresses of K, so they are appended
    // after the buffer is sized: reserve their room now
    const size_t gbytes = 2 * (size_t)nseg * sizeof(GemmGroup);
    const size_t tab_bytes = o_gx + gbytes;
    const size_t base = (kbytes + rbytes + 15) / 16 * 16;
    SFH_TRY(c->bat_tab.ensure(base + tab_bytes));
    SFH_TRY(c->bat_blk.ensure((size_t)N * k * 8));
    unsigned char* d = c->bat_tab.as<unsigned char>();
    double* Kd = reinterpret_cast<double*>(d);
    double* rs = Kd + nk;
    unsigned char* dt = d + base;
    double* t0 = c->bat_blk.as<double>();
    std::vector<GemmGroup> gx(nseg), gy(nseg);
    long long tx = 0, ty = 0;
    for (int q = 0; q < nseg; ++q) {
        const size_t r0 = (size_t)segr[q];
        const int n1 = nx[q], n2 = ny[q];
        GemmGroup& a = gx[q];
        a.A = Kd + koff[2 * q]; a.B = t0 + r0 * k; a.C = V + r0 * k;
        a.strideB = a.strideC = (long long)n1 * k;
        a.M = n1; a.N = k; a.K = n1; a.lda = n1; a.ldb = k; a.ldc = k;
        a.kz = n1; a.kstep = 0; a.nM = (n1 + TM - 1) / TM; a.nN = (k + TN - 1) / TN; a.nZ = n2;
        a.tile0 = (int)tx;
        tx += (long long)a.nM * a.nN * a.nZ;
        GemmGroup& b = gy[q];
        b.A = Kd + koff[2 * q + 1]; b.B = V + r0 * k; b.C = t0 + r0 * k;
        b.strideB = b.strideC = 0;
        b.M = n2; b.N = n1 * k; b.K = n2; b.lda = n2; b.ldb = n1 * k; b.ldc = n1 * k;
        b.kz = n2; b.kstep = 0; b.nM = (n2 + TM - 1) / TM; b.nN = (n1 * k + TN - 1) / TN; b.nZ = 1;
        b.tile0 = (int)ty;
        ty += (long long)b.nM * b.nN;
    }
    pk.h.resize(tab_bytes);
    std::memcpy(pk.h.data() + o_gx, gx.data(), nseg * sizeof(GemmGroup));
    std::memcpy(pk.h.data() + o_gx + nseg * sizeof(GemmGroup), gy.data(), nseg * sizeof(GemmGroup));
    SFH_CHECK_HIP(hipMemcpyAsync(dt, pk.h.data(), tab_bytes, hipMemcpyHostToDevice, c->stream));
    const double* dax = reinterpret_cast<const double*>(dt + o_ax);
    const int* dan = reinterpret_cast<const int*>(dt + o_an);
    const long long* dao = reinterpret_cast<const long long*>(dt + o_ao);
    const long long* dko = reinterpret_cast<const long long*>(dt + o_ko);
    const double* del = reinterpret_cast<const double*>(dt + o_el);
    const int* dsg = reinterpret_cast<const int*>(dt + o_sg);
    const double* dsi = reinterpret_cast<const double*>(dt + o_si);
    const GemmGroup* dgx = reinterpret_cast<const GemmGroup*>(dt + o_gx);
    const GemmGroup* dgy = dgx + nseg;
    k_axis_kernels<<<dim3(nblocks((long long)maxn * maxn, 256), 2 * nseg), 256, 0, c->stream>>>(dan, dao, dko, del,
                                                                                               dax, Kd);
    SFH_LAUNCH_CHECK();
    const long long total = (long long)N * k;
    k_row_scale<<<nblocks(total, 256), 256, 0, c->stream>>>(total, k, 1.0, cv, W, t0);
    SFH_LAUNCH_CHECK();
    SFH_TRY(grouped_launch(c, false, dgx, nseg, tx));
    SFH_TRY(grouped_launch(c, false, dgy, nseg, ty));
    k_seg_row_factor<<<nblocks(N, 256), 256, 0, c->stream>>>(N, nseg, dsg, dsi, cv, rs);
    SFH_LAUNCH_CHECK();
    k_row_scale<<<nblocks(total, 256), 256, 0, c->stream>>>(total, k, 1.0, rs, t0, V);
    SFH_LAUNCH_CHECK();
    SFH_CHECK_HIP(hipStreamSynchronize(c->stream));
    return SFH_OK;
}

int loglik_sets(sfh_ctx* c, int ns, int no, int dim, const double* xs, const double* Cu, const double* mean,
                int mstride, const double* Y, double eps, const double* eps_v, int M, const double* theta,
                double* out, int* status, double* logdet, DevBuf& work) {
    hipStream_t st = c->stream;
    const int na = ns + no;
    if (M <= 0) return SFH_OK;
    const int tiles = (ns + 15) / 16 + (no + 15) / 16;
    if (ns >= kLLMfmaMin && tiles <= kLLT) {
        // blocked MFMA factorisation (sensor block padded to whole 16 x 16 tiles)
        k_loglik_mfma<<<M, 1024, 0, st>>>(ns, no, dim, xs, Cu, mean, mstride, Y, eps, eps_v, theta, out, status,
                                        logdet);
        SFH_LAUNCH_CHECK();
        return SFH_OK;
    }
    if (na <= 224) {
        // register-resident factorisation (sizes up to 224 augmented rows)
#define SFH_LLREG(NB)                                                                                           \
    k_loglik_reg<NB><<<M, 1024, 0, st>>>(ns, no, dim, xs, Cu, mean, mstride, Y, eps, eps_v, theta, out, status, \
                                         logdet)
        if (na <= 64) SFH_LLREG(2);
        else if (na <= 128) SFH_LLREG(4);
        else SFH_LLREG(7);
#undef SFH_LLREG
        SFH_LAUNCH_CHECK();
        return SFH_OK;
    }
    SFH_TRY(work.ensure((size_t)M * ns * ns * 8));
    const size_t lds = sizeof(double) * ((size_t)ns + (size_t)ns * no + 256);
    if (lds > c->lds_per_block) { set_error("ns * no too large for LDS"); return SFH_EARG; }
    k_loglik<<<M, 256, lds, st>>>(ns, no, dim, xs, Cu, mean, mstride, Y, eps, eps_v, theta, work.as<double>(), out,
                                 status, logdet);
    SFH_LAUNCH_CHECK();
    return SFH_OK;
}

int loglik_batch(sfh_lik* L, int M, const double* theta, double* out, int* status, double* logdet) {
    return loglik_sets(L->ctx, L->ns, L->no, L->dim, L->xs.as<double>(), L->Cu.as<double>(), L->mean.as<double>(), 0,
                       L->Y.as<double>(), L->eps, nullptr, M, theta, out, status, logdet, L->work);
}

int gpq_sensor_loglik(sfh_ctx* c, int ns, int no, const double* Graw, double g_scale, int M, const double* eps_dev,
                      const double* Y, double* out, int* status, double* wout, DevBuf& work) {
    SFH_TRY(work.ensure((size_t)M * ns * ns * 8));
    const size_t lds = sizeof(double) * (2 * (size_t)ns + 256);
    if (lds > c->lds_per_block) { set_error("%d sensors too many for LDS", ns); return SFH_EARG; }
    k_gpq_loglik<<<M, 256, lds, c->stream>>>(ns, no, Graw, g_scale, eps_dev, Y, work.as<double>(), out, status, wout);
    SFH_LAUNCH_CHECK();
    return SFH_OK;
}

int gpq_hmat(sfh_ctx* c, int ns, int no, const double* Graw, double g_scale, double eps, double* h) {
    k_gpq_hmat<<<nblocks((long long)ns * ns, 256), 256, 0, c->stream>>>(ns, no, Graw, g_scale, eps, h);
    SFH_LAUNCH_CHECK();
    return SFH_OK;
}

int dense_spd_solve(sfh_ctx* c, int n, double* A, int m, double* B, int* status) {
    const size_t lds = sizeof(double) * n;
    k_spd_solve<<<1, 256, lds, c->stream>>>(n, A, m, B, status);
    SFH_LAUNCH_CHECK();
    return SFH_OK;
}

int mcmc_run(sfh_lik* L, int C, int steps, const double* theta0, const double* propsd, uint64_t seed, int mode,
             const double* pm, const double* pv, double* samples, double* acc) {
    hipStream_t st = L->ctx->stream;
    SFH_TRY(L->mc_theta.ensure(sizeof(double) * 3 * C));
    SFH_TRY(L->mc_prop.ensure(sizeof(double) * 3 * C));
    SFH_TRY(L->theta.ensure(sizeof(double) * 3 * C));
    SFH_TRY(L->mc_y.ensure(sizeof(double) * C));
    SFH_TRY(L->out.ensure(sizeof(double) * C));
    SFH_TRY(L->status.ensure(sizeof(int) * C));
    SFH_TRY(L->mc_acc.ensure(sizeof(int) * C));
    double* dsamp = nullptr;
    if (samples) {
        SFH_TRY(L->mc_samples.ensure(sizeof(double) * 4 * (size_t)C * steps));
        dsamp = L->mc_samples.as<double>();
    }
    const double m0 = pm ? pm[0] : 0, m1 = pm ? pm[1] : 0, m2 = pm ? pm[2] : 0;
    const double v0 = pv ? pv[0] : 1, v1 = pv ? pv[1] : 1, v2 = pv ? pv[2] : 1;
    SFH_CHECK_HIP(hipMemcpyAsync(L->mc_theta.p, theta0, sizeof(double) * 3 * C, hipMemcpyHostToDevice, st));
    SFH_CHECK_HIP(hipMemsetAsync(L->mc_acc.p, 0, sizeof(int) * C, st));
    const unsigned g = nblocks(C, 64);
    k_mc_natural<<<nblocks(3 * C, 64), 64, 0, st>>>(C, L->mc_theta.as<double>(), L->theta.as<double>());
    SFH_LAUNCH_CHECK();
    SFH_TRY(loglik_batch(L, C, L->theta.as<double>(), L->out.as<double>(), L->status.as<int>()));
    k_mc_init<<<g, 64, 0, st>>>(C, L->mc_theta.as<double>(), L->out.as<double>(), L->status.as<int>(), mode, m0, m1,
                                m2, v0, v1, v2, L->mc_y.as<double>());
    SFH_LAUNCH_CHECK();
    for (int t = 0; t < steps; ++t) {
        k_mc_propose<<<g, 64, 0, st>>>(C, t, seed, L->mc_theta.as<double>(), propsd[0], propsd[1], propsd[2],
                                       L->mc_prop.as<double>(), L->theta.as<double>());
        SFH_TRY(loglik_batch(L, C, L->theta.as<double>(), L->out.as<double>(), L->status.as<int>()));
        k_mc_accept<<<g, 64, 0, st>>>(C, t, steps, seed, L->mc_theta.as<double>(), L->mc_y.as<double>(),
                                      L->mc_prop.as<double>(), L->out.as<double>(), L->status.as<int>(), mode, m0,
                                      m1, m2, v0, v1, v2, L->mc_acc.as<int>(), dsamp);
        if ((t & 255) == 255) SFH_LAUNCH_CHECK();
    }
    SFH_LAUNCH_CHECK();
    std::vector<int> na(C);
    SFH_CHECK_HIP(hipMemcpyAsync(na.data(), L->mc_acc.p, sizeof(int) * C, hipMemcpyDeviceToHost, st));
    if (samples)
        SFH_CHECK_HIP(hipMemcpyAsync(samples, dsamp, sizeof(double) * 4 * (size_t)C * steps,
                                     hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    if (acc)
        for (int c = 0; c < C; ++c) acc[c] = na[c] / (double)steps;
    return SFH_OK;
}

int typical_readings(sfh_lik* L, double rho, double sd, double ld, int nr, uint64_t seed, double* Y) {
    hipStream_t st = L->ctx->stream;
    const int ns = L->ns;
    DevBuf a, y, dst;
    SFH_TRY(a.ensure(sizeof(double) * ns * ns));
    SFH_TRY(y.ensure(sizeof(double) * (size_t)nr * ns));
    SFH_TRY(dst.ensure(sizeof(int)));
    k_typical<<<1, 256, sizeof(double) * ns, st>>>(ns, nr, L->dim, L->xs.as<double>(), L->Cu.as<double>(),
                                                  L->mean.as<double>(), L->eps, rho, sd, ld, seed, a.as<double>(),
                                                  y.as<double>(), dst.as<int>());
    SFH_LAUNCH_CHECK();
    int h = 0;
    SFH_CHECK_HIP(hipMemcpyAsync(&h, dst.p, sizeof(int), hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipMemcpyAsync(Y, y.p, sizeof(double) * (size_t)nr * ns, hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    if (h != SFH_OK) { set_error("total covariance not SPD"); return h; }
    return SFH_OK;
}

int gp_run(sfh_ctx* c, int n, int dim, const double* X, const double* y, double sigma, double ell, double eps,
           int np, const double* Xp, double* logl, double* mean, double* var, double* cov) {
    hipStream_t st = c->stream;
    DevBuf dX, dy, dL, da, dXp, dw, dm, dv, dl, ds;
    SFH_TRY(dX.ensure(sizeof(double) * n * dim));
    SFH_TRY(dy.ensure(sizeof(double) * n));
    SFH_TRY(dL.ensure(sizeof(double) * n * n));
    SFH_TRY(da.ensure(sizeof(double) * n));
    SFH_TRY(dl.ensure(sizeof(double)));
    SFH_TRY(ds.ensure(sizeof(int)));
    SFH_CHECK_HIP(hipMemcpyAsync(dX.p, X, sizeof(double) * n * dim, hipMemcpyHostToDevice, st));
    SFH_CHECK_HIP(hipMemcpyAsync(dy.p, y, sizeof(double) * n, hipMemcpyHostToDevice, st));
    const size_t lds = sizeof(double) * (n + 256);
    if (lds > c->lds_per_block) { set_error("GP too large"); return SFH_EARG; }
    k_gp_fit<<<1, 256, lds, st>>>(n, dim, dX.as<double>(), dy.as<double>(), sigma, ell, eps, dL.as<double>(),
                                  da.as<double>(), dl.as<double>(), ds.as<int>());
    SFH_LAUNCH_CHECK();
    int hst = 0;
    SFH_CHECK_HIP(hipMemcpyAsync(&hst, ds.p, sizeof(int), hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipMemcpyAsync(logl, dl.p, sizeof(double), hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    if (hst != SFH_OK) { set_error("GP covariance not SPD"); return hst; }
    if (np > 0) {
        SFH_TRY(dXp.ensure(sizeof(double) * np * dim));
        SFH_TRY(dw.ensure(sizeof(double) * (size_t)np * n));
        SFH_TRY(dm.ensure(sizeof(double) * np));
        SFH_TRY(dv.ensure(sizeof(double) * np));
        SFH_CHECK_HIP(hipMemcpyAsync(dXp.p, Xp, sizeof(double) * np * dim, hipMemcpyHostToDevice, st));
        k_gp_predict<<<nblocks(np, 64), 64, 0, st>>>(n, dim, np, dX.as<double>(), dXp.as<double>(), sigma, ell,
                                                     dL.as<double>(), da.as<double>(), ds.as<int>(), dw.as<double>(),
                                                     dm.as<double>(), dv.as<double>());
        SFH_LAUNCH_CHECK();
        if (cov) {
            // GaussianProcess::predictCovariance (statX/infer/GaussianProcess.hpp:160-190)
            DevBuf dC, dcov, gw;
            SFH_TRY(dC.ensure(sizeof(double) * (size_t)np * np));
            SFH_TRY(dcov.ensure(sizeof(double) * (size_t)np * np));
            SFH_TRY(gram(c, n, np, np, dw.as<double>(), np, dw.as<double>(), np, dC.as<double>(), gw));
            k_gp_cov<<<nblocks((long long)np * np, 256), 256, 0, st>>>(np, dim, dXp.as<double>(), sigma, ell,
                                                                       dC.as<double>(), dcov.as<double>());
            SFH_LAUNCH_CHECK();
            SFH_CHECK_HIP(hipMemcpyAsync(cov, dcov.p, sizeof(double) * (size_t)np * np, hipMemcpyDeviceToHost, st));
        }
        SFH_CHECK_HIP(hipMemcpyAsync(mean, dm.p, sizeof(double) * np, hipMemcpyDeviceToHost, st));
        SFH_CHECK_HIP(hipMemcpyAsync(var, dv.p, sizeof(double) * np, hipMemcpyDeviceToHost, st));
        SFH_CHECK_HIP(hipStreamSynchronize(st));
    }
    return SFH_OK;
}


// ------------------------------------------- allocation-free GP (fixed scratch)
// The scratch of the GP kernels -- the Cholesky factor L (n x n), the weights
// alpha (n) and the prediction block L^-1 k* (n x np) -- in one file-scope
// device array, so that sfh_gp_noalloc allocates nothing. Each GPU holds its
// own copy; its device address is looked up once per device when a context is
// created there (gp_scratch_init) and only that address is passed to kernels.
// Every call on a GPU reuses the array: the kernels write every entry they
// read within the call, and calls that overlap on one GPU race.
static __device__ __align__(256) double scratch[kGpFixedScratch];

namespace {
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};
}  // namespace

int gp_scratch_init(int device, double** ptr) {
    if (device < 0 || device >= kMaxDevices) { set_error("device %d beyond %d", device, kMaxDevices); return SFH_EARG; }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            return SFH_EDEVICE;
        }
        g_scratch_dev[device] = static_cast<double*>(p);
    }
    *ptr = g_scratch_dev[device];
    return SFH_OK;
}

int gp_run_fixed(sfh_ctx* c, int n, int dim, const double* X, const double* y, double sigma, double ell, double eps,
                 int np, const double* Xp, double* logl, double* mean, double* var, int* status) {
    hipStream_t st = c->stream;
    double* L = c->gp_scratch;
    double* alpha = L + (size_t)kGpFixedN * kGpFixedN;
    double* work = alpha + kGpFixedN;
    const size_t lds = sizeof(double) * (n + 256);
    if (lds > c->lds_per_block) { set_error("GP too large"); return SFH_EARG; }
    k_gp_fit<<<1, 256, lds, st>>>(n, dim, X, y, sigma, ell, eps, L, alpha, logl, status);
    SFH_LAUNCH_CHECK();
    if (np > 0) {
        k_gp_predict<<<nblocks(np, 64), 64, 0, st>>>(n, dim, np, X, Xp, sigma, ell, L, alpha, status, work, mean, var);
        SFH_LAUNCH_CHECK();
    }
    return SFH_OK;
}

}  // namespace sfh
