// Multi-right-hand-side Jacobi-PCG with FP64 residual refinement.
//
// Replaces the reference's direct factorisation + per-RHS transposed solves
// (corlib/SystemSolveEigenSparse.hpp:273-404 called in a loop by
// statX/infer/ForwardGaussRHS.hpp:105-112) with one batched iterative solve:
//
//   outer (FP64): r = b - A x, per-RHS ||r||/||b|| test, x += d / scale
//   inner (FP32): Jacobi-PCG on A d = scale * r, every RHS with its own alpha/beta
//
// Layout: blocks are N x kp row-major (kp = k rounded up to 4), so one CSR row
// feeds a contiguous run of right-hand sides; a group of LPR lanes handles one
// row with 4 RHS per lane (16-byte loads in FP32). The matrix is read once per
// RHS tile. Work-groups own contiguous row chunks and are remapped so that each
// of the 8 XCDs walks one contiguous eighth of the rows (its L2 keeps the
// neighbour rows the stencil re-reads).
//
// Dirichlet rows (row-zero form, corlib/SystemSolveEigenSparse.hpp:197-215) are
// eliminated: constrained components of the iterate stay exactly zero, so the
// free block A_ff is what the PCG sees; constrained values and, for transposed
// solves, r_c = (p_c - sum_f A[f][c] r_f) / A[c][c] are restored afterwards.
// All reductions are deterministic (fixed-order two-stage sums).
#include "common.hpp"

namespace sfh {
namespace {

constexpr int kB = 256;
constexpr int kRedB = 256;

template <class T> struct V4 { T x, y, z, w; };

__device__ inline V4<float> ld4(const float* p) {
    const float4 v = *reinterpret_cast<const float4*>(p);
    return {v.x, v.y, v.z, v.w};
}
__device__ inline void st4(float* p, const V4<float>& v) {
    *reinterpret_cast<float4*>(p) = make_float4(v.x, v.y, v.z, v.w);
}
__device__ inline V4<double> ld4(const double* p) {
    const double2 a = *reinterpret_cast<const double2*>(p);
    const double2 b = *reinterpret_cast<const double2*>(p + 2);
    return {a.x, a.y, b.x, b.y};
}
__device__ inline void st4(double* p, const V4<double>& v) {
    *reinterpret_cast<double2*>(p) = make_double2(v.x, v.y);
    *reinterpret_cast<double2*>(p + 2) = make_double2(v.z, v.w);
}

// logical block id: XCD x (hardware round-robin b % 8) gets a contiguous range
__device__ inline int xcd_block(int b, int nb) {
    if (nb % 8) return b;
    return (b % 8) * (nb / 8) + b / 8;
}

struct Tile {
    int lane, grp, ngrp, col0, r0, r1, bid;
};

template <int LPR>
__device__ inline Tile tile(int n, int kp, int rpb) {
    Tile t;
    t.lane = threadIdx.x % LPR;
    t.grp = threadIdx.x / LPR;
    t.ngrp = blockDim.x / LPR;
    t.col0 = (blockIdx.y * LPR + t.lane) * 4;
    t.bid = xcd_block(blockIdx.x, gridDim.x);
    t.r0 = t.bid * rpb;
    t.r1 = min(n, t.r0 + rpb);
    return t;
}

// block-level sum over row groups of per-lane 4-column partials -> part[bid*kp + col]
template <int LPR>
__device__ inline void block_partial(const Tile& t, int kp, const V4<double>& a, double* red,
                                     double* part) {
    const int w = 4 * LPR;
    red[t.grp * w + t.lane * 4 + 0] = a.x;
    red[t.grp * w + t.lane * 4 + 1] = a.y;
    red[t.grp * w + t.lane * 4 + 2] = a.z;
    red[t.grp * w + t.lane * 4 + 3] = a.w;
    __syncthreads();
    for (int c = threadIdx.x; c < w; c += blockDim.x) {
        double s = 0.0;
        for (int g = 0; g < t.ngrp; ++g) s += red[g * w + c];
        const int col = blockIdx.y * w + c;
        if (col < kp) part[(size_t)t.bid * kp + col] = s;
    }
}

// q = A p ; part0 = sum_rows p.q
template <class T, int LPR>
__global__ __launch_bounds__(kB) void k_spmv_dot(int n, int kp, int rpb, const int* __restrict__ rp,
                                                 const int* __restrict__ ci, const T* __restrict__ va,
                                                 const T* __restrict__ p, T* __restrict__ q,
                                                 double* __restrict__ part) {
    __shared__ double red[kB * 4];
    const Tile t = tile<LPR>(n, kp, rpb);
    const bool ok = t.col0 < kp;
    V4<T> acc = {0, 0, 0, 0};
    if (ok) {
        for (int r = t.r0 + t.grp; r < t.r1; r += t.ngrp) {
            V4<T> s = {0, 0, 0, 0};
            const int e1 = rp[r + 1];
            for (int e = rp[r]; e < e1; ++e) {
                const T a = va[e];
                const V4<T> v = ld4(p + (size_t)ci[e] * kp + t.col0);
                s.x += a * v.x; s.y += a * v.y; s.z += a * v.z; s.w += a * v.w;
            }
            st4(q + (size_t)r * kp + t.col0, s);
            const V4<T> pr = ld4(p + (size_t)r * kp + t.col0);
            acc.x += pr.x * s.x; acc.y += pr.y * s.y; acc.z += pr.z * s.z; acc.w += pr.w * s.w;
        }
    }
    const V4<double> ad = {(double)acc.x, (double)acc.y, (double)acc.z, (double)acc.w};
    block_partial<LPR>(t, kp, ad, red, part);
}

// x += alpha p ; r -= alpha q ; z = D^-1 r ; part0 = r.z ; part1 = r.r
template <class T, int LPR>
__global__ __launch_bounds__(kB) void k_update(int n, int kp, int rpb, const double* __restrict__ alpha,
                                               const T* __restrict__ dinv, T* __restrict__ x,
                                               T* __restrict__ r, const T* __restrict__ p,
                                               const T* __restrict__ q, double* __restrict__ part0,
                                               double* __restrict__ part1) {
    __shared__ double red[kB * 4];
    const Tile t = tile<LPR>(n, kp, rpb);
    const bool ok = t.col0 < kp;
    V4<T> rz = {0, 0, 0, 0}, rr = {0, 0, 0, 0};
    if (ok) {
        const T a0 = (T)alpha[t.col0], a1 = (T)alpha[t.col0 + 1], a2 = (T)alpha[t.col0 + 2],
                a3 = (T)alpha[t.col0 + 3];
        for (int row = t.r0 + t.grp; row < t.r1; row += t.ngrp) {
            const size_t o = (size_t)row * kp + t.col0;
            V4<T> xv = ld4(x + o), rv = ld4(r + o);
            const V4<T> pv = ld4(p + o), qv = ld4(q + o);
            xv.x += a0 * pv.x; xv.y += a1 * pv.y; xv.z += a2 * pv.z; xv.w += a3 * pv.w;
            rv.x -= a0 * qv.x; rv.y -= a1 * qv.y; rv.z -= a2 * qv.z; rv.w -= a3 * qv.w;
            st4(x + o, xv);
            st4(r + o, rv);
            const T d = dinv[row];
            rz.x += rv.x * rv.x * d; rz.y += rv.y * rv.y * d; rz.z += rv.z * rv.z * d; rz.w += rv.w * rv.w * d;
            rr.x += rv.x * rv.x; rr.y += rv.y * rv.y; rr.z += rv.z * rv.z; rr.w += rv.w * rv.w;
        }
    }
    V4<double> a = {(double)rz.x, (double)rz.y, (double)rz.z, (double)rz.w};
    block_partial<LPR>(t, kp, a, red, part0);
    __syncthreads();
    V4<double> b = {(double)rr.x, (double)rr.y, (double)rr.z, (double)rr.w};
    block_partial<LPR>(t, kp, b, red, part1);
}

// p = D^-1 r + beta p
template <class T, int LPR>
__global__ __launch_bounds__(kB) void k_pupdate(int n, int kp, int rpb, const double* __restrict__ beta,
                                                const T* __restrict__ dinv, const T* __restrict__ r,
                                                T* __restrict__ p) {
    const Tile t = tile<LPR>(n, kp, rpb);
    if (t.col0 >= kp) return;
    const T b0 = (T)beta[t.col0], b1 = (T)beta[t.col0 + 1], b2 = (T)beta[t.col0 + 2],
            b3 = (T)beta[t.col0 + 3];
    for (int row = t.r0 + t.grp; row < t.r1; row += t.ngrp) {
        const size_t o = (size_t)row * kp + t.col0;
        const T d = dinv[row];
        const V4<T> rv = ld4(r + o);
        V4<T> pv = ld4(p + o);
        pv.x = rv.x * d + b0 * pv.x; pv.y = rv.y * d + b1 * pv.y;
        pv.z = rv.z * d + b2 * pv.z; pv.w = rv.w * d + b3 * pv.w;
        st4(p + o, pv);
    }
}

// inner start: r = scale * r64 ; x = 0 ; p = D^-1 r ; part0 = r.z ; part1 = r.r
template <class T, int LPR>
__global__ __launch_bounds__(kB) void k_init(int n, int kp, int rpb, const double* __restrict__ scale,
                                             const T* __restrict__ dinv, const double* __restrict__ r64,
                                             T* __restrict__ x, T* __restrict__ r, T* __restrict__ p,
                                             double* __restrict__ part0, double* __restrict__ part1) {
    __shared__ double red[kB * 4];
    const Tile t = tile<LPR>(n, kp, rpb);
    const bool ok = t.col0 < kp;
    V4<T> rz = {0, 0, 0, 0}, rr = {0, 0, 0, 0};
    if (ok) {
        const double s0 = scale[t.col0], s1 = scale[t.col0 + 1], s2 = scale[t.col0 + 2],
                     s3 = scale[t.col0 + 3];
        for (int row = t.r0 + t.grp; row < t.r1; row += t.ngrp) {
            const size_t o = (size_t)row * kp + t.col0;
            const V4<double> rd = ld4(r64 + o);
            const V4<T> rv = {(T)(rd.x * s0), (T)(rd.y * s1), (T)(rd.z * s2), (T)(rd.w * s3)};
            const T d = dinv[row];
            const V4<T> z0 = {0, 0, 0, 0};
            st4(x + o, z0);
            st4(r + o, rv);
            const V4<T> pv = {rv.x * d, rv.y * d, rv.z * d, rv.w * d};
            st4(p + o, pv);
            rz.x += rv.x * pv.x; rz.y += rv.y * pv.y; rz.z += rv.z * pv.z; rz.w += rv.w * pv.w;
            rr.x += rv.x * rv.x; rr.y += rv.y * rv.y; rr.z += rv.z * rv.z; rr.w += rv.w * rv.w;
        }
    }
    V4<double> a = {(double)rz.x, (double)rz.y, (double)rz.z, (double)rz.w};
    block_partial<LPR>(t, kp, a, red, part0);
    __syncthreads();
    V4<double> b = {(double)rr.x, (double)rr.y, (double)rr.z, (double)rr.w};
    block_partial<LPR>(t, kp, b, red, part1);
}

// FP64 residual r = b - A x ; part0 = r.r
template <int LPR>
__global__ __launch_bounds__(kB) void k_resid64(int n, int kp, int rpb, const int* __restrict__ rp,
                                                const int* __restrict__ ci, const double* __restrict__ va,
                                                const double* __restrict__ x, const double* __restrict__ b,
                                                double* __restrict__ r, double* __restrict__ part) {
    __shared__ double red[kB * 4];
    const Tile t = tile<LPR>(n, kp, rpb);
    const bool ok = t.col0 < kp;
    V4<double> acc = {0, 0, 0, 0};
    if (ok) {
        for (int row = t.r0 + t.grp; row < t.r1; row += t.ngrp) {
            V4<double> s = ld4(b + (size_t)row * kp + t.col0);
            const int e1 = rp[row + 1];
            for (int e = rp[row]; e < e1; ++e) {
                const double a = va[e];
                const V4<double> v = ld4(x + (size_t)ci[e] * kp + t.col0);
                s.x -= a * v.x; s.y -= a * v.y; s.z -= a * v.z; s.w -= a * v.w;
            }
            st4(r + (size_t)row * kp + t.col0, s);
            acc.x += s.x * s.x; acc.y += s.y * s.y; acc.z += s.z * s.z; acc.w += s.w * s.w;
        }
    }
    block_partial<LPR>(t, kp, acc, red, part);
}

// x64 += x_inner / scale on active columns (scale > 0)
template <class T>
__global__ void k_refine(long long total, int kp, const double* __restrict__ scale,
                         const T* __restrict__ xin, double* __restrict__ x64) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const double s = scale[i % kp];
    if (s > 0.0) x64[i] += (double)xin[i] / s;
}

// ------------------------------------------------------------ reductions
// scalar slots (each kp doubles)
enum { S_RZ = 0, S_ALPHA, S_BETA, S_RR0, S_ACTIVE, S_SCALE, S_RN2, S_BN2, S_TOL2, S_NSLOT };

__device__ inline double sum_parts(int nblk, int kp, const double* part, int col, double* red) {
    const int g = threadIdx.x / 64, l = threadIdx.x % 64;
    double s = 0.0;
    if (col < kp)
        for (int b = g; b < nblk; b += kRedB / 64) s += part[(size_t)b * kp + col];
    red[g * 64 + l] = s;
    __syncthreads();
    double tot = 0.0;
    if (g == 0)
        for (int q = 0; q < kRedB / 64; ++q) tot += red[q * 64 + l];
    __syncthreads();
    return tot;
}

__global__ __launch_bounds__(kRedB) void k_red_alpha(int kp, int nblk, const double* part, double* S) {
    __shared__ double red[kRedB];
    const int col = blockIdx.x * 64 + threadIdx.x % 64;
    const double pq = sum_parts(nblk, kp, part, col, red);
    if (threadIdx.x < 64 && col < kp) {
        const double rz = S[S_RZ * kp + col];
        S[S_ALPHA * kp + col] = (S[S_ACTIVE * kp + col] > 0.0 && pq > 0.0) ? rz / pq : 0.0;
    }
}

__global__ __launch_bounds__(kRedB) void k_red_beta(int kp, int nblk, const double* part0,
                                                    const double* part1, double* S, int* nact) {
    __shared__ double red[kRedB];
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    const int col = blockIdx.x * 64 + threadIdx.x % 64;
    const double rzn = sum_parts(nblk, kp, part0, col, red);
    const double rrn = sum_parts(nblk, kp, part1, col, red);
    if (threadIdx.x < 64 && col < kp) {
        double* active = S + S_ACTIVE * kp;
        if (active[col] > 0.0) {
            const double rz = S[S_RZ * kp + col];
            S[S_BETA * kp + col] = rz > 0.0 ? rzn / rz : 0.0;
            S[S_RZ * kp + col] = rzn;
            if (!(rrn > S[S_TOL2 * kp + col] * S[S_RR0 * kp + col])) active[col] = 0.0;
            else atomicAdd(&cnt, 1);
        } else {
            S[S_BETA * kp + col] = 0.0;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) nact[blockIdx.x] = cnt;
}

__global__ __launch_bounds__(kRedB) void k_red_init(int kp, int nblk, const double* part0,
                                                    const double* part1, double* S) {
    __shared__ double red[kRedB];
    const int col = blockIdx.x * 64 + threadIdx.x % 64;
    const double rz = sum_parts(nblk, kp, part0, col, red);
    const double rr = sum_parts(nblk, kp, part1, col, red);
    if (threadIdx.x < 64 && col < kp) {
        S[S_RZ * kp + col] = rz;
        S[S_RR0 * kp + col] = rr;
        S[S_BETA * kp + col] = 0.0;
        S[S_ALPHA * kp + col] = 0.0;
        S[S_ACTIVE * kp + col] = (S[S_SCALE * kp + col] > 0.0 && rr > 0.0) ? 1.0 : 0.0;
    }
}

__global__ __launch_bounds__(kRedB) void k_red_norm(int kp, int nblk, const double* part, double* out) {
    __shared__ double red[kRedB];
    const int col = blockIdx.x * 64 + threadIdx.x % 64;
    const double s = sum_parts(nblk, kp, part, col, red);
    if (threadIdx.x < 64 && col < kp) out[col] = s;
}

// ---------------------------------------------------- constraint handling
__global__ void k_dinv(int n, const int* diag, const double* va, float* d32, double* d64) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = va[diag[i]];
    d64[i] = 1.0 / d;
    d32[i] = (float)(1.0 / d);
}

// b64 (N x kp): free rows carry B (non-transposed: minus the constrained-column
// couplings A[f][c] x_c with x_c = B_c / A[c][c]); constrained rows are zero.
__global__ void k_prepare(int n, int k, int kp, int trans, const double* __restrict__ B, int ldb,
                          const int* __restrict__ rp, const int* __restrict__ ci,
                          const double* __restrict__ va, const int* __restrict__ diag,
                          const unsigned char* __restrict__ cmask, double* __restrict__ b64) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * kp) return;
    const int i = (int)(t / kp), c = (int)(t % kp);
    double v = 0.0;
    if (c < k && !cmask[i]) {
        v = B[(size_t)i * ldb + c];
        if (!trans) {
            for (int e = rp[i]; e < rp[i + 1]; ++e) {
                const int j = ci[e];
                if (cmask[j]) v -= va[e] * (B[(size_t)j * ldb + c] / va[diag[j]]);
            }
        }
    }
    b64[t] = v;
}

// X from the eliminated solution; see the file comment for the constrained rows.
__global__ void k_finalize(int n, int k, int kp, int trans, const double* __restrict__ B, int ldb,
                           const double* __restrict__ x64, const int* __restrict__ rp,
                           const int* __restrict__ ci, const double* __restrict__ va,
                           const int* __restrict__ diag, const unsigned char* __restrict__ cmask,
                           double* __restrict__ X, int ldx) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * k) return;
    const int i = (int)(t / k), c = (int)(t % k);
    double v;
    if (!cmask[i]) {
        v = x64[(size_t)i * kp + c];
    } else if (!trans) {
        v = B[(size_t)i * ldb + c] / va[diag[i]];
    } else {
        double s = B[(size_t)i * ldb + c];
        for (int e = rp[i]; e < rp[i + 1]; ++e) {
            const int j = ci[e];
            if (cmask[j]) continue;
            int lo = rp[j], hi = rp[j + 1] - 1;   // A[j][i]
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (ci[mid] < i) lo = mid + 1; else hi = mid;
            }
            if (ci[lo] == i) s -= va[lo] * x64[(size_t)j * kp + c];
        }
        v = s / va[diag[i]];
    }
    X[(size_t)i * ldx + c] = v;
}

inline int pow2_at_least(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

struct Plan {
    int n, kp, lpr, ntile, rpb, nblk, nred;
};

Plan make_plan(int n, int k, int num_cu) {
    Plan p;
    p.n = n;
    p.kp = (k + 3) / 4 * 4;
    p.lpr = std::min(64, pow2_at_least(p.kp / 4));
    p.ntile = (p.kp + 4 * p.lpr - 1) / (4 * p.lpr);
    const int rows_per_pass = kB / p.lpr;
    int target = std::max(8, 4 * num_cu / p.ntile);
    p.rpb = std::max(rows_per_pass, (n + target - 1) / target);
    p.nblk = (n + p.rpb - 1) / p.rpb;
    p.nblk = (p.nblk + 7) / 8 * 8;
    p.nred = (p.kp + 63) / 64;
    return p;
}

template <class T, int LPR>
struct Kern {
    static void spmv(const Plan& P, hipStream_t st, const int* rp, const int* ci, const T* va, const T* p,
                     T* q, double* part) {
        k_spmv_dot<T, LPR><<<dim3(P.nblk, P.ntile), kB, 0, st>>>(P.n, P.kp, P.rpb, rp, ci, va, p, q, part);
    }
    static void update(const Plan& P, hipStream_t st, const double* alpha, const T* dinv, T* x, T* r,
                       const T* p, const T* q, double* p0, double* p1) {
        k_update<T, LPR><<<dim3(P.nblk, P.ntile), kB, 0, st>>>(P.n, P.kp, P.rpb, alpha, dinv, x, r, p, q, p0, p1);
    }
    static void pupdate(const Plan& P, hipStream_t st, const double* beta, const T* dinv, const T* r, T* p) {
        k_pupdate<T, LPR><<<dim3(P.nblk, P.ntile), kB, 0, st>>>(P.n, P.kp, P.rpb, beta, dinv, r, p);
    }
    static void init(const Plan& P, hipStream_t st, const double* scale, const T* dinv, const double* r64,
                     T* x, T* r, T* p, double* p0, double* p1) {
        k_init<T, LPR><<<dim3(P.nblk, P.ntile), kB, 0, st>>>(P.n, P.kp, P.rpb, scale, dinv, r64, x, r, p, p0, p1);
    }
    static void resid(const Plan& P, hipStream_t st, const int* rp, const int* ci, const double* va,
                      const double* x, const double* b, double* r, double* part) {
        k_resid64<LPR><<<dim3(P.nblk, P.ntile), kB, 0, st>>>(P.n, P.kp, P.rpb, rp, ci, va, x, b, r, part);
    }
};

template <class T>
int run_solve(sfh_system* s, const Plan& P, int trans, int k, const double* B, int ldb, double* X,
              int ldx, const sfh_solve_opts& o, sfh_solve_stats* stats);

template <class T, int LPR>
int inner_loop(sfh_system* s, const Plan& P, const sfh_solve_opts& o, const T* va, const T* dinv,
               int& iters) {
    hipStream_t st = s->ctx->stream;
    SolveWork& w = s->work;
    double* S = w.scal.as<double>();
    const int kp = P.kp;
    T* x = w.x32.as<T>();
    T* r = w.r32.as<T>();
    T* p = w.p32.as<T>();
    T* q = w.q32.as<T>();
    double* p0 = w.part0.as<double>();
    double* p1 = w.part1.as<double>();
    const int* rp = s->rowptr.as<int>();
    const int* ci = s->colind.as<int>();
    Kern<T, LPR>::init(P, st, S + S_SCALE * kp, dinv, w.r64.as<double>(), x, r, p, p0, p1);
    k_red_init<<<P.nred, kRedB, 0, st>>>(kp, P.nblk, p0, p1, S);
    SFH_LAUNCH_CHECK();
    std::vector<int> nact(P.nred);
    const int check = std::max(1, o.check_every);
    for (int it = 0; it < o.max_inner; ++it) {
        Kern<T, LPR>::spmv(P, st, rp, ci, va, p, q, p0);
        k_red_alpha<<<P.nred, kRedB, 0, st>>>(kp, P.nblk, p0, S);
        Kern<T, LPR>::update(P, st, S + S_ALPHA * kp, dinv, x, r, p, q, p0, p1);
        k_red_beta<<<P.nred, kRedB, 0, st>>>(kp, P.nblk, p0, p1, S, w.nact.as<int>());
        Kern<T, LPR>::pupdate(P, st, S + S_BETA * kp, dinv, r, p);
        ++iters;
        if ((it + 1) % check == 0 || it + 1 == o.max_inner) {
            SFH_LAUNCH_CHECK();
            SFH_CHECK_HIP(hipMemcpyAsync(nact.data(), w.nact.p, sizeof(int) * P.nred,
                                         hipMemcpyDeviceToHost, st));
            SFH_CHECK_HIP(hipStreamSynchronize(st));
            int tot = 0;
            for (int v : nact) tot += v;
            if (tot == 0) break;
        }
    }
    SFH_LAUNCH_CHECK();
    return SFH_OK;
}

template <class T>
int dispatch_inner(sfh_system* s, const Plan& P, const sfh_solve_opts& o, const T* va, const T* dinv,
                   int& iters) {
    switch (P.lpr) {
        case 1: return inner_loop<T, 1>(s, P, o, va, dinv, iters);
        case 2: return inner_loop<T, 2>(s, P, o, va, dinv, iters);
        case 4: return inner_loop<T, 4>(s, P, o, va, dinv, iters);
        case 8: return inner_loop<T, 8>(s, P, o, va, dinv, iters);
        case 16: return inner_loop<T, 16>(s, P, o, va, dinv, iters);
        case 32: return inner_loop<T, 32>(s, P, o, va, dinv, iters);
        default: return inner_loop<T, 64>(s, P, o, va, dinv, iters);
    }
}

void launch_resid(const Plan& P, hipStream_t st, const int* rp, const int* ci, const double* va,
                  const double* x, const double* b, double* r, double* part) {
    switch (P.lpr) {
        case 1: Kern<double, 1>::resid(P, st, rp, ci, va, x, b, r, part); break;
        case 2: Kern<double, 2>::resid(P, st, rp, ci, va, x, b, r, part); break;
        case 4: Kern<double, 4>::resid(P, st, rp, ci, va, x, b, r, part); break;
        case 8: Kern<double, 8>::resid(P, st, rp, ci, va, x, b, r, part); break;
        case 16: Kern<double, 16>::resid(P, st, rp, ci, va, x, b, r, part); break;
        case 32: Kern<double, 32>::resid(P, st, rp, ci, va, x, b, r, part); break;
        default: Kern<double, 64>::resid(P, st, rp, ci, va, x, b, r, part); break;
    }
}

template <class T>
int run_solve(sfh_system* s, const Plan& P, int trans, int k, const double* B, int ldb, double* X,
              int ldx, const sfh_solve_opts& o, sfh_solve_stats* stats) {
    hipStream_t st = s->ctx->stream;
    SolveWork& w = s->work;
    const int n = P.n, kp = P.kp;
    const size_t nk = (size_t)n * kp;
    SFH_TRY(w.b64.ensure(nk * 8));
    SFH_TRY(w.x64.ensure(nk * 8));
    SFH_TRY(w.r64.ensure(nk * 8));
    SFH_TRY(w.x32.ensure(nk * sizeof(T)));
    SFH_TRY(w.r32.ensure(nk * sizeof(T)));
    SFH_TRY(w.p32.ensure(nk * sizeof(T)));
    SFH_TRY(w.q32.ensure(nk * sizeof(T)));
    SFH_TRY(w.part0.ensure((size_t)P.nblk * kp * 8));
    SFH_TRY(w.part1.ensure((size_t)P.nblk * kp * 8));
    SFH_TRY(w.scal.ensure((size_t)S_NSLOT * kp * 8));
    SFH_TRY(w.nact.ensure(sizeof(int) * P.nred));
    double* S = w.scal.as<double>();
    const int* rp = s->rowptr.as<int>();
    const int* ci = s->colind.as<int>();
    const double* va = s->vals.as<double>();

    SFH_CHECK_HIP(hipEventRecord(s->ctx->ev0, st));
    if (s->dirty) {
        k_dinv<<<(n + 255) / 256, 256, 0, st>>>(n, s->diagpos.as<int>(), va, s->dinv32.as<float>(),
                                                s->dinv64.as<double>());
        SFH_LAUNCH_CHECK();
        s->dirty = false;
    }
    k_prepare<<<(unsigned)((nk + 255) / 256), 256, 0, st>>>(n, k, kp, trans, B, ldb, rp, ci, va,
                                                          s->diagpos.as<int>(), s->cmask.as<unsigned char>(),
                                                          w.b64.as<double>());
    SFH_LAUNCH_CHECK();
    SFH_CHECK_HIP(hipMemsetAsync(w.x64.p, 0, nk * 8, st));
    launch_resid(P, st, rp, ci, va, w.x64.as<double>(), w.b64.as<double>(), w.r64.as<double>(),
                 w.part0.as<double>());
    k_red_norm<<<P.nred, kRedB, 0, st>>>(kp, P.nblk, w.part0.as<double>(), S + S_BN2 * kp);
    SFH_LAUNCH_CHECK();
    std::vector<double> bn2(kp), rn2(kp), scale(kp), tol2(kp, o.inner_tol * o.inner_tol);
    SFH_CHECK_HIP(hipMemcpyAsync(bn2.data(), S + S_BN2 * kp, kp * 8, hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipMemcpyAsync(S + S_TOL2 * kp, tol2.data(), kp * 8, hipMemcpyHostToDevice, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    rn2 = bn2;
    const T* vaT = std::is_same<T, float>::value ? (const T*)s->vals32.p : (const T*)s->vals.p;
    const T* dinvT = std::is_same<T, float>::value ? (const T*)s->dinv32.p : (const T*)s->dinv64.p;
    int outer = 0, iters = 0;
    double maxrel = 0.0;
    bool done = false;
    for (;;) {
        int nact = 0;
        maxrel = 0.0;
        for (int c = 0; c < kp; ++c) {
            const double rel = bn2[c] > 0.0 ? std::sqrt(rn2[c] / bn2[c]) : 0.0;
            maxrel = std::max(maxrel, rel);
            const bool act = bn2[c] > 0.0 && rel > o.tol;
            scale[c] = act ? 1.0 / std::sqrt(rn2[c]) : 0.0;
            nact += act;
        }
        if (nact == 0) { done = true; break; }
        if (outer >= o.max_outer) break;
        SFH_CHECK_HIP(hipMemcpyAsync(S + S_SCALE * kp, scale.data(), kp * 8, hipMemcpyHostToDevice, st));
        SFH_TRY(dispatch_inner<T>(s, P, o, vaT, dinvT, iters));
        k_refine<T><<<(unsigned)((nk + 255) / 256), 256, 0, st>>>((long long)nk, kp, S + S_SCALE * kp,
                                                                 w.x32.as<T>(), w.x64.as<double>());
        launch_resid(P, st, rp, ci, va, w.x64.as<double>(), w.b64.as<double>(), w.r64.as<double>(),
                     w.part0.as<double>());
        k_red_norm<<<P.nred, kRedB, 0, st>>>(kp, P.nblk, w.part0.as<double>(), S + S_RN2 * kp);
        SFH_LAUNCH_CHECK();
        SFH_CHECK_HIP(hipMemcpyAsync(rn2.data(), S + S_RN2 * kp, kp * 8, hipMemcpyDeviceToHost, st));
        SFH_CHECK_HIP(hipStreamSynchronize(st));
        ++outer;
    }
    k_finalize<<<(unsigned)(((size_t)n * k + 255) / 256), 256, 0, st>>>(
        n, k, kp, trans, B, ldb, w.x64.as<double>(), rp, ci, va, s->diagpos.as<int>(),
        s->cmask.as<unsigned char>(), X, ldx);
    SFH_LAUNCH_CHECK();
    SFH_CHECK_HIP(hipEventRecord(s->ctx->ev1, st));
    SFH_CHECK_HIP(hipEventSynchronize(s->ctx->ev1));
    float ms = 0.f;
    SFH_CHECK_HIP(hipEventElapsedTime(&ms, s->ctx->ev0, s->ctx->ev1));
    if (stats) {
        stats->outer = outer;
        stats->inner_iters = iters;
        stats->max_relres = maxrel;
        stats->ms = ms;
        stats->converged = done ? 1 : 0;
    }
    if (!done) {
        set_error("solve did not reach tol %.3g (max rel. residual %.3g after %d steps, %d inner iterations)",
                  o.tol, maxrel, outer, iters);
        return SFH_ENOCONV;
    }
    return SFH_OK;
}

template <class T, int LPR>
int time_kernels(sfh_system* s, const Plan& P, int reps, double* ms) {
    hipStream_t st = s->ctx->stream;
    SolveWork& w = s->work;
    const int kp = P.kp;
    double* S = w.scal.as<double>();
    hipEvent_t e[4];
    for (auto& x : e) SFH_CHECK_HIP(hipEventCreate(&x));
    const T* va = (const T*)s->vals32.p;
    const T* dinv = (const T*)s->dinv32.p;
    SFH_CHECK_HIP(hipEventRecord(e[0], st));
    for (int r = 0; r < reps; ++r)
        Kern<T, LPR>::spmv(P, st, s->rowptr.as<int>(), s->colind.as<int>(), va, w.p32.as<T>(), w.q32.as<T>(),
                           w.part0.as<double>());
    SFH_CHECK_HIP(hipEventRecord(e[1], st));
    for (int r = 0; r < reps; ++r)
        Kern<T, LPR>::update(P, st, S + S_ALPHA * kp, dinv, w.x32.as<T>(), w.r32.as<T>(), w.p32.as<T>(),
                             w.q32.as<T>(), w.part0.as<double>(), w.part1.as<double>());
    SFH_CHECK_HIP(hipEventRecord(e[2], st));
    for (int r = 0; r < reps; ++r)
        Kern<T, LPR>::pupdate(P, st, S + S_BETA * kp, dinv, w.r32.as<T>(), w.p32.as<T>());
    SFH_CHECK_HIP(hipEventRecord(e[3], st));
    SFH_LAUNCH_CHECK();
    SFH_CHECK_HIP(hipEventSynchronize(e[3]));
    for (int i = 0; i < 3; ++i) {
        float f = 0.f;
        SFH_CHECK_HIP(hipEventElapsedTime(&f, e[i], e[i + 1]));
        ms[i] = f / reps;
    }
    for (auto& x : e) (void)hipEventDestroy(x);
    return SFH_OK;
}

}  // namespace

// Average device time per launch of the three inner-PCG kernels on the
// workspace of the last k-RHS solve (whose buffers hold finite values).
int profile_pcg(sfh_system* s, int k, int reps, double* ms) {
    const Plan P = make_plan(s->ndof, k, s->ctx->num_cu);
    if (s->work.x32.bytes < (size_t)P.n * P.kp * sizeof(float) || s->work.part0.bytes < (size_t)P.nblk * P.kp * 8) {
        set_error("run a %d-RHS FP32 solve before profiling", k);
        return SFH_ESTATE;
    }
    switch (P.lpr) {
        case 1: return time_kernels<float, 1>(s, P, reps, ms);
        case 2: return time_kernels<float, 2>(s, P, reps, ms);
        case 4: return time_kernels<float, 4>(s, P, reps, ms);
        case 8: return time_kernels<float, 8>(s, P, reps, ms);
        case 16: return time_kernels<float, 16>(s, P, reps, ms);
        case 32: return time_kernels<float, 32>(s, P, reps, ms);
        default: return time_kernels<float, 64>(s, P, reps, ms);
    }
}

int solve(sfh_system* s, int trans, int k, const double* B, int ldb, double* X, int ldx,
          const sfh_solve_opts* opts, sfh_solve_stats* stats) {
    if (!s->assembled) { set_error("solve before assembly"); return SFH_ESTATE; }
    if (k <= 0 || ldb < k || ldx < k || !B || !X) { set_error("bad solve arguments"); return SFH_EARG; }
    sfh_solve_opts o;
    sfh_solve_opts_default(&o);
    if (opts) o = *opts;
    if (o.tol <= 0) o.tol = 1e-10;
    if (o.inner_tol <= 0) o.inner_tol = 1e-5;
    if (o.max_outer <= 0) o.max_outer = 8;
    if (o.max_inner <= 0) o.max_inner = 20000;
    if (o.check_every <= 0) o.check_every = 32;
    const Plan P = make_plan(s->ndof, k, s->ctx->num_cu);
    if (o.precision == 1) {
        if (opts == nullptr || opts->inner_tol <= 0) o.inner_tol = o.tol * 0.5;
        return run_solve<double>(s, P, trans, k, B, ldb, X, ldx, o, stats);
    }
    return run_solve<float>(s, P, trans, k, B, ldb, X, ldx, o, stats);
}

}  // namespace sfh

extern "C" void sfh_solve_opts_default(sfh_solve_opts* o) {
    o->tol = 1e-10;
    o->inner_tol = 1e-5;
    o->max_outer = 8;
    o->max_inner = 20000;
    o->check_every = 32;
    o->precision = 0;
}
