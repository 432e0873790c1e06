// Device-side building blocks shared by the PCG translation units: vector
// types, work-group tiles, the per-column scalar and accumulator slots of the
// solver, direction and step-length coefficients, and work-group reductions.
#pragma once

#include "common.hpp"

namespace sfh {
namespace {

constexpr int kB = 1024;   // 16 waves per work-group: 16 waves per SIMD at 1024 work-groups
constexpr int kRedB = 1024;   // reduction blocks: 16 row groups x 64 columns
constexpr int kU = 3;         // gathered neighbours in flight per lane in the direction kernel (9-point rows = 3 x 3)

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
    int lane, grp, ngrp, col0, r0, r1, bid, so;
};

// Row partition of a (segmented) solve: work-group b (logical id) owns rows
// [brow[b], brow[b+1]) of segment bseg[b]; a segment is one independent
// problem of a batch, with its own per-column scalars at offset seg * kp of
// each scalar slot (slot stride ks = nseg * kp).
struct Seg {
    const int* brow;
    const int* bseg;
    int ks;
};

template <int LPR>
__device__ inline Tile tile(int kp, const Seg& g) {
    Tile t;
    t.lane = threadIdx.x % LPR;
    t.grp = threadIdx.x / LPR;
    t.ngrp = blockDim.x / LPR;
    t.col0 = (blockIdx.y * LPR + t.lane) * 4;
    t.bid = xcd_block(blockIdx.x, gridDim.x);
    t.r0 = g.brow[t.bid];
    t.r1 = g.brow[t.bid + 1];
    t.so = g.bseg[t.bid] * kp;
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

// per-RHS scalar slots (each ks doubles). The first five are what the host
// reads back at every convergence check, in one copy: the state flags, the
// target trigger and S_CHKRR, the latest recursive r.r copied there for the check
enum { S_ACTIVE = 0, S_FROZEN, S_SCALE, S_TRIG, S_CHKRR, S_RZ, S_RZPREV, S_BETA, S_TGT, S_RATIO, S_RN2, S_BN2,
       S_NSLOT };

// Per-column FP64 accumulators of the iteration kernels (ks doubles each),
// rotating with the iteration index j of the current refinement cycle:
// p.Ap of iteration j in A_PQ + j % 2, r.z and r.r after iteration j in
// A_RZ + j % 3 and A_RR + j % 3. A slot is cleared by work-group 0 of the
// kernel launched before the one that accumulates into it, and no work-group of
// that launch reads it.
enum { A_PQ = 0, A_RZ = 2, A_RR = 5, A_N = 8 };

// r.z that the step length of iteration j uses
__device__ inline double rz_at(int j, int ks, int c, const double* __restrict__ S, const double* __restrict__ acc) {
    return j == 0 ? S[S_RZ * ks + c] : acc[(A_RZ + (j - 1) % 3) * ks + c];
}

// Step length of iteration j - 1 (j >= 1) of column c, formed again from its
// accumulator slots exactly as the update kernel of j - 1 formed it: zero for
// a column that was not live then (inactive, or frozen at an iteration before
// j: S_FROZEN holds the freezing iteration + 1)
__device__ inline double alpha_before(int j, int ks, int c, const double* __restrict__ S,
                                      const double* __restrict__ acc) {
    const double fz = S[S_FROZEN * ks + c];
    const bool live = S[S_ACTIVE * ks + c] > 0.0 && !(fz > 0.0 && fz <= (double)j);
    const double pq = acc[(A_PQ + ((j - 1) & 1)) * ks + c];
    return (live && pq > 0.0) ? rz_at(j - 1, ks, c, S, acc) / pq : 0.0;
}

// Direction coefficients of column c at iteration j: p = m z + b p_old with
// b = beta, m = 1 for a live column and b = 1, m = 0 for a frozen or inactive
// one. A live column whose recursive residual r.r reached its trigger, or whose
// r.z is no longer positive (the FP32 recurrence has reached its rounding
// floor), is frozen here (until the next FP64 replacement); every work-group
// takes the same decision from the same inputs, so the flag and the r.z of the
// column's last step are written with the same values by all of them.
template <class T>
__device__ inline void dir_coef(int j, int ks, int c, double* S, const double* __restrict__ acc, T& b, T& m) {
    const bool act = S[S_ACTIVE * ks + c] > 0.0;
    bool frozen = false;
    double beta = S[S_BETA * ks + c];
    if (j > 0 && act) {
        frozen = S[S_FROZEN * ks + c] > 0.0;
        if (!frozen && (!(acc[(A_RR + (j - 1) % 3) * ks + c] > S[S_TRIG * ks + c]) ||
                        !(rz_at(j, ks, c, S, acc) > 0.0))) {
            frozen = true;
            S[S_FROZEN * ks + c] = (double)(j + 1);   // the freezing iteration + 1
            S[S_RZPREV * ks + c] = rz_at(j - 1, ks, c, S, acc);
        }
        const double rzo = rz_at(j - 1, ks, c, S, acc);
        beta = rzo > 0.0 ? rz_at(j, ks, c, S, acc) / rzo : 0.0;
    }
    const bool live = act && !frozen;
    b = live ? (T)beta : (T)1;
    m = live ? (T)1 : (T)0;
}

// Per-column direction coefficients (b, m) or step lengths of the work-group's
// 4*LPR columns, computed once per work-group into LDS and read by every row
// group (the inputs are a handful of scattered scalars per column).
template <class T, int LPR>
__device__ inline void shared_dir_coef(int it, int kp, int ks, int so, double* S, const double* __restrict__ acc,
                                       T* cb, T* cm) {
    const int w = 4 * LPR;
    for (int c = threadIdx.x; c < w; c += blockDim.x) {
        const int col = blockIdx.y * w + c;
        T b = (T)1, m = (T)0;
        if (col < kp) dir_coef<T>(it, ks, so + col, S, acc, b, m);
        cb[c] = b;
        cm[c] = m;
    }
    __syncthreads();
}

template <class T, int LPR>
__device__ inline void shared_alpha(int it, int kp, int ks, int so, const double* __restrict__ S,
                                    const double* __restrict__ acc, T* ca) {
    const int w = 4 * LPR;
    for (int c = threadIdx.x; c < w; c += blockDim.x) {
        const int col = blockIdx.y * w + c;
        double a = 0.0;
        if (col < kp) {
            const int k = so + col;
            const double pq = acc[(A_PQ + (it & 1)) * ks + k];
            const bool live = S[S_ACTIVE * ks + k] > 0.0 && !(S[S_FROZEN * ks + k] > 0.0);
            a = (live && pq > 0.0) ? rz_at(it, ks, k, S, acc) / pq : 0.0;
        }
        ca[c] = (T)a;
    }
    __syncthreads();
}

// work-group 0 clears `nslot` consecutive accumulator slots starting at slot a
__device__ inline void clear_slots(double* acc, int a, int nslot, int ks) {
    if (blockIdx.x != 0 || blockIdx.y != 0) return;
    for (int i = threadIdx.x; i < nslot * ks; i += blockDim.x) acc[(size_t)a * ks + i] = 0.0;
}

// block-level sum over row groups of per-lane 4-column partials, added into
// acc[col] with one hardware FP64 atomic per column and block (the order of the
// additions across blocks is not fixed)
template <int LPR>
__device__ inline void block_accumulate(const Tile& t, int kp, const V4<double>& a, double* red, double* acc) {
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
        if (col < kp) unsafeAtomicAdd(acc + col, s);
    }
}

// true when none of the columns the work-group handles is still being solved
// (converged, zero or padding columns have scale 0): the whole work-group then
// skips the iteration. Frozen columns (scale > 0) keep the work-group live.
__device__ inline bool block_done(const Tile& t, int kp, int ks, const double* __restrict__ S) {
    bool live = false;
    if (t.col0 < kp) {
        const double* sc = S + S_SCALE * ks + t.so + t.col0;
        live = sc[0] > 0.0 || sc[1] > 0.0 || sc[2] > 0.0 || sc[3] > 0.0;
    }
    return !__syncthreads_or(live);
}

// The multigrid iteration kernels also skip columns frozen at an iteration
// before `before` (S_FROZEN holds the freezing iteration + 1; 0: not frozen):
// the freezing iteration wrote p_new = p_old, so both direction buffers hold
// the column's kept direction, and its step length is 0, so the iteration
// changes nothing of it. Columns that freeze in the current launch (their flag
// is 0 or is being set to it + 1 by sibling work-groups) stay live.
// `lane` receives whether the lane's own four columns are live: a live
// work-group's idle lanes neither load nor store (a batch's narrower problems
// leave whole lanes of padding columns).
__device__ inline bool block_idle(const Tile& t, int kp, int ks, const double* __restrict__ S, int before,
                                  bool& lane) {
    bool live = false;
    if (t.col0 < kp) {
        const double* sc = S + S_SCALE * ks + t.so + t.col0;
        const double* fz = S + S_FROZEN * ks + t.so + t.col0;
#pragma unroll
        for (int q = 0; q < 4; ++q) live = live || (sc[q] > 0.0 && !(fz[q] > 0.0 && fz[q] <= (double)before));
    }
    lane = live;
    return !__syncthreads_or(live);
}

// row index of the lane group; wave-uniform (scalar) when one row spans the wave
template <int LPR>
__device__ inline int uniform_row(int r) {
    if (LPR == 64) return __builtin_amdgcn_readfirstlane(r);
    return r;
}

// operator entries of one CSR row: FP32 as packed (column, value bits) pairs
// (one 8-byte load per entry), FP64 from the separate column / value arrays
template <class T> struct Entries;
template <> struct Entries<float> {
    const int2* cv;
    const int* ci;
    const float* va;
    __device__ inline void get(int e, int& c, float& a) const {
        const int2 v = cv[e];
        c = v.x;
        a = __int_as_float(v.y);
    }
};
template <> struct Entries<double> {
    const int2* cv;
    const int* ci;
    const double* va;
    __device__ inline void get(int e, int& c, double& a) const {
        c = ci[e];
        a = va[e];
    }
};

// preconditioned direction entry p = m z + b p_old (m = 0 freezes a column),
// evaluated identically wherever it is needed (fused into the SpMV gathers)
template <class T>
__device__ inline void pdir4(const V4<T>& z, const V4<T>& m, const V4<T>& b, const V4<T>& po, V4<T>& out) {
    out.x = b.x * po.x + m.x * z.x; out.y = b.y * po.y + m.y * z.y;
    out.z = b.z * po.z + m.z * z.z; out.w = b.w * po.w + m.w * z.w;
}

// r.z and r.r of one row from z = D^-1 r (r = g z, g = A_ii)
template <class T>
__device__ inline void res_dots(const V4<T>& z, T g, V4<T>& rz, V4<T>& rr) {
    const V4<T> r = {g * z.x, g * z.y, g * z.z, g * z.w};
    rz.x += r.x * z.x; rz.y += r.y * z.y; rz.z += r.z * z.z; rz.w += r.w * z.w;
    rr.x += r.x * r.x; rr.y += r.y * r.y; rr.z += r.z * r.z; rr.w += r.w * r.w;
}

__device__ inline void fma4(float a, const float4& v, float4& s) {
    s.x += a * v.x; s.y += a * v.y; s.z += a * v.z; s.w += a * v.w;
}
__device__ inline float4 ldf4(const float* p) { return *reinterpret_cast<const float4*>(p); }
__device__ inline float4 ldf4(const double* p) {
    const V4<double> v = ld4(p);
    return make_float4((float)v.x, (float)v.y, (float)v.z, (float)v.w);
}
__device__ inline float4 zero4() { return make_float4(0.f, 0.f, 0.f, 0.f); }

// e = C^-1 r on the coarsest level of every grid (dense FP32 inverses,
// row-major): work-group (column group x, grid y) stages C^-1 and its 16
// residual columns in LDS. Table per grid: nodes, C^-1 offset, first float of
// the grid's residual and correction in the coarse-level buffer.
constexpr int kCoarseCols = 16;
inline size_t mg_coarse_lds(int n) {
    return sizeof(float) * (((size_t)n * n + 3) / 4 * 4) + sizeof(float4) * (size_t)n * (kCoarseCols / 4);
}
__global__ __launch_bounds__(256) void k_mg_coarse(const int* __restrict__ ctab, int kp, const float* __restrict__ cinv,
                                                   float* __restrict__ mgv) {
    extern __shared__ float4 sm[];
    const int* d = ctab + 4 * blockIdx.y;
    const int n = d[0];
    const float* ci_g = cinv + d[1];
    const float* r = mgv + d[2];
    float* e = mgv + d[3];
    float* ci = reinterpret_cast<float*>(sm);
    float4* rs = sm + ((size_t)n * n + 3) / 4;
    constexpr int nq = kCoarseCols / 4;
    const int c0 = blockIdx.x * kCoarseCols;
    for (int i = threadIdx.x; i < n * n; i += blockDim.x) ci[i] = ci_g[i];
    for (int i = threadIdx.x; i < n * nq; i += blockDim.x) {
        const int j = i / nq, col = c0 + 4 * (i % nq);
        rs[i] = col < kp ? *reinterpret_cast<const float4*>(r + (size_t)j * kp + col) : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    __syncthreads();
    for (int o = threadIdx.x; o < n * nq; o += blockDim.x) {
        const int i = o / nq, q = o % nq, col = c0 + 4 * q;
        if (col >= kp) continue;
        float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int j = 0; j < n; ++j) fma4(ci[i * n + j], rs[j * nq + q], s);
        *reinterpret_cast<float4*>(e + (size_t)i * kp + col) = s;
    }
}

}  // namespace
}  // namespace sfh
