// Multi-right-hand-side Jacobi-PCG with FP64 residual refinement.
//
// Replaces the reference's direct factorisation + per-RHS transposed solves
// (corlib/SystemSolveEigenSparse.hpp:273-404 called in a loop by
// statX/infer/ForwardGaussRHS.hpp:105-112) with one batched iterative solve:
//
//   FP32: Jacobi-PCG recurrences, every RHS with its own alpha / beta, two
//         fused kernels per iteration (direction + SpMV + p.Ap; SpMV + x/r
//         update + r.z, r.r), 32 bytes of vector traffic per DOF-RHS;
//   FP64: residual replacement -- the FP32 correction is flushed into the FP64
//         iterate and r = b - A x is recomputed in FP64 whenever a column's
//         recursive residual fell by inner_tol since the last replacement; the
//         search direction is kept, so no Krylov information is lost.
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
// The iteration kernels' dot products are added into per-column FP64
// accumulators with hardware atomics (one per column and work-group); the
// refinement-time reductions are fixed-order two-stage sums.
#include "pcg_dev.hpp"

namespace sfh {
namespace {

// K1: p_new = z + beta p_old (z = D^-1 r is what the solver stores; every row,
// recomputed on the fly at the neighbours the stencil touches), q = A p_new,
// acc += sum_rows p_new . q. q itself is not stored.
template <class T, int LPR>
__global__ __launch_bounds__(kB) void k_dir_spmv(int kp, Seg G, int it, const int* __restrict__ rp,
                                                 const Entries<T> M, double* __restrict__ S,
                                                 const T* __restrict__ z, const T* __restrict__ pold,
                                                 T* __restrict__ pnew, double* __restrict__ acc) {
    __shared__ double red[kB * 4];
    const Tile t = tile<LPR>(kp, G);
    clear_slots(acc, A_RZ + it % 3, 1, G.ks);
    clear_slots(acc, A_RR + it % 3, 1, G.ks);
    if (block_done(t, kp, G.ks, S)) return;
    __shared__ T cb[4 * LPR], cm[4 * LPR];
    shared_dir_coef<T, LPR>(it, kp, G.ks, t.so, S, acc, cb, cm);
    const bool ok = t.col0 < kp;
    V4<T> pq = {0, 0, 0, 0};
    if (ok) {
        // frozen columns (waiting for residual replacement) keep their direction:
        // p = 1 * p_old + 0 * z
        const int lc = t.lane * 4;
        const V4<T> b = {cb[lc], cb[lc + 1], cb[lc + 2], cb[lc + 3]};
        const V4<T> m = {cm[lc], cm[lc + 1], cm[lc + 2], cm[lc + 3]};
        for (int row0 = t.r0 + t.grp; row0 < t.r1; row0 += t.ngrp) {
            const int row = uniform_row<LPR>(row0);
            V4<T> q = {0, 0, 0, 0};
            int e = rp[row];
            const int e1 = rp[row + 1];
            for (; e + kU <= e1; e += kU) {
                int c[kU];
                T a[kU];
                V4<T> zv[kU], pv[kU];
#pragma unroll
                for (int u = 0; u < kU; ++u) M.get(e + u, c[u], a[u]);
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    zv[u] = ld4(z + (size_t)c[u] * kp + t.col0);
                    pv[u] = ld4(pold + (size_t)c[u] * kp + t.col0);
                }
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    V4<T> pj;
                    pdir4(zv[u], m, b, pv[u], pj);
                    q.x += a[u] * pj.x; q.y += a[u] * pj.y; q.z += a[u] * pj.z; q.w += a[u] * pj.w;
                }
            }
            for (; e < e1; ++e) {
                int c;
                T a;
                M.get(e, c, a);
                V4<T> pj;
                pdir4(ld4(z + (size_t)c * kp + t.col0), m, b, ld4(pold + (size_t)c * kp + t.col0), pj);
                q.x += a * pj.x; q.y += a * pj.y; q.z += a * pj.z; q.w += a * pj.w;
            }
            const size_t o = (size_t)row * kp + t.col0;
            V4<T> pi;
            pdir4(ld4(z + o), m, b, ld4(pold + o), pi);
            st4(pnew + o, pi);
            pq.x += pi.x * q.x; pq.y += pi.y * q.y; pq.z += pi.z * q.z; pq.w += pi.w * q.w;
        }
    }
    const V4<double> ad = {(double)pq.x, (double)pq.y, (double)pq.z, (double)pq.w};
    block_accumulate<LPR>(t, kp, ad, red, acc + (A_PQ + (it & 1)) * G.ks + t.so);
}

// K2: q = A p (recomputed), x += alpha p, z -= alpha D^-1 q (z = D^-1 r),
// acc += r.z, r.r with r = diag z
template <class T, int LPR>
__global__ __launch_bounds__(kB) void k_cg_update(int kp, Seg G, int it, const int* __restrict__ rp,
                                                  const Entries<T> M, const T* __restrict__ dinv,
                                                  const T* __restrict__ diag, const double* __restrict__ S,
                                                  const T* __restrict__ p, T* __restrict__ x, T* __restrict__ z,
                                                  double* __restrict__ acc) {
    __shared__ double red[kB * 4];
    const Tile t = tile<LPR>(kp, G);
    clear_slots(acc, A_PQ + ((it + 1) & 1), 1, G.ks);
    if (block_done(t, kp, G.ks, S)) return;
    __shared__ T ca[4 * LPR];
    shared_alpha<T, LPR>(it, kp, G.ks, t.so, S, acc, ca);
    const bool ok = t.col0 < kp;
    V4<T> rz = {0, 0, 0, 0}, rr = {0, 0, 0, 0};
    if (ok) {
        const V4<T> al = {ca[t.lane * 4], ca[t.lane * 4 + 1], ca[t.lane * 4 + 2], ca[t.lane * 4 + 3]};
        for (int row0 = t.r0 + t.grp; row0 < t.r1; row0 += t.ngrp) {
            const int row = uniform_row<LPR>(row0);
            V4<T> q = {0, 0, 0, 0};
            int e = rp[row];
            const int e1 = rp[row + 1];
            for (; e + 4 <= e1; e += 4) {
                int c[4];
                T a[4];
                V4<T> pv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) M.get(e + u, c[u], a[u]);
#pragma unroll
                for (int u = 0; u < 4; ++u) pv[u] = ld4(p + (size_t)c[u] * kp + t.col0);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    q.x += a[u] * pv[u].x; q.y += a[u] * pv[u].y; q.z += a[u] * pv[u].z; q.w += a[u] * pv[u].w;
                }
            }
            for (; e < e1; ++e) {
                int c;
                T a;
                M.get(e, c, a);
                const V4<T> pv = ld4(p + (size_t)c * kp + t.col0);
                q.x += a * pv.x; q.y += a * pv.y; q.z += a * pv.z; q.w += a * pv.w;
            }
            const size_t o = (size_t)row * kp + t.col0;
            const V4<T> pi = ld4(p + o);
            V4<T> xv = ld4(x + o), zv = ld4(z + o);
            const T d = dinv[row], g = diag[row];
            xv.x += al.x * pi.x; xv.y += al.y * pi.y; xv.z += al.z * pi.z; xv.w += al.w * pi.w;
            zv.x -= al.x * d * q.x; zv.y -= al.y * d * q.y; zv.z -= al.z * d * q.z; zv.w -= al.w * d * q.w;
            st4(x + o, xv);
            st4(z + o, zv);
            res_dots(zv, g, rz, rr);
        }
    }
    V4<double> a = {(double)rz.x, (double)rz.y, (double)rz.z, (double)rz.w};
    block_accumulate<LPR>(t, kp, a, red, acc + (A_RZ + it % 3) * G.ks + t.so);
    __syncthreads();
    V4<double> bb = {(double)rr.x, (double)rr.y, (double)rr.z, (double)rr.w};
    block_accumulate<LPR>(t, kp, bb, red, acc + (A_RR + it % 3) * G.ks + t.so);
}

// Residual replacement: z = D^-1 (s' r64), p *= ratio (s'/s), part0 = r.z,
// part1 = r.r
template <class T, int LPR>
__global__ __launch_bounds__(kB) void k_rescale(int kp, Seg G, const double* __restrict__ scale,
                                                const double* __restrict__ ratio, const double* __restrict__ dinv64,
                                                const T* __restrict__ diag, const double* __restrict__ r64,
                                                T* __restrict__ z, T* __restrict__ p, double* __restrict__ part0,
                                                double* __restrict__ part1) {
    __shared__ double red[kB * 4];
    const Tile t = tile<LPR>(kp, G);
    const bool ok = t.col0 < kp;
    V4<T> rz = {0, 0, 0, 0}, rr = {0, 0, 0, 0};
    if (ok) {
        const double* sc = scale + t.so + t.col0;
        const double* ra = ratio + t.so + t.col0;
        const double s0 = sc[0], s1 = sc[1], s2 = sc[2], s3 = sc[3];
        const V4<T> q = {(T)ra[0], (T)ra[1], (T)ra[2], (T)ra[3]};
        for (int row = t.r0 + t.grp; row < t.r1; row += t.ngrp) {
            const size_t o = (size_t)row * kp + t.col0;
            const V4<double> rd = ld4(r64 + o);
            const double di = dinv64[row];
            const V4<T> zv = {(T)(rd.x * s0 * di), (T)(rd.y * s1 * di), (T)(rd.z * s2 * di), (T)(rd.w * s3 * di)};
            V4<T> pv = ld4(p + o);
            // a zero ratio starts the direction afresh (p = 0, whatever p held)
            pv.x = q.x != (T)0 ? pv.x * q.x : (T)0;
            pv.y = q.y != (T)0 ? pv.y * q.y : (T)0;
            pv.z = q.z != (T)0 ? pv.z * q.z : (T)0;
            pv.w = q.w != (T)0 ? pv.w * q.w : (T)0;
            st4(z + o, zv);
            st4(p + o, pv);
            res_dots(zv, diag[row], rz, rr);
        }
    }
    V4<double> a = {(double)rz.x, (double)rz.y, (double)rz.z, (double)rz.w};
    block_partial<LPR>(t, kp, a, red, part0);
    __syncthreads();
    V4<double> bb = {(double)rr.x, (double)rr.y, (double)rr.z, (double)rr.w};
    block_partial<LPR>(t, kp, bb, red, part1);
}

// ------------------------------------------------ 2D tensor-grid (DIA-9) path
// Nodes are numbered x fastest on a W x H grid and the operator is stored as 9
// coefficients per node (coefficient (dy+1)*3 + (dx+1) couples (x+dx, y+dy)).
// A work-group owns a kTX x kTY tile of nodes: it stages the tile plus a
// one-node halo of the gathered vector in LDS with coalesced 16-byte loads
// (each node row of kp columns is contiguous), then forms the 9-point products
// from LDS. Every vector element is read from memory about (kTX+2)(kTY+2) /
// (kTX kTY) times instead of once per stencil neighbour, so the iteration is
// bound by streaming bandwidth rather than by gather instructions. Column tiles
// of up to 64 right-hand sides (LPR <= 16 lanes of 4 columns per node), FP32.
constexpr int kTX = 32, kTY = 6, kG = 512;
constexpr int kHN = (kTX + 2) * (kTY + 2);   // staged nodes per tile

// threads per work-group of the tile kernels: kG, or 128 lane groups when
// LPR < 4 (a tile then keeps every lane busy; at kG threads and one lane per
// node most lane groups would have no node of the tile)
template <int LPR> struct GB { static constexpr int v = LPR >= 4 ? kG : 128 * LPR; };
inline int grid_block(int lpr) { return lpr >= 4 ? kG : 128 * lpr; }

inline size_t grid_lds_bytes(int lpr) {
    const size_t tile = (size_t)kHN * lpr * sizeof(float4);
    const size_t red = (size_t)grid_block(lpr) * 4 * sizeof(double);   // block_accumulate after the tile is consumed
    return tile > red ? tile : red;
}

// Work-group tile table, 6 ints per (logical) work-group: first node of the
// grid's rows, grid width W and height H (W = 0: idle padding work-group),
// tile origin x0, y0, and the segment (problem) whose per-column scalars apply.
constexpr int kGT = 6;

template <int LPR>
__device__ inline Tile grid_tile_t(int kp, const int* __restrict__ gtab, int* d) {
    Tile t;
    t.lane = threadIdx.x % LPR;
    t.grp = threadIdx.x / LPR;
    t.ngrp = blockDim.x / LPR;
    t.col0 = (blockIdx.y * LPR + t.lane) * 4;   // column tile blockIdx.y of 4 * LPR columns
    t.bid = xcd_block(blockIdx.x, gridDim.x);
#pragma unroll
    for (int i = 0; i < kGT; ++i) d[i] = gtab[kGT * t.bid + i];
    t.r0 = t.r1 = 0;
    t.so = d[5] * kp;
    return t;
}

__device__ inline float4 f4(const V4<float>& v) { return make_float4(v.x, v.y, v.z, v.w); }
__device__ inline V4<float> v4(const float4& v) { return {v.x, v.y, v.z, v.w}; }

// q = 9-point product at tile node (ox, oy) from the staged values
template <int LPR>
__device__ inline V4<float> stencil9(const float4* sm, const float* __restrict__ c, int ox, int oy, int lane) {
    V4<float> q = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int dy = 0; dy < 3; ++dy) {
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
            const float a = c[dy * 3 + dx];
            const float4 v = sm[((oy + dy) * (kTX + 2) + (ox + dx)) * LPR + lane];
            q.x += a * v.x; q.y += a * v.y; q.z += a * v.z; q.w += a * v.w;
        }
    }
    return q;
}

template <int LPR>
__global__ __launch_bounds__(GB<LPR>::v) void k_dir_spmv_t2(const int* __restrict__ gtab, int kp, int ks, int it,
                                                    const float* __restrict__ coef, double* __restrict__ S,
                                                    const float* __restrict__ z, const float* __restrict__ pold,
                                                    float* __restrict__ pnew, double* __restrict__ acc) {
    extern __shared__ float4 sm[];
    int d[kGT];
    clear_slots(acc, A_RZ + it % 3, 1, ks);
    clear_slots(acc, A_RR + it % 3, 1, ks);
    const Tile t = grid_tile_t<LPR>(kp, gtab, d);
    const int base = d[0], W = d[1], H = d[2], tx0 = d[3], ty0 = d[4];
    if (W == 0) return;   // padding work-groups (whole group)
    if (block_done(t, kp, ks, S)) return;
    __shared__ float cb[4 * LPR], cm[4 * LPR];
    shared_dir_coef<float, LPR>(it, kp, ks, t.so, S, acc, cb, cm);
    const bool ok = t.col0 < kp;
    const int lc = t.lane * 4;
    const V4<float> b = {cb[lc], cb[lc + 1], cb[lc + 2], cb[lc + 3]};
    const V4<float> m = {cm[lc], cm[lc + 1], cm[lc + 2], cm[lc + 3]};
    // stage p_new = m z + b p_old on the tile and its halo; write the tile's own nodes
    for (int h = t.grp; h < kHN; h += t.ngrp) {
        const int hx = h % (kTX + 2), hy = h / (kTX + 2);
        const int gx = tx0 + hx - 1, gy = ty0 + hy - 1;
        V4<float> v = {0.f, 0.f, 0.f, 0.f};
        if (ok && gx >= 0 && gx < W && gy >= 0 && gy < H) {
            const size_t o = (size_t)(base + gy * W + gx) * kp + t.col0;
            pdir4(ld4(z + o), m, b, ld4(pold + o), v);
            if (hx >= 1 && hx <= kTX && hy >= 1 && hy <= kTY) st4(pnew + o, v);
        }
        sm[h * LPR + t.lane] = f4(v);
    }
    __syncthreads();
    V4<float> pq = {0.f, 0.f, 0.f, 0.f};
    if (ok) {
        for (int o = t.grp; o < kTX * kTY; o += t.ngrp) {
            const int ox = o % kTX, oy = o / kTX;
            const int gx = tx0 + ox, gy = ty0 + oy;
            if (gx >= W || gy >= H) continue;
            const V4<float> q = stencil9<LPR>(sm, coef + (size_t)(base + gy * W + gx) * 9, ox, oy, t.lane);
            const V4<float> pi = v4(sm[((oy + 1) * (kTX + 2) + (ox + 1)) * LPR + t.lane]);
            pq.x += pi.x * q.x; pq.y += pi.y * q.y; pq.z += pi.z * q.z; pq.w += pi.w * q.w;
        }
    }
    __syncthreads();   // the staging area becomes the reduction buffer
    const V4<double> ad = {(double)pq.x, (double)pq.y, (double)pq.z, (double)pq.w};
    block_accumulate<LPR>(t, kp, ad, reinterpret_cast<double*>(sm), acc + (A_PQ + (it & 1)) * ks + t.so);
}

template <int LPR>
__global__ __launch_bounds__(GB<LPR>::v) void k_cg_update_t2(const int* __restrict__ gtab, int kp, int ks, int it,
                                                     const float* __restrict__ coef, const float* __restrict__ dinv,
                                                     const float* __restrict__ diag, const double* __restrict__ S,
                                                     const float* __restrict__ p, float* __restrict__ x,
                                                     float* __restrict__ z, double* __restrict__ acc) {
    extern __shared__ float4 sm[];
    int d[kGT];
    clear_slots(acc, A_PQ + ((it + 1) & 1), 1, ks);
    const Tile t = grid_tile_t<LPR>(kp, gtab, d);
    const int base = d[0], W = d[1], H = d[2], tx0 = d[3], ty0 = d[4];
    if (W == 0) return;
    if (block_done(t, kp, ks, S)) return;
    __shared__ float ca[4 * LPR];
    shared_alpha<float, LPR>(it, kp, ks, t.so, S, acc, ca);
    const bool ok = t.col0 < kp;
    for (int h = t.grp; h < kHN; h += t.ngrp) {
        const int hx = h % (kTX + 2), hy = h / (kTX + 2);
        const int gx = tx0 + hx - 1, gy = ty0 + hy - 1;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (ok && gx >= 0 && gx < W && gy >= 0 && gy < H)
            v = *reinterpret_cast<const float4*>(p + (size_t)(base + gy * W + gx) * kp + t.col0);
        sm[h * LPR + t.lane] = v;
    }
    __syncthreads();
    V4<float> rz = {0.f, 0.f, 0.f, 0.f}, rr = {0.f, 0.f, 0.f, 0.f};
    if (ok) {
        const V4<float> al = {ca[t.lane * 4], ca[t.lane * 4 + 1], ca[t.lane * 4 + 2], ca[t.lane * 4 + 3]};
        for (int o = t.grp; o < kTX * kTY; o += t.ngrp) {
            const int ox = o % kTX, oy = o / kTX;
            const int gx = tx0 + ox, gy = ty0 + oy;
            if (gx >= W || gy >= H) continue;
            const int nd = base + gy * W + gx;
            const V4<float> q = stencil9<LPR>(sm, coef + (size_t)nd * 9, ox, oy, t.lane);
            const V4<float> pi = v4(sm[((oy + 1) * (kTX + 2) + (ox + 1)) * LPR + t.lane]);
            const size_t off = (size_t)nd * kp + t.col0;
            V4<float> xv = ld4(x + off), zv = ld4(z + off);
            const float dd = dinv[nd], g = diag[nd];
            xv.x += al.x * pi.x; xv.y += al.y * pi.y; xv.z += al.z * pi.z; xv.w += al.w * pi.w;
            zv.x -= al.x * dd * q.x; zv.y -= al.y * dd * q.y; zv.z -= al.z * dd * q.z; zv.w -= al.w * dd * q.w;
            st4(x + off, xv);
            st4(z + off, zv);
            res_dots(zv, g, rz, rr);
        }
    }
    __syncthreads();
    double* red = reinterpret_cast<double*>(sm);
    V4<double> a = {(double)rz.x, (double)rz.y, (double)rz.z, (double)rz.w};
    block_accumulate<LPR>(t, kp, a, red, acc + (A_RZ + it % 3) * ks + t.so);
    __syncthreads();
    V4<double> bb = {(double)rr.x, (double)rr.y, (double)rr.z, (double)rr.w};
    block_accumulate<LPR>(t, kp, bb, red, acc + (A_RR + it % 3) * ks + t.so);
}

// ------------------------------------------------ multigrid-preconditioned iteration (2D grids)
// With the multigrid preconditioner one iteration reduces the residual about
// tenfold, so the recurrences must stay accurate far below FP32's reach: for a
// smooth right-hand side on a fine grid, FP32 iterates and FP32 stencil
// products (whose 9 terms cancel down to the small A p of a smooth p) lose
// the true residual at about eps32 times the condition number. Here the
// iterate x and residual r are FP64 and q = A p is formed from the FP64
// stencil with FP64 sums; the direction p and the preconditioned residual z
// stay FP32 (p is exact as stored, so x += alpha p and r -= alpha A p remain
// consistent to FP64 rounding whatever p's precision). Vector traffic per
// DOF-RHS and iteration: p (staged), x and r read and written.

// stage an FP32 block on a kTX x kTY tile and its one-node halo (zeros off the
// grid and on idle lanes); all of a lane group's loads are issued first
template <int LPR>
__device__ inline void stage_grid_f32(float4* sm, const float* __restrict__ v, int kp, int lc, bool ok, int base,
                                      int W, int H, int tx0, int ty0, int grp, int lane) {
    constexpr int ngrp = GB<LPR>::v / LPR, kSt = (kHN + ngrp - 1) / ngrp;
    float4 a[kSt];
    bool in[kSt];
#pragma unroll
    for (int u = 0; u < kSt; ++u) {
        const int h = grp + u * ngrp;
        const int gx = tx0 + h % (kTX + 2) - 1, gy = ty0 + h / (kTX + 2) - 1;
        in[u] = ok && h < kHN && gx >= 0 && gx < W && gy >= 0 && gy < H;
        a[u] = *reinterpret_cast<const float4*>(v + (size_t)(in[u] ? base + gy * W + gx : 0) * kp + lc);
    }
#pragma unroll
    for (int u = 0; u < kSt; ++u) {
        const int h = grp + u * ngrp;
        if (h < kHN) sm[h * LPR + lane] = in[u] ? a[u] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// q = FP64 9-point product at tile node (ox, oy) from the staged FP32 values
template <int LPR>
__device__ inline V4<double> stencil9d(const float4* sm, const double* __restrict__ c, int ox, int oy, int lane) {
    V4<double> q = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int dy = 0; dy < 3; ++dy) {
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
            const double a = c[dy * 3 + dx];
            const float4 v = sm[((oy + dy) * (kTX + 2) + (ox + dx)) * LPR + lane];
            q.x += a * v.x; q.y += a * v.y; q.z += a * v.z; q.w += a * v.w;
        }
    }
    return q;
}

// p_new = m z + b p_old (FP32), p.Ap in FP64 from the FP64 stencil
template <int LPR>
__global__ __launch_bounds__(GB<LPR>::v) void k_dir_spmv_mg(const int* __restrict__ gtab, int kp, int ks, int it,
                                                    const double* __restrict__ coef, double* __restrict__ S,
                                                    const float* __restrict__ z, const float* __restrict__ pold,
                                                    float* __restrict__ pnew, double* __restrict__ acc) {
    extern __shared__ float4 sm[];
    int d[kGT];
    clear_slots(acc, A_RZ + it % 3, 1, ks);
    clear_slots(acc, A_RR + it % 3, 1, ks);
    const Tile t = grid_tile_t<LPR>(kp, gtab, d);
    const int base = d[0], W = d[1], H = d[2], tx0 = d[3], ty0 = d[4];
    if (W == 0) return;
    bool lane_live;
    if (block_idle(t, kp, ks, S, it, lane_live)) return;
    __shared__ float cb[4 * LPR], cm[4 * LPR];
    shared_dir_coef<float, LPR>(it, kp, ks, t.so, S, acc, cb, cm);
    const bool ok = t.col0 < kp && lane_live;
    const int lc = t.lane * 4;
    const V4<float> b = {cb[lc], cb[lc + 1], cb[lc + 2], cb[lc + 3]};
    const V4<float> m = {cm[lc], cm[lc + 1], cm[lc + 2], cm[lc + 3]};
    {
        constexpr int ngrp = GB<LPR>::v / LPR, kSt = (kHN + ngrp - 1) / ngrp;
        const int l0 = ok ? t.col0 : 0;
        V4<float> za[kSt], pa[kSt];
        bool in[kSt];
#pragma unroll
        for (int u = 0; u < kSt; ++u) {
            const int h = t.grp + u * ngrp;
            const int gx = tx0 + h % (kTX + 2) - 1, gy = ty0 + h / (kTX + 2) - 1;
            in[u] = ok && h < kHN && gx >= 0 && gx < W && gy >= 0 && gy < H;
            const size_t o = (size_t)(in[u] ? base + gy * W + gx : 0) * kp + l0;
            za[u] = ld4(z + o);
            pa[u] = ld4(pold + o);
        }
#pragma unroll
        for (int u = 0; u < kSt; ++u) {
            const int h = t.grp + u * ngrp;
            if (h >= kHN) continue;
            const int hx = h % (kTX + 2), hy = h / (kTX + 2);
            V4<float> v = {0.f, 0.f, 0.f, 0.f};
            if (in[u]) {
                pdir4(za[u], m, b, pa[u], v);
                if (hx >= 1 && hx <= kTX && hy >= 1 && hy <= kTY)
                    st4(pnew + (size_t)(base + (ty0 + hy - 1) * W + tx0 + hx - 1) * kp + t.col0, v);
            }
            sm[h * LPR + t.lane] = f4(v);
        }
    }
    __syncthreads();
    V4<double> pq = {0.0, 0.0, 0.0, 0.0};
    if (ok) {
        // kDN nodes per pass, all their stencil coefficients loaded first
        constexpr int kDN = 2;
        for (int o0 = t.grp; o0 < kTX * kTY; o0 += kDN * t.ngrp) {
            int o[kDN];
            bool in[kDN];
            double c[kDN][9];
#pragma unroll
            for (int u = 0; u < kDN; ++u) {
                o[u] = o0 + u * t.ngrp;
                const int gx = tx0 + o[u] % kTX, gy = ty0 + o[u] / kTX;
                in[u] = o[u] < kTX * kTY && gx < W && gy < H;
                const int nd = in[u] ? base + gy * W + gx : base;
#pragma unroll
                for (int k = 0; k < 9; ++k) c[u][k] = coef[(size_t)nd * 9 + k];
            }
#pragma unroll
            for (int u = 0; u < kDN; ++u) {
                if (!in[u]) continue;
                const int ox = o[u] % kTX, oy = o[u] / kTX;
                V4<double> q = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                    for (int dx = 0; dx < 3; ++dx) {
                        const double a = c[u][dy * 3 + dx];
                        const float4 v = sm[((oy + dy) * (kTX + 2) + (ox + dx)) * LPR + t.lane];
                        q.x += a * v.x; q.y += a * v.y; q.z += a * v.z; q.w += a * v.w;
                    }
                const float4 pi = sm[((oy + 1) * (kTX + 2) + (ox + 1)) * LPR + t.lane];
                pq.x += pi.x * q.x; pq.y += pi.y * q.y; pq.z += pi.z * q.z; pq.w += pi.w * q.w;
            }
        }
    }
    __syncthreads();
    block_accumulate<LPR>(t, kp, pq, reinterpret_cast<double*>(sm), acc + (A_PQ + (it & 1)) * ks + t.so);
}

// x += (alpha / s) p, r -= alpha A p, r.r: FP64 x, r, stencil and sums; r is
// the residual scaled by the column's s (unit norm at the last replacement), x
// the iterate itself. x moves every other iteration: on odd iterations j by
// (alpha_{j-1} p_{j-1} + alpha_j p_j) / s, p_{j-1} read from the other
// direction buffer (pprev), alpha_{j-1} formed again from its accumulator slots
// (the p.Ap slot of j - 1 is cleared only by the V-cycle after this launch);
// a cycle that ends after an even iteration flushes its last step with
// k_x_flush_mg. Work-groups skip as in block_idle, on odd iterations only
// when every column was frozen before j - 1's step too.
// XUP = (it odd): the instantiation with the x step holds x, p_{j-1} and the
// two step lengths; the other one needs none of their registers.
template <int LPR, bool XUP>
__global__ __launch_bounds__(GB<LPR>::v) __attribute__((amdgpu_waves_per_eu(4))) void k_cg_update_mg(const int* __restrict__ gtab, int kp, int ks, int it,
                                                     const double* __restrict__ coef, const double* __restrict__ S,
                                                     const float* __restrict__ p, const float* __restrict__ pprev,
                                                     double* __restrict__ x, double* __restrict__ r,
                                                     double* __restrict__ acc) {
    extern __shared__ float4 sm[];
    int d[kGT];
    constexpr bool xup = XUP;
    const Tile t = grid_tile_t<LPR>(kp, gtab, d);
    const int base = d[0], W = d[1], H = d[2], tx0 = d[3], ty0 = d[4];
    if (W == 0) return;
    bool lane_live;
    if (block_idle(t, kp, ks, S, xup ? it : it + 1, lane_live)) return;
    __shared__ double ca[4 * LPR], cx[4 * LPR], cx1[4 * LPR];
    shared_alpha<double, LPR>(it, kp, ks, t.so, S, acc, ca);
    if (xup)
        for (int c = threadIdx.x; c < 4 * LPR; c += blockDim.x) {
            const int col = blockIdx.y * 4 * LPR + c;
            const double sc = col < kp ? S[S_SCALE * ks + t.so + col] : 0.0;
            cx[c] = sc > 0.0 ? ca[c] / sc : 0.0;
            cx1[c] = sc > 0.0 ? alpha_before(it, ks, t.so + col, S, acc) / sc : 0.0;
        }
    const bool ok = t.col0 < kp && lane_live;
    // The thread's nodes o = grp, grp + ngrp, ... of the tile are software
    // pipelined: the next node's x, r, p_prev and stencil coefficients are in
    // flight while the current node is computed, and the first node's loads
    // are issued right after the staging loads of p, before the barrier (an
    // off-grid node loads the tile's first node's data and stores nothing).
    struct Node {
        V4<double> xv, rv;
        float4 pv;
        double c[9];
        int nd;
        bool in;
    };
    auto fetch = [&](int o, Node& n) {
        const int gx = tx0 + o % kTX, gy = ty0 + o / kTX;
        n.in = gx < W && gy < H;
        n.nd = n.in ? base + gy * W + gx : base;
        const size_t off = (size_t)n.nd * kp + t.col0;
        if (xup) {
            n.xv = ld4(x + off);
            n.pv = *reinterpret_cast<const float4*>(pprev + off);
        }
        n.rv = ld4(r + off);
#pragma unroll
        for (int k = 0; k < 9; ++k) n.c[k] = coef[(size_t)n.nd * 9 + k];
    };
    Node cur;
    {
        constexpr int ngrp = GB<LPR>::v / LPR, kSt = (kHN + ngrp - 1) / ngrp;
        const int lc = ok ? t.col0 : 0;
        float4 a[kSt];
        bool sin[kSt];
#pragma unroll
        for (int u = 0; u < kSt; ++u) {
            const int h = t.grp + u * ngrp;
            const int gx = tx0 + h % (kTX + 2) - 1, gy = ty0 + h / (kTX + 2) - 1;
            sin[u] = ok && h < kHN && gx >= 0 && gx < W && gy >= 0 && gy < H;
            a[u] = *reinterpret_cast<const float4*>(p + (size_t)(sin[u] ? base + gy * W + gx : 0) * kp + lc);
        }
        if (ok) fetch(t.grp, cur);   // t.grp < kTX * kTY: every lane group has a first node
#pragma unroll
        for (int u = 0; u < kSt; ++u) {
            const int h = t.grp + u * ngrp;
            if (h < kHN) sm[h * LPR + t.lane] = sin[u] ? a[u] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
    __syncthreads();
    V4<double> rr = {0.0, 0.0, 0.0, 0.0};
    if (ok) {
        const V4<double> al = {ca[t.lane * 4], ca[t.lane * 4 + 1], ca[t.lane * 4 + 2], ca[t.lane * 4 + 3]};
        V4<double> ax = {0.0, 0.0, 0.0, 0.0}, ax1 = {0.0, 0.0, 0.0, 0.0};
        if (xup) {
            ax = {cx[t.lane * 4], cx[t.lane * 4 + 1], cx[t.lane * 4 + 2], cx[t.lane * 4 + 3]};
            ax1 = {cx1[t.lane * 4], cx1[t.lane * 4 + 1], cx1[t.lane * 4 + 2], cx1[t.lane * 4 + 3]};
        }
        for (int o = t.grp; o < kTX * kTY; o += t.ngrp) {
            Node nxt;
            if (o + t.ngrp < kTX * kTY) fetch(o + t.ngrp, nxt);
            if (cur.in) {
                const int ox = o % kTX, oy = o / kTX;
                V4<double> q = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                    for (int dx = 0; dx < 3; ++dx) {
                        const double a = cur.c[dy * 3 + dx];
                        const float4 v = sm[((oy + dy) * (kTX + 2) + (ox + dx)) * LPR + t.lane];
                        q.x += a * v.x; q.y += a * v.y; q.z += a * v.z; q.w += a * v.w;
                    }
                const float4 pi = sm[((oy + 1) * (kTX + 2) + (ox + 1)) * LPR + t.lane];
                const size_t off = (size_t)cur.nd * kp + t.col0;
                V4<double> rn = cur.rv;
                rn.x -= al.x * q.x; rn.y -= al.y * q.y; rn.z -= al.z * q.z; rn.w -= al.w * q.w;
                if (xup) {
                    V4<double> xn = cur.xv;
                    const float4 pp = cur.pv;
                    xn.x += ax1.x * pp.x + ax.x * pi.x; xn.y += ax1.y * pp.y + ax.y * pi.y;
                    xn.z += ax1.z * pp.z + ax.z * pi.z; xn.w += ax1.w * pp.w + ax.w * pi.w;
                    st4(x + off, xn);
                }
                st4(r + off, rn);
                rr.x += rn.x * rn.x; rr.y += rn.y * rn.y; rr.z += rn.z * rn.z; rr.w += rn.w * rn.w;
            }
            cur = nxt;
        }
    }
    __syncthreads();
    block_accumulate<LPR>(t, kp, rr, reinterpret_cast<double*>(sm), acc + (A_RR + it % 3) * ks + t.so);
}

// ------------------------------------------------ multigrid V-cycle (2D grids)
// z = M r with M one V(1,1)-cycle of the hierarchy built by mg_build: on each
// level l, pre-smoothing from zero e = S r (S = omega D_l1^-1), restriction of
// the residual r - A e to level l+1, recursion, prolongation of the coarse
// correction, one post-smoothing sweep; a dense solve on the coarsest level.
// Two launches per level: the restriction kernel recomputes e = S r on the
// fly, the prolongation kernel rebuilds it and adds P e_{l+1} before the post
// sweep, so no level stores its pre-smoothed iterate. Level-0 constrained
// nodes carry neither residual nor correction, so M acts on the free block.
// Tiles stage their vectors (kp columns per node, 4 per lane) with a halo in
// LDS exactly like the iteration kernels above.
constexpr int kRX = 8, kRY = 4;                       // coarse nodes per restriction tile
constexpr int kRW = 2 * kRX + 3, kRH = 2 * kRY + 3;   // fine nodes it stages (residual support + halo)
constexpr int kPX = 16, kPY = 8;                      // fine nodes per prolongation tile
constexpr int kPW = kPX + 2, kPH = kPY + 2;
constexpr int kCW = kPX / 2 + 2, kCH = kPY / 2 + 2;   // coarse nodes a prolongation tile stages
// tile table, kMT ints per work-group: fine level first node, W, H, coarse
// level first node, Wc, Hc, tile origin x, y (restriction: coarse nodes;
// prolongation: fine nodes), segment; W = 0 marks padding work-groups
constexpr int kMT = 9;

// restriction tile in LDS: the staged residual and smoother weights, the
// stencils (9 floats each) and constraint flags of the nodes whose residual
// it restricts
constexpr int kRIn = (2 * kRX + 1) * (2 * kRY + 1);
constexpr int kRCf = kRIn * 9;
constexpr int kRCm = (kRIn + 15) / 16 * 16;
inline size_t mg_restrict_lds(int lpr) {
    return (size_t)kRW * kRH * (lpr * sizeof(float4) + sizeof(float)) + kRCf * sizeof(float) + kRCm;
}
// prolongation tile in LDS: the staged fine and coarse vectors, the inner
// nodes' stencils (9 floats each) and the staged nodes' constraint flags
constexpr int kPCf = kPX * kPY * 9;
constexpr int kPCm = (kPW * kPH + 15) / 16 * 16;
inline size_t mg_prolong_lds(int lpr) {
    const size_t tile = (size_t)(kPW * kPH + kCW * kCH) * lpr * sizeof(float4) + kPCf * sizeof(float) + kPCm;
    const size_t red = (size_t)grid_block(lpr) * 4 * sizeof(double);
    return tile > red ? tile : red;
}

// weight of coarse index X in fine index x of an n-node line (bilinear P,
// constant continuation at an odd last node)
__device__ inline float mg_w(int x, int X, int n) { return mg_axis_w(x, X, n); }



// The staging loops below issue all of a thread's loads of a phase before
// using any of them: out-of-grid nodes and idle lanes load from node 0 /
// column 0 (always valid) and are masked afterwards, so no load sits behind a
// branch and the memory latencies of a phase overlap.

// r_{l+1} = P^T (r - A (S r)) on one tile of coarse nodes
// work-group skip of the V-cycle kernels inside an iteration: none of the
// work-group's columns of its problem is still iterating (scale 0: converged,
// zero or padding columns of a narrower problem of a batch; frozen: at the
// target, its z unused until the next replacement recomputes it); S == NULL
// never skips
// (`lane`: the lane's own columns are live, as in block_idle)
__device__ inline bool mg_cols_done(const double* __restrict__ S, int ks, int so, int col0, bool ok, bool& lane) {
    lane = ok;
    if (!S) return false;
    bool live = false;
    if (ok) {
        const double* sc = S + S_SCALE * ks + so + col0;
        const double* fz = S + S_FROZEN * ks + so + col0;
        live = (sc[0] > 0.0 && !(fz[0] > 0.0)) || (sc[1] > 0.0 && !(fz[1] > 0.0)) ||
               (sc[2] > 0.0 && !(fz[2] > 0.0)) || (sc[3] > 0.0 && !(fz[3] > 0.0));
    }
    lane = live;
    return !__syncthreads_or(live);
}

template <int LPR, class TR>
__global__ __launch_bounds__(GB<LPR>::v) void k_mg_restrict(const int* __restrict__ tab, int kp, const float* __restrict__ coef,
                                                    const float* __restrict__ dsm,
                                                    const unsigned char* __restrict__ cmask,
                                                    const TR* __restrict__ r, float* __restrict__ rc,
                                                    const double* __restrict__ S, int ks) {
    extern __shared__ float4 sm[];
    float* sd = reinterpret_cast<float*>(sm + kRW * kRH * LPR);
    float* cs = sd + kRW * kRH;                                          // stencils of the kIn nodes
    unsigned char* cmk = reinterpret_cast<unsigned char*>(cs + kRCf);   // their constraint flags
    constexpr int ngrp = GB<LPR>::v / LPR;
    constexpr int kIW = 2 * kRX + 1, kIn = kRIn;   // fine nodes whose residual is restricted
    constexpr int kPer = (kIn + ngrp - 1) / ngrp;
    constexpr int kSt = (kRW * kRH + ngrp - 1) / ngrp;
    constexpr int kCf = (kRCf + GB<LPR>::v - 1) / GB<LPR>::v;   // stencil floats per thread
    constexpr int kCm = (kIn + GB<LPR>::v - 1) / GB<LPR>::v;    // flags per thread
    const int lane = threadIdx.x % LPR, grp = threadIdx.x / LPR;
    const int* d = tab + kMT * xcd_block(blockIdx.x, gridDim.x);
    const int fbase = d[0], W = d[1], H = d[2], cbase = d[3], Wc = d[4], Hc = d[5], cx0 = d[6], cy0 = d[7];
    if (W == 0) return;
    const int col0 = (blockIdx.y * LPR + lane) * 4;   // column tile blockIdx.y of 4 * LPR columns
    bool ok = col0 < kp;
    if (mg_cols_done(S, ks, d[8] * kp, col0, ok, ok)) return;
    const int lc = ok ? col0 : 0;
    const int fx0 = 2 * cx0 - 2, fy0 = 2 * cy0 - 2;
    {
        // the stencils (contiguous rows of kIW nodes, 9 coefficients each) and
        // constraint flags of the residual nodes load with the residual, so
        // the residual phase waits on LDS only
        float cv[kCf];
        unsigned char cmv[kCm];
#pragma unroll
        for (int u = 0; u < kCf; ++u) {
            const int i = threadIdx.x + u * GB<LPR>::v, ii = i / 9;
            const int gx = fx0 + ii % kIW + 1, gy = fy0 + ii / kIW + 1;
            const bool in = i < kRCf && gx >= 0 && gx < W && gy >= 0 && gy < H;
            cv[u] = coef[in ? (size_t)(fbase + gy * W + gx) * 9 + i % 9 : 0];
        }
#pragma unroll
        for (int u = 0; u < kCm; ++u) {
            const int i = threadIdx.x + u * GB<LPR>::v;
            const int gx = fx0 + i % kIW + 1, gy = fy0 + i / kIW + 1;
            const bool in = cmask && i < kIn && gx >= 0 && gx < W && gy >= 0 && gy < H;
            cmv[u] = in ? cmask[fbase + gy * W + gx] : 0;
        }
        float4 v[kSt];
        float w[kSt];
        bool in[kSt];
#pragma unroll
        for (int u = 0; u < kSt; ++u) {
            const int h = grp + u * ngrp;
            const int gx = fx0 + h % kRW, gy = fy0 + h / kRW;
            in[u] = h < kRW * kRH && gx >= 0 && gx < W && gy >= 0 && gy < H;
            const int nd = in[u] ? fbase + gy * W + gx : 0;
            v[u] = ldf4(r + (size_t)nd * kp + lc);
            w[u] = dsm[nd];
        }
#pragma unroll
        for (int u = 0; u < kSt; ++u) {
            const int h = grp + u * ngrp;
            if (h < kRW * kRH) {
                sm[h * LPR + lane] = (in[u] && ok) ? v[u] : zero4();
                if (lane == 0) sd[h] = in[u] ? w[u] : 0.f;
            }
        }
#pragma unroll
        for (int u = 0; u < kCf; ++u)
            if (threadIdx.x + u * GB<LPR>::v < kRCf) cs[threadIdx.x + u * GB<LPR>::v] = cv[u];
#pragma unroll
        for (int u = 0; u < kCm; ++u)
            if (threadIdx.x + u * GB<LPR>::v < kIn) cmk[threadIdx.x + u * GB<LPR>::v] = cmv[u];
    }
    __syncthreads();
    float4 res[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int i = grp + u * ngrp;
        const int hx = i % kIW + 1, hy = i / kIW + 1;
        const int gx = fx0 + hx, gy = fy0 + hy;
        float4 q = zero4();
        if (i < kIn && gx >= 0 && gx < W && gy >= 0 && gy < H && !cmk[i]) {
            const float* c = cs + i * 9;
            q = sm[(hy * kRW + hx) * LPR + lane];
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                const int h = (hy + k / 3 - 1) * kRW + hx + k % 3 - 1;
                fma4(-c[k] * sd[h], sm[h * LPR + lane], q);
            }
        }
        res[u] = q;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int i = grp + u * ngrp;
        if (i < kIn) sm[((i / kIW + 1) * kRW + i % kIW + 1) * LPR + lane] = res[u];
    }
    __syncthreads();
    if (!ok) return;
    for (int o = grp; o < kRX * kRY; o += ngrp) {
        const int X = cx0 + o % kRX, Y = cy0 + o / kRX;
        if (X >= Wc || Y >= Hc) continue;
        float4 acc4 = zero4();
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy) {
            const int fy = 2 * Y + dy;
            if (fy < 0 || fy >= H) continue;
            const float wy = mg_w(fy, Y, H);
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                const int fx = 2 * X + dx;
                if (fx < 0 || fx >= W) continue;
                const float w = wy * mg_w(fx, X, W);
                if (w != 0.f) fma4(w, sm[((fy - fy0) * kRW + (fx - fx0)) * LPR + lane], acc4);
            }
        }
        *reinterpret_cast<float4*>(rc + (size_t)(cbase + Y * Wc + X) * kp + col0) = acc4;
    }
}

// e = e0 + S (r - A e0), e0 = S r + P e_{l+1}, on one tile of fine nodes. TOP
// (level 0, FP64 residual read as FP32): P acts on free nodes only, e is zero
// on constrained nodes, and the column sums r.e (and r.r) are added in FP64:
// mode 0 adds r.e into the iteration's accumulator slot, mode 1 writes
// per-tile partial rows of r.e and r.r (after a residual replacement; padding
// tiles write zeros)
template <int LPR, bool TOP, class TR>
__global__ __launch_bounds__(GB<LPR>::v) void k_mg_prolong(const int* __restrict__ tab, int kp, int ks, int it, int mode,
                                                   const float* __restrict__ coef, const float* __restrict__ dsm,
                                                   const unsigned char* __restrict__ cmask,
                                                   const TR* __restrict__ r, const float* __restrict__ ec,
                                                   float* __restrict__ e, double* __restrict__ acc,
                                                   double* __restrict__ part0, double* __restrict__ part1,
                                                   const double* __restrict__ S) {
    extern __shared__ float4 sm[];
    float4* sc = sm + kPW * kPH * LPR;
    float* cs = reinterpret_cast<float*>(sc + kCW * kCH * LPR);    // inner nodes' stencils
    unsigned char* cmk = reinterpret_cast<unsigned char*>(cs + kPCf);   // staged nodes' flags (level 0)
    constexpr int ngrp = GB<LPR>::v / LPR;
    constexpr int kCs = (kCW * kCH + ngrp - 1) / ngrp;   // staged coarse nodes per lane group
    constexpr int kSt = (kPW * kPH + ngrp - 1) / ngrp;   // staged fine nodes per lane group
    constexpr int kCf = (kPCf + GB<LPR>::v - 1) / GB<LPR>::v;        // stencil floats per thread
    constexpr int kCm = (kPW * kPH + GB<LPR>::v - 1) / GB<LPR>::v;   // flags per thread
    Tile t;
    t.lane = threadIdx.x % LPR;
    t.grp = threadIdx.x / LPR;
    t.ngrp = ngrp;
    t.col0 = (blockIdx.y * LPR + t.lane) * 4;
    t.bid = xcd_block(blockIdx.x, gridDim.x);
    t.r0 = t.r1 = 0;
    const int* d = tab + kMT * t.bid;
    const int fbase = d[0], W = d[1], H = d[2], cbase = d[3], Wc = d[4], Hc = d[5], tx0 = d[6], ty0 = d[7];
    t.so = d[8] * kp;
    // the p.Ap slot of the next iteration (the update kernel of this one has
    // read the slot's previous contents, the direction kernel of the next
    // accumulates into it)
    if (TOP && mode == 0) clear_slots(acc, A_PQ + ((it + 1) & 1), 1, ks);
    if (W == 0 && !(TOP && mode == 1)) return;
    bool ok = t.col0 < kp;
    if (mg_cols_done(S, ks, t.so, t.col0, ok, ok)) return;   // (S == NULL for residual replacements)
    const int lc = ok ? t.col0 : 0;
    V4<double> rz = {0.0, 0.0, 0.0, 0.0}, rr = {0.0, 0.0, 0.0, 0.0};
    if (W > 0) {
        const int cx0 = tx0 / 2 - 1, cy0 = ty0 / 2 - 1;
        // the inner nodes' stencils (kPY rows of kPX nodes, 9 coefficients per
        // node: contiguous rows, coalesced) and the staged nodes' constraint
        // flags are loaded with the vectors, so the later phases wait on LDS only
        float cv[kCf];
        unsigned char cmv[kCm];
#pragma unroll
        for (int u = 0; u < kCf; ++u) {
            const int i = threadIdx.x + u * GB<LPR>::v;
            const int row = i / (kPX * 9), off = i % (kPX * 9);
            const int gx = tx0 + off / 9, gy = ty0 + row;
            const bool in = i < kPCf && gx < W && gy < H;
            cv[u] = coef[in ? (size_t)(fbase + gy * W + gx) * 9 + off % 9 : 0];
        }
#pragma unroll
        for (int u = 0; u < kCm; ++u) {
            const int h = threadIdx.x + u * GB<LPR>::v;
            const int gx = tx0 - 1 + h % kPW, gy = ty0 - 1 + h / kPW;
            const bool in = TOP && h < kPW * kPH && gx >= 0 && gx < W && gy >= 0 && gy < H;
            cmv[u] = in ? cmask[fbase + gy * W + gx] : 0;
        }
        {
            float4 v[kCs];
            bool in[kCs];
#pragma unroll
            for (int u = 0; u < kCs; ++u) {
                const int h = t.grp + u * ngrp;
                const int X = cx0 + h % kCW, Y = cy0 + h / kCW;
                in[u] = h < kCW * kCH && X >= 0 && X < Wc && Y >= 0 && Y < Hc;
                v[u] = ldf4(ec + (size_t)(in[u] ? cbase + Y * Wc + X : 0) * kp + lc);
            }
#pragma unroll
            for (int u = 0; u < kCs; ++u) {
                const int h = t.grp + u * ngrp;
                if (h < kCW * kCH) sc[h * LPR + t.lane] = (in[u] && ok) ? v[u] : zero4();
            }
        }
        // phase 1 and 2 walk the same staged nodes (h = grp + u * ngrp): each
        // lane group keeps its nodes' residuals in registers for the smoothing
        // step instead of fetching them again
        float4 v[kSt];
        float w[kSt];
        bool in[kSt];
#pragma unroll
        for (int u = 0; u < kSt; ++u) {
            const int h = t.grp + u * ngrp;
            const int gx = tx0 - 1 + h % kPW, gy = ty0 - 1 + h / kPW;
            const int nd = fbase + gy * W + gx;
            in[u] = h < kPW * kPH && gx >= 0 && gx < W && gy >= 0 && gy < H;
            v[u] = ldf4(r + (size_t)(in[u] ? nd : 0) * kp + lc);
            w[u] = dsm[in[u] ? nd : 0];
        }
#pragma unroll
        for (int u = 0; u < kCf; ++u)
            if (threadIdx.x + u * GB<LPR>::v < kPCf) cs[threadIdx.x + u * GB<LPR>::v] = cv[u];
#pragma unroll
        for (int u = 0; u < kCm; ++u)
            if (threadIdx.x + u * GB<LPR>::v < kPW * kPH) cmk[threadIdx.x + u * GB<LPR>::v] = cmv[u];
        __syncthreads();   // the coarse values, stencils and flags are staged
#pragma unroll
        for (int u = 0; u < kSt; ++u) {
            const int h = t.grp + u * ngrp;
            if (h >= kPW * kPH) continue;
            const int gx = tx0 - 1 + h % kPW, gy = ty0 - 1 + h / kPW;
            float4 x = zero4();
            if (in[u] && ok && !(TOP && cmk[h])) {
                x = make_float4(w[u] * v[u].x, w[u] * v[u].y, w[u] * v[u].z, w[u] * v[u].w);
                int ax, ay;   // first coarse parents
                float fx, fy;
                const int nx = mg_axis_parents(gx, W, ax, fx);
                const int ny = mg_axis_parents(gy, H, ay, fy);
                for (int a = 0; a < ny; ++a)
                    for (int b = 0; b < nx; ++b)
                        fma4(fx * fy, sc[((ay + a - cy0) * kCW + (ax + b - cx0)) * LPR + t.lane], x);
            }
            sm[h * LPR + t.lane] = x;
        }
        __syncthreads();
        // smoothing on the tile's own nodes (the inner part of the staged region)
#pragma unroll
        for (int u = 0; u < kSt; ++u) {
            const int h = t.grp + u * ngrp;
            const int hx = h % kPW, hy = h / kPW;
            if (!(in[u] && ok && h < kPW * kPH && hx >= 1 && hx <= kPX && hy >= 1 && hy <= kPY)) continue;
            const int ox = hx - 1, oy = hy - 1;
            const int nd = fbase + (ty0 + oy) * W + tx0 + ox;
            const float* c = cs + (oy * kPX + ox) * 9;
            float4 q = v[u];
#pragma unroll
            for (int k = 0; k < 9; ++k)
                fma4(-c[k], sm[((oy + k / 3) * kPW + ox + k % 3) * LPR + t.lane], q);
            float4 z = sm[h * LPR + t.lane];
            fma4(w[u], q, z);
            if (TOP && cmk[h]) z = zero4();
            *reinterpret_cast<float4*>(e + (size_t)nd * kp + t.col0) = z;
            if constexpr (TOP) {
                const float4 a = v[u];
                rz.x += (double)a.x * z.x; rz.y += (double)a.y * z.y; rz.z += (double)a.z * z.z;
                rz.w += (double)a.w * z.w;
                rr.x += (double)a.x * a.x; rr.y += (double)a.y * a.y; rr.z += (double)a.z * a.z;
                rr.w += (double)a.w * a.w;
            }
        }
    }
    if constexpr (TOP) {
        __syncthreads();   // the staging area becomes the reduction buffer
        double* red = reinterpret_cast<double*>(sm);
        if (mode == 0) {
            block_accumulate<LPR>(t, kp, rz, red, acc + (A_RZ + it % 3) * ks + t.so);
        } else {
            block_partial<LPR>(t, kp, rz, red, part0);
            __syncthreads();
            block_partial<LPR>(t, kp, rr, red, part1);
        }
    }
}

// The coarse end of the V-cycle in one launch: from level lt (the first level
// below which the rest of every grid's hierarchy fits in one work-group's LDS)
// down to the coarsest and back, four columns of one grid per work-group
// (float4 per node, so the index arithmetic and the stencil loads serve four
// columns). The level-lt residual comes from the restriction launch above;
// the restricted residuals and corrections of the lower levels live in LDS
// only; the level-lt correction goes back to the level buffer for the
// prolongation launch below. The operations are those of the per-level
// launches: down, t = S r, r - A t and its restriction P^T; on the coarsest
// level the dense inverse; up, e0 = S r + P e_{l+1} and e = e0 + S (r - A e0).
// Each thread issues all loads of a level phase (its nodes' smoother weights
// and coefficient-major stencils) at once. The coarsest product reads the
// inverse column by column (coalesced over the rows the threads own): the
// inverse of the symmetric coarsest operator, equal to its transpose up to
// FP32 rounding. Work-groups of one 128-byte line of columns run on one XCD
// (xcd_block over the column groups), so their stores merge in that XCD's L2.
constexpr int kTailB = 1024;
constexpr int kTailVec = 4096;                           // LDS of a work-group in float4: 64 KB
constexpr int kTailRed = kTailB;                         // partial sums of the coarsest product
constexpr int kTailPer = 2;                              // nodes per thread: levels of at most 2,048 nodes
constexpr int kTailLev = 8;
// table per grid: levels nl, offset of the coarsest inverse, LDS offset (in
// float4) of the partial sums, 0; per level u = 0 (lt) .. nl - 1 (coarsest):
// W, H, LDS offset of r (e follows it), the level's node count (the row
// stride of the coefficient-major stencil), then as 64-bit pairs the grid's
// first coefficient in that stencil, its smoother weights, and the float
// offsets of its r and e (column 0) in the level buffer
constexpr int kTLev = 12;
constexpr int kTT = 4 + kTLev * kTailLev;

__device__ inline unsigned long long tab_u64(const int* t) {
    return ((unsigned long long)(unsigned)t[1] << 32) | (unsigned)t[0];
}

// the smoother weights and stencils of the thread's nodes i = tid + k kTailB
// of a level of n nodes (coefficient-major stencil, row stride ld; nodes past
// n load node 0 and are not used)
__device__ inline void tail_load(const float* __restrict__ cf, int ld, const float* __restrict__ ds, int n,
                                 float (&a)[kTailPer][9], float (&w)[kTailPer]) {
#pragma unroll
    for (int k = 0; k < kTailPer; ++k) {
        const int i = threadIdx.x + k * kTailB;
        const int j = i < n ? i : 0;
        w[k] = ds[j];
#pragma unroll
        for (int m = 0; m < 9; ++m) a[k][m] = cf[(size_t)m * ld + j];
    }
}

// q_k = R_i - sum_j a_ij X_j for the thread's nodes of a W x H level
// (out-of-grid neighbours skipped)
__device__ inline void tail_residual(const float4* R, const float4* X, const float (&a)[kTailPer][9], int W, int H,
                                     float4 (&q)[kTailPer]) {
    const int n = W * H;
#pragma unroll
    for (int k = 0; k < kTailPer; ++k) {
        const int i = threadIdx.x + k * kTailB;
        float4 s = zero4();
        if (i < n) {
            const int x = i % W, y = i / W;
            s = R[i];
#pragma unroll
            for (int m = 0; m < 9; ++m) {
                const int nx = x + m % 3 - 1, ny = y + m / 3 - 1;
                if (nx >= 0 && nx < W && ny >= 0 && ny < H) fma4(-a[k][m], X[ny * W + nx], s);
            }
        }
        q[k] = s;
    }
}

__global__ __launch_bounds__(kTailB) void k_mg_tail(const int* __restrict__ tab, int kp, int ks,
                                                     const float* __restrict__ cinv, float* __restrict__ mgv,
                                                     const double* __restrict__ S) {
    extern __shared__ float4 sm[];
    const int c = 4 * xcd_block(blockIdx.x, gridDim.x), q = blockIdx.y, tid = threadIdx.x;
    if (c >= kp) return;
    bool lane_live;
    if (mg_cols_done(S, ks, q * kp, c, true, lane_live)) return;   // (S == NULL for residual replacements)
    const int* d = tab + kTT * q;
    const int nl = d[0];
    {   // the top level's residual
        const int* e = d + 4;
        const int n = e[0] * e[1];
        const float* rg = mgv + tab_u64(e + 8) + c;
        float4* R = sm + e[2];
        float4 v[kTailPer];
#pragma unroll
        for (int k = 0; k < kTailPer; ++k) {
            const int i = tid + k * kTailB;
            v[k] = ldf4(rg + (size_t)(i < n ? i : 0) * kp);
        }
#pragma unroll
        for (int k = 0; k < kTailPer; ++k)
            if (tid + k * kTailB < n) R[tid + k * kTailB] = v[k];
    }
    __syncthreads();
    for (int u = 0; u + 1 < nl; ++u) {
        const int* e = d + 4 + kTLev * u;
        const int W = e[0], H = e[1], n = W * H;
        float4* R = sm + e[2];
        float4* E = R + n;
        float a[kTailPer][9], w[kTailPer];
        float4 t[kTailPer];
        tail_load(reinterpret_cast<const float*>(tab_u64(e + 4)), e[3], reinterpret_cast<const float*>(tab_u64(e + 6)),
                  n, a, w);
#pragma unroll
        for (int k = 0; k < kTailPer; ++k) {
            const int i = tid + k * kTailB;
            if (i < n) {
                const float4 r = R[i];
                E[i] = make_float4(w[k] * r.x, w[k] * r.y, w[k] * r.z, w[k] * r.w);
            }
        }
        __syncthreads();
        tail_residual(R, E, a, W, H, t);   // r - A S r
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kTailPer; ++k) {
            const int i = tid + k * kTailB;
            if (i < n) E[i] = t[k];
        }
        __syncthreads();
        const int* ec = e + kTLev;
        const int Wc = ec[0], Hc = ec[1];
        float4* Rc = sm + ec[2];
        for (int I = tid; I < Wc * Hc; I += kTailB) {
            const int X = I % Wc, Y = I / Wc;
            float4 acc = zero4();
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy) {
                const int fy = 2 * Y + dy;
                if (fy < 0 || fy >= H) continue;
                const float wy = mg_axis_w(fy, Y, H);
#pragma unroll
                for (int dx = -1; dx <= 1; ++dx) {
                    const int fx = 2 * X + dx;
                    if (fx < 0 || fx >= W) continue;
                    const float wt = wy * mg_axis_w(fx, X, W);
                    if (wt != 0.f) fma4(wt, E[fy * W + fx], acc);
                }
            }
            Rc[I] = acc;
        }
        __syncthreads();
    }
    {   // coarsest level: e = C^-1 r; g = kTailB / n partial sums per row, added in a fixed order
        const int* e = d + 4 + kTLev * (nl - 1);
        const int n = e[0] * e[1];
        const float4* R = sm + e[2];
        float4* red = sm + d[2];
        const float* ci = cinv + d[1];
        const int ng = kTailB / n, g = tid / n, i = tid % n;
        if (g < ng) {
            float4 s = zero4();
#pragma unroll 4
            for (int j = g; j < n; j += ng) fma4(ci[(size_t)j * n + i], R[j], s);
            red[tid] = s;
        }
        __syncthreads();
        if (tid < n) {
            float4 s = zero4();
            for (int h = 0; h < ng; ++h) {
                const float4 v = red[h * n + tid];
                s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
            }
            if (nl == 1)
                *reinterpret_cast<float4*>(mgv + tab_u64(e + 10) + (size_t)tid * kp + c) = s;
            else
                sm[e[2] + n + tid] = s;
        }
        __syncthreads();
    }
    for (int u = nl - 2; u >= 0; --u) {
        const int* e = d + 4 + kTLev * u;
        const int W = e[0], H = e[1], n = W * H;
        float4* R = sm + e[2];
        float4* E = R + n;
        const int* ec = e + kTLev;
        const int Wc = ec[0];
        const float4* Ec = sm + ec[2] + Wc * ec[1];
        float a[kTailPer][9], w[kTailPer];
        float4 z[kTailPer];
        tail_load(reinterpret_cast<const float*>(tab_u64(e + 4)), e[3], reinterpret_cast<const float*>(tab_u64(e + 6)),
                  n, a, w);
#pragma unroll
        for (int k = 0; k < kTailPer; ++k) {   // e0 = S r + P e_{l+1}
            const int i = tid + k * kTailB;
            if (i >= n) continue;
            const int x = i % W, y = i / W;
            const float4 r = R[i];
            float4 v = make_float4(w[k] * r.x, w[k] * r.y, w[k] * r.z, w[k] * r.w);
            int ax, ay;
            float fx, fy;
            const int nx = mg_axis_parents(x, W, ax, fx);
            const int ny = mg_axis_parents(y, H, ay, fy);
            for (int pa = 0; pa < ny; ++pa)
                for (int pb = 0; pb < nx; ++pb) fma4(fx * fy, Ec[(ay + pa) * Wc + ax + pb], v);
            E[i] = v;
        }
        __syncthreads();
        tail_residual(R, E, a, W, H, z);
#pragma unroll
        for (int k = 0; k < kTailPer; ++k) {
            const int i = tid + k * kTailB;
            if (i < n) {
                float4 y = E[i];
                fma4(w[k], z[k], y);
                z[k] = y;
            }
        }
        if (u == 0) {
            float* eg = mgv + tab_u64(e + 10) + c;
#pragma unroll
            for (int k = 0; k < kTailPer; ++k) {
                const int i = tid + k * kTailB;
                if (i < n) *reinterpret_cast<float4*>(eg + (size_t)i * kp) = z[k];
            }
        } else {
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kTailPer; ++k) {
                const int i = tid + k * kTailB;
                if (i < n) E[i] = z[k];
            }
            __syncthreads();
        }
    }
}

// residual replacement with the multigrid preconditioner: r *= s' in place
// (FP64) and, when rf is given, its FP32 copy, p *= s'/s; the row's problem
// (srow: nseg + 1 row bounds) selects the scalars
// the segment of a row: the last seg with srow[seg] <= row
__device__ inline int row_segment(int row, int nseg, const int* __restrict__ srow) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (srow[mid] <= row) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// The element-wise kernels below take four consecutive columns of one row
// per thread (kp is a multiple of 4): one segment search per four entries and
// 16-byte accesses. total4 = N kp / 4.
// r = src * scale (src: r itself, or b at the first replacement of a solve);
// p *= ratio, and p = 0 where the ratio is 0 (a new start)
__global__ void k_mg_rescale(long long total4, int kp, int nseg, const int* __restrict__ srow,
                             const double* __restrict__ scale, const double* __restrict__ ratio,
                             const double* src, double* r, float* __restrict__ rf, float* __restrict__ p) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total4) return;
    const long long i = 4 * t;
    const int row = (int)(i / kp);
    const size_t c = (size_t)row_segment(row, nseg, srow) * kp + i % kp;
    V4<double> v = ld4(src + i);
    const V4<double> sc = ld4(scale + c), ra = ld4(ratio + c);
    v.x *= sc.x; v.y *= sc.y; v.z *= sc.z; v.w *= sc.w;
    st4(r + i, v);
    if (rf) *reinterpret_cast<float4*>(rf + i) = make_float4((float)v.x, (float)v.y, (float)v.z, (float)v.w);
    float4 pv = *reinterpret_cast<const float4*>(p + i);
    pv.x = ra.x != 0.0 ? pv.x * (float)ra.x : 0.f;
    pv.y = ra.y != 0.0 ? pv.y * (float)ra.y : 0.f;
    pv.z = ra.z != 0.0 ? pv.z * (float)ra.z : 0.f;
    pv.w = ra.w != 0.0 ? pv.w * (float)ra.w : 0.f;
    *reinterpret_cast<float4*>(p + i) = pv;
}

// x += alpha_j p_j / s on every node and column: the last step of a
// multigrid cycle that ended after an even iteration j (the update kernel adds
// the steps of x in pairs, on odd iterations); alpha_j from its accumulator
// slots as alpha_before forms it
__global__ void k_x_flush_mg(long long total4, int kp, int nseg, const int* __restrict__ srow, int j, int ks,
                             const double* __restrict__ S, const double* __restrict__ acc,
                             const float* __restrict__ p, double* __restrict__ x) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total4) return;
    const long long i0 = 4 * t;
    const int row = (int)(i0 / kp);
    const int c0 = row_segment(row, nseg, srow) * kp + (int)(i0 % kp);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int c = c0 + u;
        const double sc = S[S_SCALE * ks + c];
        const double a = alpha_before(j + 1, ks, c, S, acc);
        if (sc > 0.0 && a != 0.0) x[i0 + u] += (a / sc) * p[i0 + u];
    }
}

// ------------------------------------------------ vector DOFs: 3 x 3 blocks
// Interleaved vector DOFs (3 per node) as a node-level block operator: for node
// pair (i, j) one column index and 9 FP32 values, a(row comp, col comp). A lane
// group owns a node and forms its 3 rows at once, so each neighbour node's 3
// DOFs are gathered once instead of once per row.
__global__ void k_pack_bsr3(int nnodes, const int* __restrict__ rpn, const float* __restrict__ va32,
                            float* __restrict__ bval) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnodes) return;
    const int b = rpn[i], len = rpn[i + 1] - b;
    for (int q = 0; q < len; ++q)
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) bval[(size_t)(b + q) * 9 + a * 3 + c] = va32[(size_t)9 * b + a * 3 * len + 3 * q + c];
}

struct NodeTile {
    int lane, grp, ngrp, col0, n0, n1;
};

template <int LPR>
__device__ inline NodeTile node_tile(int nnodes, int npb) {
    NodeTile t;
    t.lane = threadIdx.x % LPR;
    t.grp = threadIdx.x / LPR;
    t.ngrp = blockDim.x / LPR;
    t.col0 = (blockIdx.y * LPR + t.lane) * 4;
    const int bid = xcd_block(blockIdx.x, gridDim.x);
    t.n0 = bid * npb;
    t.n1 = min(nnodes, t.n0 + npb);
    return t;
}

template <int LPR>
__device__ inline Tile as_row_tile(const NodeTile& n) {
    Tile t;
    t.lane = n.lane; t.grp = n.grp; t.ngrp = n.ngrp; t.col0 = n.col0;
    t.bid = xcd_block(blockIdx.x, gridDim.x);
    t.r0 = 3 * n.n0; t.r1 = 3 * n.n1; t.so = 0;
    return t;
}

__device__ inline void fma4(V4<float>& q, float a, const V4<float>& v) {
    q.x += a * v.x; q.y += a * v.y; q.z += a * v.z; q.w += a * v.w;
}

// Direction update as a streaming pass: p_new = m z + b p_old on every row
// (the 27-node neighbourhoods of 3D meshes outgrow the L2, so gathering both
// z and p_old at every neighbour is replaced by one gathered vector below).
template <int LPR>
__global__ __launch_bounds__(kB) void k_pdir_stream(int kp, int nrow, int rpb, int it, double* __restrict__ S,
                                                    const float* __restrict__ z, const float* __restrict__ pold,
                                                    float* __restrict__ pnew, double* __restrict__ acc) {
    Tile t;
    t.lane = threadIdx.x % LPR;
    t.grp = threadIdx.x / LPR;
    t.ngrp = blockDim.x / LPR;
    t.col0 = (blockIdx.y * LPR + t.lane) * 4;
    t.bid = xcd_block(blockIdx.x, gridDim.x);
    t.r0 = t.bid * rpb;
    t.r1 = min(nrow, t.r0 + rpb);
    t.so = 0;
    clear_slots(acc, A_RZ + it % 3, 1, kp);
    clear_slots(acc, A_RR + it % 3, 1, kp);
    if (block_done(t, kp, kp, S)) return;
    __shared__ float cb[4 * LPR], cm[4 * LPR];
    shared_dir_coef<float, LPR>(it, kp, kp, 0, S, acc, cb, cm);
    if (t.col0 >= kp) return;
    const int lc = t.lane * 4;
    const V4<float> b = {cb[lc], cb[lc + 1], cb[lc + 2], cb[lc + 3]};
    const V4<float> m = {cm[lc], cm[lc + 1], cm[lc + 2], cm[lc + 3]};
    for (int r = t.r0 + t.grp; r < t.r1; r += t.ngrp) {
        const size_t o = (size_t)r * kp + t.col0;
        V4<float> v;
        pdir4(ld4(z + o), m, b, ld4(pold + o), v);
        st4(pnew + o, v);
    }
}

// q = A p (3 x 3 blocks, one gathered vector), acc += p . q
template <int LPR>
__global__ __launch_bounds__(kB) void k_spmv_pq_b3(int kp, int nnodes, int npb, int it, const int* __restrict__ rpn,
                                                   const int* __restrict__ cin, const float* __restrict__ bval,
                                                   const double* __restrict__ S, const float* __restrict__ p,
                                                   double* __restrict__ acc) {
    __shared__ double red[kB * 4];
    const NodeTile n = node_tile<LPR>(nnodes, npb);
    const Tile t = as_row_tile<LPR>(n);
    if (block_done(t, kp, kp, S)) return;
    V4<float> pq = {0.f, 0.f, 0.f, 0.f};
    if (n.col0 < kp) {
        for (int i0 = n.n0 + n.grp; i0 < n.n1; i0 += n.ngrp) {
            const int i = uniform_row<LPR>(i0);
            V4<float> q[3] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
            for (int e = rpn[i]; e < rpn[i + 1]; ++e) {
                const int j = cin[e];
                const float* A = bval + (size_t)e * 9;
                V4<float> pj[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) pj[c] = ld4(p + (size_t)(3 * j + c) * kp + n.col0);
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    fma4(q[a], A[a * 3], pj[0]);
                    fma4(q[a], A[a * 3 + 1], pj[1]);
                    fma4(q[a], A[a * 3 + 2], pj[2]);
                }
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const V4<float> pi = ld4(p + (size_t)(3 * i + a) * kp + n.col0);
                pq.x += pi.x * q[a].x; pq.y += pi.y * q[a].y; pq.z += pi.z * q[a].z; pq.w += pi.w * q[a].w;
            }
        }
    }
    const V4<double> ad = {(double)pq.x, (double)pq.y, (double)pq.z, (double)pq.w};
    block_accumulate<LPR>(t, kp, ad, red, acc + (A_PQ + (it & 1)) * kp);
}

template <int LPR>
__global__ __launch_bounds__(kB) void k_cg_update_b3(int kp, int nnodes, int npb, int it, const int* __restrict__ rpn,
                                                     const int* __restrict__ cin, const float* __restrict__ bval,
                                                     const float* __restrict__ dinv, const float* __restrict__ diag,
                                                     const double* __restrict__ S, const float* __restrict__ p,
                                                     float* __restrict__ x, float* __restrict__ z,
                                                     double* __restrict__ acc) {
    __shared__ double red[kB * 4];
    const NodeTile n = node_tile<LPR>(nnodes, npb);
    const Tile t = as_row_tile<LPR>(n);
    clear_slots(acc, A_PQ + ((it + 1) & 1), 1, kp);
    if (block_done(t, kp, kp, S)) return;
    __shared__ float ca[4 * LPR];
    shared_alpha<float, LPR>(it, kp, kp, 0, S, acc, ca);
    V4<float> rz = {0.f, 0.f, 0.f, 0.f}, rr = {0.f, 0.f, 0.f, 0.f};
    if (n.col0 < kp) {
        const V4<float> al = {ca[n.lane * 4], ca[n.lane * 4 + 1], ca[n.lane * 4 + 2], ca[n.lane * 4 + 3]};
        for (int i0 = n.n0 + n.grp; i0 < n.n1; i0 += n.ngrp) {
            const int i = uniform_row<LPR>(i0);
            V4<float> q[3] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
            for (int e = rpn[i]; e < rpn[i + 1]; ++e) {
                const int j = cin[e];
                const float* A = bval + (size_t)e * 9;
                V4<float> pj[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) pj[c] = ld4(p + (size_t)(3 * j + c) * kp + n.col0);
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    fma4(q[a], A[a * 3], pj[0]);
                    fma4(q[a], A[a * 3 + 1], pj[1]);
                    fma4(q[a], A[a * 3 + 2], pj[2]);
                }
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int r = 3 * i + a;
                const size_t o = (size_t)r * kp + n.col0;
                const V4<float> pi = ld4(p + o);
                V4<float> xv = ld4(x + o), zv = ld4(z + o);
                const float dd = dinv[r], g = diag[r];
                xv.x += al.x * pi.x; xv.y += al.y * pi.y; xv.z += al.z * pi.z; xv.w += al.w * pi.w;
                zv.x -= al.x * dd * q[a].x; zv.y -= al.y * dd * q[a].y; zv.z -= al.z * dd * q[a].z;
                zv.w -= al.w * dd * q[a].w;
                st4(x + o, xv);
                st4(z + o, zv);
                res_dots(zv, g, rz, rr);
            }
        }
    }
    V4<double> a = {(double)rz.x, (double)rz.y, (double)rz.z, (double)rz.w};
    block_accumulate<LPR>(t, kp, a, red, acc + (A_RZ + it % 3) * kp);
    __syncthreads();
    V4<double> bb = {(double)rr.x, (double)rr.y, (double)rr.z, (double)rr.w};
    block_accumulate<LPR>(t, kp, bb, red, acc + (A_RR + it % 3) * kp);
}

// CSR -> DIA-9. Every entry must couple a grid neighbour of its row's node;
// anything else (err) keeps the solver on the CSR kernels.
__global__ void k_build_dia9(const int* __restrict__ tab, const int* __restrict__ rp, const int* __restrict__ ci,
                             const double* __restrict__ va, float* d32, double* d64, int* err) {
    const int* t = tab + 3 * blockIdx.y;   // grid: first row, W, H (W >= 3)
    const int base = t[0], W = t[1], H = t[2];
    const int loc = blockIdx.x * blockDim.x + threadIdx.x;
    if (loc >= W * H) return;
    const int nd = base + loc;
    const int x = loc % W, y = loc / W;
    double a[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) a[k] = 0.0;
    for (int e = rp[nd]; e < rp[nd + 1]; ++e) {
        // column offset -> (dy, dx) without a division: |dx| <= 1 < W - 1
        const int off = ci[e] - nd;
        const int dy = off < -1 ? -1 : (off > 1 ? 1 : 0);
        const int dx = off - dy * W;
        if (dx < -1 || dx > 1 || x + dx < 0 || x + dx >= W || y + dy < 0 || y + dy >= H) {
            atomicOr(err, 1);
            continue;
        }
        const int k = (dy + 1) * 3 + (dx + 1);
#pragma unroll
        for (int q = 0; q < 9; ++q)
            if (q == k) a[q] = va[e];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        d64[(size_t)nd * 9 + k] = a[k];
        d32[(size_t)nd * 9 + k] = (float)a[k];
    }
}

// FP64 residual r = b - A x (x == NULL: a zero iterate); part0 = r.r
template <int LPR>
__global__ __launch_bounds__(kB) void k_resid64(int kp, Seg G, const int* __restrict__ rp,
                                                const int* __restrict__ ci, const double* __restrict__ va,
                                                const double* __restrict__ x, const double* __restrict__ b,
                                                double* __restrict__ r, double* __restrict__ part) {
    __shared__ double red[kB * 4];
    const Tile t = tile<LPR>(kp, G);
    const bool ok = t.col0 < kp;
    V4<double> acc = {0, 0, 0, 0};
    if (ok) {
        for (int row = t.r0 + t.grp; row < t.r1; row += t.ngrp) {
            V4<double> s = ld4(b + (size_t)row * kp + t.col0);
            const int e1 = x ? rp[row + 1] : 0;   // x == NULL: a zero iterate, r = b
            for (int e = x ? rp[row] : 0; e < e1; ++e) {
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

// The start of a solve in one pass: the eliminated right-hand side b (the
// value k_prepare forms: B on free rows, minus the couplings to constrained
// DOFs on rows flagged in cnb for a non-transposed solve, 0 on constrained
// rows and padding columns) and |b|^2 per column as per-block partial rows, the
// layout k_resid64 gives the initial residual r = b. r itself is not written:
// the first residual replacement scales b into it.
template <int LPR>
__global__ __launch_bounds__(kB) void k_prepare_norm(int k, int kp, Seg G, int trans, const double* __restrict__ B,
                                                     int ldb, const int* __restrict__ rp, const int* __restrict__ ci,
                                                     const double* __restrict__ va, const int* __restrict__ diag,
                                                     const unsigned char* __restrict__ cmask,
                                                     const unsigned char* __restrict__ cnb, double* __restrict__ b64,
                                                     double* __restrict__ part) {
    __shared__ double red[kB * 4];
    const Tile t = tile<LPR>(kp, G);
    const bool ok = t.col0 < kp;
    V4<double> acc = {0, 0, 0, 0};
    if (ok) {
        for (int row = t.r0 + t.grp; row < t.r1; row += t.ngrp) {
            double v[4] = {0.0, 0.0, 0.0, 0.0};
            if (!cmask[row]) {
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (t.col0 + u < k) v[u] = B[(size_t)row * ldb + t.col0 + u];
                if (!trans && cnb && cnb[row]) {
                    for (int e = rp[row]; e < rp[row + 1]; ++e) {
                        const int j = ci[e];
                        if (!cmask[j]) continue;
                        const double a = va[e], dj = va[diag[j]];
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (t.col0 + u < k) v[u] -= a * (B[(size_t)j * ldb + t.col0 + u] / dj);
                    }
                }
            }
            st4(b64 + (size_t)row * kp + t.col0, V4<double>{v[0], v[1], v[2], v[3]});
            acc.x += v[0] * v[0]; acc.y += v[1] * v[1]; acc.z += v[2] * v[2]; acc.w += v[3] * v[3];
        }
    }
    block_partial<LPR>(t, kp, acc, red, part);
}

// FP64 residual r = b - A x on the declared grids (tile table of the iteration
// kernels) from the FP64 stencil, each row's products in the order of its CSR
// entries (ascending columns), so r is the CSR residual's; |r|^2 is added into
// rn2 per segment and column (cleared before the launch). A lane group walks
// a strip of SY nodes down one tile column and keeps the 3 x 3 window of x in
// registers: one new row of three values per node instead of nine.
template <int LPR>
__global__ __launch_bounds__(GB<LPR>::v) void k_resid_grid64(const int* __restrict__ gtab, int kp,
                                                            const double* __restrict__ coef,
                                                            const double* __restrict__ x,
                                                            const double* __restrict__ b, double* __restrict__ r,
                                                            double* __restrict__ rn2) {
    extern __shared__ float4 sm[];
    constexpr int NG = GB<LPR>::v / LPR;
    constexpr int SY = NG <= 32 ? kTY : (NG <= 64 ? kTY / 2 : 1);
    constexpr int NS = kTX * (kTY / SY);
    int d[kGT];
    const Tile t = grid_tile_t<LPR>(kp, gtab, d);
    const int base = d[0], W = d[1], H = d[2], tx0 = d[3], ty0 = d[4];
    if (W == 0) return;
    V4<double> acc = {0.0, 0.0, 0.0, 0.0};
    if (t.col0 < kp) {
        for (int sidx = t.grp; sidx < NS; sidx += t.ngrp) {
            const int gx = tx0 + sidx % kTX, gy0 = ty0 + (sidx / kTX) * SY;
            if (gx >= W || gy0 >= H) continue;
            V4<double> wv[3][3];   // [row y - 1, y, y + 1][column x - 1, x, x + 1]
            auto load_row = [&](int gy, V4<double>* dst) {
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) {
                    const int jx = gx + dx - 1;
                    dst[dx] = (jx >= 0 && jx < W && gy >= 0 && gy < H)
                                  ? ld4(x + ((size_t)base + (size_t)gy * W + jx) * kp + t.col0)
                                  : V4<double>{0.0, 0.0, 0.0, 0.0};
                }
            };
            load_row(gy0 - 1, wv[0]);
            load_row(gy0, wv[1]);
            for (int u = 0; u < SY; ++u) {
                const int gy = gy0 + u;
                if (gy >= H) break;
                load_row(gy + 1, wv[2]);
                const size_t nd = (size_t)base + (size_t)gy * W + gx;
                V4<double> sv = ld4(b + nd * kp + t.col0);
#pragma unroll
                for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                    for (int dx = 0; dx < 3; ++dx) {
                        const int jx = gx + dx - 1, jy = gy + dy - 1;
                        if (jx < 0 || jx >= W || jy < 0 || jy >= H) continue;
                        const double a = coef[nd * 9 + dy * 3 + dx];
                        const V4<double> v = wv[dy][dx];
                        sv.x -= a * v.x; sv.y -= a * v.y; sv.z -= a * v.z; sv.w -= a * v.w;
                    }
                st4(r + nd * kp + t.col0, sv);
                acc.x += sv.x * sv.x; acc.y += sv.y * sv.y; acc.z += sv.z * sv.z; acc.w += sv.w * sv.w;
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) {
                    wv[0][dx] = wv[1][dx];
                    wv[1][dx] = wv[2][dx];
                }
            }
        }
    }
    __syncthreads();
    block_accumulate<LPR>(t, kp, acc, reinterpret_cast<double*>(sm), rn2 + t.so);
}

// x64 += x_inner / scale on active columns (scale > 0); x_inner restarts at 0.
// srow: nseg + 1 segment row bounds (the row's segment selects its scales).
template <class T>
__global__ void k_refine(long long total4, int kp, int nseg, const int* __restrict__ srow,
                         const double* __restrict__ scale, T* __restrict__ xin, double* __restrict__ x64) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total4) return;
    const long long i0 = 4 * t;
    const int row = (int)(i0 / kp);
    const size_t c0 = (size_t)row_segment(row, nseg, srow) * kp + i0 % kp;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const double s = scale[c0 + u];
        if (s > 0.0) x64[i0 + u] += (double)xin[i0 + u] / s;
        xin[i0 + u] = (T)0;
    }
}

// ------------------------------------------------------------ reductions
// Fixed-order sum over the nblk per-block partial rows of one column: 16
// groups of 64 lanes, each group strides over the rows with 8 independent
// loads in flight, then a fixed-order combine of the 16 group sums.
__device__ inline double sum_parts(int b0, int nblk, int kp, const double* __restrict__ part, int col, double* red) {
    constexpr int G = kRedB / 64;
    const int g = threadIdx.x / 64, l = threadIdx.x % 64;
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (col < kp) {
        int b = b0 + g;
        for (; b + 7 * G < nblk; b += 8 * G) {
#pragma unroll
            for (int u = 0; u < 8; ++u) a[u] += part[(size_t)(b + u * G) * kp + col];
        }
        for (; b < nblk; b += G) a[0] += part[(size_t)b * kp + col];
    }
    red[g * 64 + l] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    __syncthreads();
    double tot = 0.0;
    if (g == 0)
        for (int q = 0; q < G; ++q) tot += red[q * 64 + l];
    __syncthreads();
    return tot;
}

// After residual replacement: rz from the new residual; beta against the
// rescaled r.z of the column's last step (jlast = last iteration of the cycle
// that ended, -1 for none; a frozen column's is kept in S_RZPREV) since the
// direction is kept -- unless beta exceeds kBetaRestart: the true residual is
// then far above the recursive one the direction was built for (the FP32
// recurrence ran into its rounding floor) and the cycle restarts from p = z;
// next trigger level; a fresh cycle (no frozen columns, all accumulator slots
// cleared). One column group of one segment per work-group
// (blockIdx.y = segment, whose partial rows are work-groups sb[seg] ..
// sb[seg+1]-1); slot stride ks.
constexpr double kBetaRestart = 4.0;

__global__ __launch_bounds__(kRedB) void k_red_reset(int kp, int ks, const int* __restrict__ sb, const double* part0,
                                                     const double* part1, double* S, double* acc, double delta2,
                                                     int jlast) {
    __shared__ double red[kRedB];
    const int c = blockIdx.x * 64 + threadIdx.x % 64;
    const int b0 = sb[blockIdx.y], b1 = sb[blockIdx.y + 1];
    const double rzn = sum_parts(b0, b1, kp, part0, c, red);
    const double rrn = sum_parts(b0, b1, kp, part1, c, red);
    if (threadIdx.x < 64 && c < kp) {
        const int col = blockIdx.y * kp + c;
        double last = 0.0;
        if (jlast >= 0 && S[S_ACTIVE * ks + col] > 0.0)
            last = S[S_FROZEN * ks + col] > 0.0 ? S[S_RZPREV * ks + col] : rz_at(jlast, ks, col, S, acc);
        const bool act = S[S_SCALE * ks + col] > 0.0 && rrn > 0.0;
        const double q = S[S_RATIO * ks + col];
        const double prev = last * q * q;
        S[S_ACTIVE * ks + col] = act ? 1.0 : 0.0;
        const double beta = (act && prev > 0.0) ? rzn / prev : 0.0;
        S[S_BETA * ks + col] = beta > kBetaRestart ? 0.0 : beta;
        S[S_RZ * ks + col] = act ? rzn : 0.0;
        S[S_RZPREV * ks + col] = 0.0;
        S[S_FROZEN * ks + col] = 0.0;
        S[S_TRIG * ks + col] = fmax(delta2 * rrn, S[S_TGT * ks + col]);
        for (int a = 0; a < A_N; ++a) acc[(size_t)a * ks + col] = 0.0;
    }
}

__global__ __launch_bounds__(kRedB) void k_red_norm(int kp, const int* __restrict__ sb, const double* part,
                                                    double* out) {
    __shared__ double red[kRedB];
    const int col = blockIdx.x * 64 + threadIdx.x % 64;
    const double s = sum_parts(sb[blockIdx.y], sb[blockIdx.y + 1], kp, part, col, red);
    if (threadIdx.x < 64 && col < kp) out[(size_t)blockIdx.y * kp + col] = s;
}

// ---------------------------------------------------- constraint handling
__global__ void k_dinv(int n, const int* diagpos, const double* va, float* d32, double* d64, float* g32,
                       double* g64) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = va[diagpos[i]];
    d64[i] = 1.0 / d;
    d32[i] = (float)(1.0 / d);
    g64[i] = d;
    g32[i] = (float)d;
}

__global__ void k_pack_entries(long long nnz, const int* __restrict__ ci, const float* __restrict__ va32,
                               int2* __restrict__ cv) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nnz) cv[e] = make_int2(ci[e], __float_as_int(va32[e]));
}

// rows (not constrained themselves) with a constrained neighbour, gathered
// row by row (any pattern)
__global__ void k_cnbr(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                       const unsigned char* __restrict__ cmask, unsigned char* __restrict__ cnb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned char f = 0;
    if (!cmask[i])
        for (int e = rp[i]; e < rp[i + 1]; ++e) f |= cmask[ci[e]];
    cnb[i] = f;
}

// the same flags scattered from the constrained rows (cnb cleared first): for a
// structurally symmetric pattern, i couples to a constrained j exactly when j's
// row holds i, and only the few constrained rows walk their entries
__global__ void k_cnbr_scatter(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                               const unsigned char* __restrict__ cmask, unsigned char* __restrict__ cnb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !cmask[i]) return;
    for (int e = rp[i]; e < rp[i + 1]; ++e) {
        const int j = ci[e];
        if (!cmask[j]) cnb[j] = 1;
    }
}

// X from the eliminated solution; see the file comment for the constrained rows.
// A thread takes four consecutive columns of one row (k4 = ceil(k / 4) column
// groups per row): one 32-byte load of x for a free row.
template <class I>
__device__ inline double finalize_constrained(int i, int c, int trans, const double* __restrict__ B, int ldb,
                                              const double* __restrict__ x64, int kp, const int* __restrict__ rp,
                                              const int* __restrict__ ci, const double* __restrict__ va,
                                              const int* __restrict__ diag, const unsigned char* __restrict__ cmask) {
    if (!trans) return B[(size_t)i * ldb + c] / va[diag[i]];
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
    return s / va[diag[i]];
}

template <class I>
__global__ void k_finalize(int n, int k, int kp, int trans, const double* __restrict__ B, int ldb,
                           const double* __restrict__ x64, const int* __restrict__ rp,
                           const int* __restrict__ ci, const double* __restrict__ va,
                           const int* __restrict__ diag, const unsigned char* __restrict__ cmask,
                           double* __restrict__ X, int ldx) {
    const int k4 = (k + 3) / 4;
    const I t = (I)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (I)n * k4) return;
    const int i = (int)(t / (I)k4), c0 = 4 * (int)(t - (I)i * k4);
    double v[4];
    if (!cmask[i]) {
        const V4<double> a = ld4(x64 + (size_t)i * kp + c0);   // c0 + 3 < kp
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    } else {
#pragma unroll
        for (int u = 0; u < 4; ++u)
            v[u] = c0 + u < k ? finalize_constrained<I>(i, c0 + u, trans, B, ldb, x64, kp, rp, ci, va, diag, cmask)
                              : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
        if (c0 + u < k) X[(size_t)i * ldx + c0 + u] = v[u];
}

inline int pow2_at_least(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

struct Plan {
    int n, kp, lpr, ntile, rpb, nblk, nred, nseg, ks;
    // row partition: work-group b owns rows [brow[b], brow[b+1]) of segment
    // bseg[b]; segment q's work-groups are sb[q] .. sb[q+1]-1, its rows
    // srow[q] .. srow[q+1]-1
    std::vector<int> brow, bseg, sb, srow;
    std::vector<int> gtab;   // 2D-grid tile table (kGT ints per work-group)
    // 2D tensor-grid path
    int grid, W, H, L, nsegx, gblk;
    const int* d_gtab = nullptr;
    // 3 x 3 block path (vector DOFs): node ranges of npb nodes per work-group
    int bsr = 0, nnodes = 0, npb = 0, bblk = 0;
    // multigrid V-cycle: tile tables (kMT ints per work-group) of the
    // restriction (coarse tiles of level l+1) and prolongation (tiles of level
    // l) launches for l < mgL, each padded to whole XCD groups; level l >= 1
    // vectors in the coarse-level buffer: r_l at mgv_off[l], e_l after it
    int mgL = 0;
    std::vector<int> mgtab;
    int mg_roff[kMGMaxLev] = {}, mg_rn[kMGMaxLev] = {}, mg_poff[kMGMaxLev] = {}, mg_pn[kMGMaxLev] = {};
    size_t mgv_off[kMGMaxLev + 1] = {}, mgv_total = 0;
    const int* d_mgtab = nullptr;
    // coarse solves (4 ints per grid, see k_mg_coarse) and the level-0
    // prolongation tiles of every grid (nseg + 1 bounds: partial-row ranges
    // of a residual replacement)
    std::vector<int> mgctab, mgsb;
    const int* d_mgctab = nullptr;
    // coarse end of the V-cycle in one launch (k_mg_tail): its first level
    // (0: one restriction and one prolongation launch per level and the
    // coarse solve), its table (kTT ints per grid) and LDS bytes
    int mg_tail = 0;
    size_t mgt_lds = 0;
    std::vector<int> mgttab;
    const int* d_mgttab = nullptr;
};

Plan make_plan(int n, int k, int num_cu, int nseg = 1, const int* seg_rows = nullptr, int min_kp = 4,
               int max_lpr = 64) {
    Plan p;
    p.n = n;
    p.kp = std::max((k + 3) / 4 * 4, min_kp);
    p.lpr = std::min(max_lpr, pow2_at_least(p.kp / 4));
    p.ntile = (p.kp + 4 * p.lpr - 1) / (4 * p.lpr);
    const int rows_per_pass = kB / p.lpr;
    int target = std::max(8, 4 * num_cu / p.ntile);
    p.rpb = std::max(rows_per_pass, (n + target - 1) / target);
    p.nseg = nseg;
    p.ks = nseg * p.kp;
    p.srow.assign(nseg + 1, 0);
    for (int q = 0; q <= nseg; ++q) p.srow[q] = seg_rows ? seg_rows[q] : (q == 0 ? 0 : n);
    for (int q = 0; q < nseg; ++q) {
        p.sb.push_back((int)p.bseg.size());
        for (int r = p.srow[q]; r < p.srow[q + 1]; r += p.rpb) {
            p.brow.push_back(r);
            p.bseg.push_back(q);
        }
    }
    while (p.bseg.size() % 8 || p.bseg.empty()) {   // empty work-groups pad to whole XCD groups
        p.brow.push_back(n);
        p.bseg.push_back(nseg - 1);
    }
    p.brow.push_back(n);
    p.sb.push_back((int)p.bseg.size());
    p.nblk = (int)p.bseg.size();
    p.nred = (p.ks + 255) / 256;
    p.grid = 0;
    p.W = p.H = p.L = p.nsegx = p.gblk = 0;
    return p;
}

void plan_bsr3(Plan& p, int nnodes, int num_cu) {
    p.bsr = 1;
    p.nnodes = nnodes;
    const int nodes_per_pass = kB / p.lpr;
    const int target = std::max(8, 4 * num_cu / p.ntile);
    p.npb = std::max(nodes_per_pass, (nnodes + target - 1) / target);
    p.bblk = (nnodes + p.npb - 1) / p.npb;
    p.bblk = (p.bblk + 7) / 8 * 8;
}

// segments of L nodes along x, one per lane group
// one work-group per kTX x kTY node tile of every grid (segments in order, tiles
// row-major within a grid), padded to whole XCD groups; grids: (first row, W, H)
void plan_grid(Plan& p, const std::vector<int>& grids) {
    p.grid = 1;
    p.gtab.clear();
    const int ng = (int)grids.size() / 3;
    for (int q = 0; q < ng; ++q) {
        const int base = grids[3 * q], W = grids[3 * q + 1], H = grids[3 * q + 2];
        for (int ty = 0; ty < H; ty += kTY)
            for (int tx = 0; tx < W; tx += kTX) {
                const int e[kGT] = {base, W, H, tx, ty, q};
                p.gtab.insert(p.gtab.end(), e, e + kGT);
            }
    }
    while ((p.gtab.size() / kGT) % 8) {
        const int e[kGT] = {0, 0, 0, 0, 0, 0};
        p.gtab.insert(p.gtab.end(), e, e + kGT);
    }
    p.gblk = (int)(p.gtab.size() / kGT);
}

void pad_mgtab(std::vector<int>& tab, int& cnt) {
    while (cnt % 8) {
        tab.insert(tab.end(), kMT, 0);
        ++cnt;
    }
}

void plan_mg(Plan& p, const MGHier& m) {
    p.mgL = m.L;
    p.mgtab.clear();
    const int ng = m.nseg;
    for (int l = 0; l < m.L; ++l) {
        p.mg_roff[l] = (int)(p.mgtab.size() / kMT);
        int cnt = 0;
        for (int q = 0; q < ng; ++q) {
            if (m.segL[q] <= l) continue;
            const size_t o = (size_t)l * ng + q, oc = (size_t)(l + 1) * ng + q;
            const int W = m.gW[o], H = m.gH[o], Wc = m.gW[oc], Hc = m.gH[oc];
            for (int cy = 0; cy < Hc; cy += kRY)
                for (int cx = 0; cx < Wc; cx += kRX, ++cnt) {
                    const int e[kMT] = {m.gbase[o], W, H, m.gbase[oc], Wc, Hc, cx, cy, q};
                    p.mgtab.insert(p.mgtab.end(), e, e + kMT);
                }
        }
        pad_mgtab(p.mgtab, cnt);
        p.mg_rn[l] = cnt;
        p.mg_poff[l] = (int)(p.mgtab.size() / kMT);
        cnt = 0;
        if (l == 0) p.mgsb.assign(ng + 1, 0);
        for (int q = 0; q < ng; ++q) {
            if (l == 0) p.mgsb[q] = cnt;
            if (m.segL[q] <= l) continue;
            const size_t o = (size_t)l * ng + q, oc = (size_t)(l + 1) * ng + q;
            const int W = m.gW[o], H = m.gH[o], Wc = m.gW[oc], Hc = m.gH[oc];
            for (int ty = 0; ty < H; ty += kPY)
                for (int tx = 0; tx < W; tx += kPX, ++cnt) {
                    const int e[kMT] = {m.gbase[o], W, H, m.gbase[oc], Wc, Hc, tx, ty, q};
                    p.mgtab.insert(p.mgtab.end(), e, e + kMT);
                }
        }
        pad_mgtab(p.mgtab, cnt);   // padding tiles count with the last grid
        p.mg_pn[l] = cnt;
        if (l == 0) p.mgsb[ng] = cnt;
    }
    size_t off = 0;
    for (int l = 1; l <= m.L; ++l) {
        p.mgv_off[l] = off;
        off += 2 * (size_t)m.N[l] * p.kp;
    }
    p.mgv_total = off;
    p.mgctab.clear();
    for (int q = 0; q < ng; ++q) {
        const int l = m.segL[q];
        const size_t o = (size_t)l * ng + q;
        const int n = m.gW[o] * m.gH[o];
        const size_t r = p.mgv_off[l] + (size_t)m.gbase[o] * p.kp;
        const size_t e = p.mgv_off[l] + ((size_t)m.N[l] + m.gbase[o]) * p.kp;
        const int c[4] = {n, m.cinv_off[q], (int)r, (int)e};
        p.mgctab.insert(p.mgctab.end(), c, c + 4);
    }
}

// k_mg_tail's first level: the finest level lt >= 1 from which every grid's
// remaining levels (from min(lt, its coarsest) down) hold r and e of four
// columns in kTailVec float4 with the partial-sum area, in at most kTailLev
// levels of at most kTailPer nodes per thread; lt = L always fits (the
// coarsest levels alone) when every coarsest level has at most kMGCoarse nodes
void plan_mg_tail(Plan& p, const MGHier& m) {
    const int ng = m.nseg;
    auto nodes = [&](int l, int q) { return (long long)m.gW[(size_t)l * ng + q] * m.gH[(size_t)l * ng + q]; };
    p.mg_tail = 0;
    for (int l0 = 1; l0 <= m.L && !p.mg_tail; ++l0) {
        bool fits = true;
        for (int q = 0; q < ng && fits; ++q) {
            const int lt = std::min(l0, m.segL[q]);
            long long tot = 0;
            for (int l = lt; l <= m.segL[q]; ++l) tot += nodes(l, q);
            fits = m.segL[q] - lt + 1 <= kTailLev && 2 * tot + kTailRed <= kTailVec &&
                   nodes(lt, q) <= (long long)kTailPer * kTailB && nodes(m.segL[q], q) <= kMGCoarse;
        }
        if (fits) p.mg_tail = l0;
    }
    p.mgttab.assign((size_t)kTT * ng, 0);
    p.mgt_lds = 0;
    if (!p.mg_tail) return;
    auto put64 = [](int* t, unsigned long long v) {
        t[0] = (int)(unsigned)(v & 0xffffffffull);
        t[1] = (int)(unsigned)(v >> 32);
    };
    for (int q = 0; q < ng; ++q) {
        int* d = p.mgttab.data() + (size_t)kTT * q;
        const int lt = std::min(p.mg_tail, m.segL[q]), nl = m.segL[q] - lt + 1;
        int off = 0;
        for (int u = 0; u < nl; ++u) {
            const int l = lt + u;
            const size_t o = (size_t)l * ng + q;
            int* e = d + 4 + kTLev * u;
            e[0] = m.gW[o];
            e[1] = m.gH[o];
            e[2] = off;
            e[3] = (int)m.N[l];
            off += 2 * e[0] * e[1];
            put64(e + 4, (unsigned long long)(uintptr_t)(m.coefT32[l].as<float>() + m.gbase[o]));
            put64(e + 6, (unsigned long long)(uintptr_t)(m.dsm[l].as<float>() + m.gbase[o]));
            put64(e + 8, (unsigned long long)(p.mgv_off[l] + (size_t)m.gbase[o] * p.kp));
            put64(e + 10, (unsigned long long)(p.mgv_off[l] + ((size_t)m.N[l] + m.gbase[o]) * p.kp));
        }
        d[0] = nl;
        d[1] = m.cinv_off[q];
        d[2] = off;
        p.mgt_lds = std::max(p.mgt_lds, sizeof(float4) * (size_t)(off + kTailRed));
    }
}

template <class T, int LPR>
struct Kern {
    static void dir_spmv(const Plan& P, const Seg& G, hipStream_t st, int it, const int* rp, const Entries<T>& M,
                         double* S, const T* z, const T* pold, T* pnew, double* acc) {
        k_dir_spmv<T, LPR><<<dim3(P.nblk, P.ntile), kB, 0, st>>>(P.kp, G, it, rp, M, S, z, pold, pnew, acc);
    }
    static void update(const Plan& P, const Seg& G, hipStream_t st, int it, const int* rp, const Entries<T>& M,
                       const T* dinv, const T* diag, const double* S, const T* p, T* x, T* z, double* acc) {
        k_cg_update<T, LPR><<<dim3(P.nblk, P.ntile), kB, 0, st>>>(P.kp, G, it, rp, M, dinv, diag, S, p, x, z, acc);
    }
    // 2D tensor-grid kernels: FP32, at most 64 columns (16 lanes) per column tile
    static void dir_spmv_g(const Plan& P, hipStream_t st, int it, const T* coef, double* S, const T* z, const T* pold,
                           T* pnew, double* acc) {
        if constexpr (std::is_same<T, float>::value && LPR <= 16)
            k_dir_spmv_t2<LPR><<<dim3(P.gblk, P.ntile), GB<LPR>::v, grid_lds_bytes(LPR), st>>>(P.d_gtab, P.kp, P.ks, it, coef, S, z, pold,
                                                                       pnew, acc);
    }
    static void update_g(const Plan& P, hipStream_t st, int it, const T* coef, const T* dinv, const T* diag,
                         const double* S, const T* p, T* x, T* z, double* acc) {
        if constexpr (std::is_same<T, float>::value && LPR <= 16)
            k_cg_update_t2<LPR><<<dim3(P.gblk, P.ntile), GB<LPR>::v, grid_lds_bytes(LPR), st>>>(P.d_gtab, P.kp, P.ks, it, coef, dinv, diag,
                                                                        S, p, x, z, acc);
    }
    // multigrid-preconditioned iteration: FP64 x, r and stencil products
    static void dir_spmv_mg(const Plan& P, hipStream_t st, int it, const double* coef64, double* S, const T* z,
                            const T* pold, T* pnew, double* acc) {
        if constexpr (std::is_same<T, float>::value && LPR <= 16)
            k_dir_spmv_mg<LPR><<<dim3(P.gblk, P.ntile), GB<LPR>::v, grid_lds_bytes(LPR), st>>>(P.d_gtab, P.kp, P.ks, it, coef64, S, z, pold,
                                                                       pnew, acc);
    }
    static void update_mg(const Plan& P, hipStream_t st, int it, const double* coef64, const double* S, const T* p,
                          const T* pprev, double* x, double* r, double* acc) {
        if constexpr (std::is_same<T, float>::value && LPR <= 16)
        {
            if (it & 1)
                k_cg_update_mg<LPR, true><<<dim3(P.gblk, P.ntile), GB<LPR>::v, grid_lds_bytes(LPR), st>>>(
                    P.d_gtab, P.kp, P.ks, it, coef64, S, p, pprev, x, r, acc);
            else
                k_cg_update_mg<LPR, false><<<dim3(P.gblk, P.ntile), GB<LPR>::v, grid_lds_bytes(LPR), st>>>(
                    P.d_gtab, P.kp, P.ks, it, coef64, S, p, pprev, x, r, acc);
        }
    }
    // z = M r (one V-cycle) with r.z into the iteration's slot (mode 0) or r.z,
    // r.r as per-tile partial rows (mode 1)
    static void vcycle(const Plan& P, const MGHier& m, hipStream_t st, int it, int mode, const T* coef0,
                       const unsigned char* cmask, const double* r0, T* z, T* mgv, double* acc, double* p0,
                       double* p1, const double* S) {
        if constexpr (std::is_same<T, float>::value && LPR >= 1 && LPR <= 16) {
            // restriction: 32 columns per work-group (half the LDS of a
            // 64-column tile, so four work-groups share a CU)
            constexpr int ML = LPR > 8 ? 8 : LPR;
            const int L = P.mgL, kp = P.kp, ntm = (kp + 4 * ML - 1) / (4 * ML);
            auto rvec = [&](int l) -> float* { return mgv + P.mgv_off[l]; };
            auto evec = [&](int l) -> float* { return l == 0 ? z : mgv + P.mgv_off[l] + (size_t)m.N[l] * kp; };
            const double* Sk = mode == 0 ? S : nullptr;   // work-group skips only inside an iteration
            // levels from lt on: one launch (k_mg_tail), else the coarse solve
            const int lt = P.mg_tail > 0 ? P.mg_tail : L;
            k_mg_restrict<ML, double><<<dim3(P.mg_rn[0], ntm), GB<ML>::v, mg_restrict_lds(ML), st>>>(
                P.d_mgtab + (size_t)kMT * P.mg_roff[0], kp, coef0, m.dsm[0].as<float>(), cmask, r0, rvec(1), Sk,
                P.ks);
            for (int l = 1; l < lt; ++l)
                k_mg_restrict<ML, float><<<dim3(P.mg_rn[l], ntm), GB<ML>::v, mg_restrict_lds(ML), st>>>(
                    P.d_mgtab + (size_t)kMT * P.mg_roff[l], kp, m.coef32[l].as<float>(), m.dsm[l].as<float>(),
                    nullptr, rvec(l), rvec(l + 1), Sk, P.ks);
            if (P.mg_tail > 0)
                k_mg_tail<<<dim3((kp / 4 + 7) / 8 * 8, m.nseg), kTailB, P.mgt_lds, st>>>(P.d_mgttab, kp, P.ks,
                                                                                       m.cinv.as<float>(), mgv, Sk);
            else
                k_mg_coarse<<<dim3((kp + kCoarseCols - 1) / kCoarseCols, m.nseg), 256, mg_coarse_lds(m.cmax), st>>>(
                    P.d_mgctab, kp, m.cinv.as<float>(), mgv);
            for (int l = lt - 1; l >= 1; --l)
                k_mg_prolong<LPR, false, float><<<dim3(P.mg_pn[l], P.ntile), GB<LPR>::v, mg_prolong_lds(LPR), st>>>(
                    P.d_mgtab + (size_t)kMT * P.mg_poff[l], kp, P.ks, it, mode, m.coef32[l].as<float>(),
                    m.dsm[l].as<float>(), nullptr, rvec(l), evec(l + 1), evec(l), nullptr, nullptr, nullptr, Sk);
            k_mg_prolong<LPR, true, double><<<dim3(P.mg_pn[0], P.ntile), GB<LPR>::v, mg_prolong_lds(LPR), st>>>(
                P.d_mgtab + (size_t)kMT * P.mg_poff[0], kp, P.ks, it, mode, coef0, m.dsm[0].as<float>(), cmask, r0,
                evec(1), z, acc, p0, p1, Sk);
        }
    }
    static void dir_spmv_b(const Plan& P, hipStream_t st, int it, const int* rpn, const int* cin, const float* bval,
                           double* S, const T* z, const T* pold, T* pnew, double* acc) {
        if constexpr (std::is_same<T, float>::value) {
            k_pdir_stream<LPR><<<dim3(P.nblk, P.ntile), kB, 0, st>>>(P.kp, P.n, P.rpb, it, S, z, pold, pnew, acc);
            k_spmv_pq_b3<LPR><<<dim3(P.bblk, P.ntile), kB, 0, st>>>(P.kp, P.nnodes, P.npb, it, rpn, cin, bval, S,
                                                                    pnew, acc);
        }
    }
    static void update_b(const Plan& P, hipStream_t st, int it, const int* rpn, const int* cin, const float* bval,
                         const T* dinv, const T* diag, const double* S, const T* p, T* x, T* z, double* acc) {
        if constexpr (std::is_same<T, float>::value)
            k_cg_update_b3<LPR><<<dim3(P.bblk, P.ntile), kB, 0, st>>>(P.kp, P.nnodes, P.npb, it, rpn, cin, bval, dinv,
                                                                      diag, S, p, x, z, acc);
    }
    static void rescale(const Plan& P, const Seg& G, hipStream_t st, const double* scale, const double* ratio,
                        const double* dinv64, const T* diag, const double* r64, T* z, T* p, double* p0, double* p1) {
        k_rescale<T, LPR><<<dim3(P.nblk, P.ntile), kB, 0, st>>>(P.kp, G, scale, ratio, dinv64, diag, r64, z, p, p0, p1);
    }
    static void resid(const Plan& P, const Seg& G, hipStream_t st, const int* rp, const int* ci, const double* va,
                      const double* x, const double* b, double* r, double* part) {
        k_resid64<LPR><<<dim3(P.nblk, P.ntile), kB, 0, st>>>(P.kp, G, rp, ci, va, x, b, r, part);
    }
};

// Stream and workspace of a solve: the context's by default; a concurrent
// solve on the same system (issued from another host thread) brings its own.
struct SolveEnv {
    hipStream_t st;
    SolveWork* w;
};
thread_local const SolveEnv* tl_env = nullptr;

// (first row, W, H) of every problem when all are declared 2D grids, else empty
std::vector<int> grids_of(const sfh_system* s, int nseg, const std::vector<int>& srow) {
    std::vector<int> grids;
    if (s->dim != 2 || s->dofpn != 1) return grids;
    const int n = s->ndof;
    if (nseg == 1 && s->gn[0] >= 3 && s->gn[1] >= 3 && (size_t)s->gn[0] * s->gn[1] == (size_t)n) {
        grids = {0, s->gn[0], s->gn[1]};
    } else if (nseg > 1 && (int)s->sgrid.size() == 3 * nseg) {
        bool match = true;
        for (int q = 0; q < nseg; ++q)
            match = match && s->sgrid[3 * q] == srow[q] &&
                    (long long)s->sgrid[3 * q + 1] * s->sgrid[3 * q + 2] == srow[q + 1] - srow[q] &&
                    s->sgrid[3 * q + 1] >= 3 && s->sgrid[3 * q + 2] >= 3;
        if (match) grids = s->sgrid;
    }
    return grids;
}

// operator data the iteration kernels read besides the CSR arrays: D^-1, diag,
// packed FP32 entries, and the stencil coefficients of declared grids
int prepare_operator_on(sfh_system* s, hipStream_t st, const std::vector<int>& grids) {
    const int n = s->ndof;
    if (s->dirty) {
        s->mg3.built = false;
        k_dinv<<<(n + 255) / 256, 256, 0, st>>>(n, s->diagpos.as<int>(), s->vals.as<double>(), s->dinv32.as<float>(),
                                                s->dinv64.as<double>(), s->diag32.as<float>(), s->diag64.as<double>());
        SFH_LAUNCH_CHECK();
        k_pack_entries<<<(unsigned)((s->nnz + 255) / 256), 256, 0, st>>>(s->nnz, s->colind.as<int>(),
                                                                         s->vals32.as<float>(), s->cv32.as<int2>());
        SFH_LAUNCH_CHECK();
        if (s->dofpn == 3 && s->kind != 0) {
            SFH_TRY(s->bsr_val.ensure(sizeof(float) * 9 * (size_t)s->nnz_node));
            k_pack_bsr3<<<(s->nnodes + 255) / 256, 256, 0, st>>>(s->nnodes, s->rowptr_node.as<int>(),
                                                                 s->vals32.as<float>(), s->bsr_val.as<float>());
            SFH_LAUNCH_CHECK();
        }
        s->dirty = false;
    }
    if (!grids.empty() && s->dia_state == 0) {
        SFH_TRY(s->dia32.ensure((size_t)n * 9 * 4));
        SFH_TRY(s->dia64.ensure((size_t)n * 9 * 8));
        DevBuf err;
        SFH_TRY(err.ensure(sizeof(int)));
        SFH_CHECK_HIP(hipMemsetAsync(err.p, 0, sizeof(int), st));
        // every grid in one launch: work-group row blockIdx.y is grid (first row, W, H)
        const int ng = (int)grids.size() / 3;
        int nmax = 0;
        for (int q = 0; q < ng; ++q) nmax = std::max(nmax, grids[3 * q + 1] * grids[3 * q + 2]);
        SFH_TRY(s->dia_tab.ensure(sizeof(int) * grids.size()));
        SFH_CHECK_HIP(hipMemcpyAsync(s->dia_tab.p, grids.data(), sizeof(int) * grids.size(), hipMemcpyHostToDevice, st));
        for (int q0 = 0; q0 < ng; q0 += 65535) {
            const int qn = std::min(65535, ng - q0);
            k_build_dia9<<<dim3((nmax + 255) / 256, qn), 256, 0, st>>>(
                s->dia_tab.as<int>() + 3 * (size_t)q0, s->rowptr.as<int>(), s->colind.as<int>(), s->vals.as<double>(),
                s->dia32.as<float>(), s->dia64.as<double>(), err.as<int>());
            SFH_LAUNCH_CHECK();
        }
        int he = 0;
        SFH_CHECK_HIP(hipMemcpyAsync(&he, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
        SFH_CHECK_HIP(hipStreamSynchronize(st));
        s->dia_state = he ? -1 : 1;
        s->mg.built = false;
        s->mg.L = 0;
    }
    // multigrid hierarchies of the grids, from the stencil copy
    if (!grids.empty() && s->dia_state == 1 && !s->mg.built) {
        SFH_TRY(mg_build(s, st, grids));
        s->mg.built = true;
    }
    return SFH_OK;
}

// Mixed-precision PCG with residual replacement ("reliable updates"):
// FP32 recurrences; whenever a column's recursive residual fell by 10x (or
// below its target) the accumulated FP32 correction is flushed into the FP64
// iterate, the true residual b - A x is recomputed in FP64 and replaces the
// FP32 one (rescaled to unit norm), and the search direction is kept.
template <class T, int LPR>
int run_solve(sfh_system* s, const Plan& P0, int trans, int k, const double* B, int ldb, double* X, int ldx,
              const sfh_solve_opts& o, sfh_solve_stats* stats) {
    hipStream_t st = tl_env ? tl_env->st : s->ctx->stream;
    SolveWork& w = tl_env ? *tl_env->w : s->work;
    Plan P = P0;
    const int n = P.n, kp = P.kp;
    const size_t nk = (size_t)n * kp;
    const int ns = P.nseg, ks = P.ks;
    // 2D tensor grids: 9-coefficient stencil layout and LDS-tiled kernels (FP32,
    // 16 to 64 columns per tile); layout 0 chooses it whenever every problem is
    // a declared 2D grid: the system's grid for one problem, the declared
    // segment grids for a batch
    const std::vector<int> grids = grids_of(s, ns, P.srow);
    // one lane per node (4-column tiles) only when a grid has a multigrid
    // hierarchy and the preconditioner is not Jacobi: narrow solves on one
    // grid (the one-column mean solve, the single-column SPDE solves) and
    // batches with a problem per column (the conductivity chains)
    bool mg_grid = false;
    for (size_t q = 0; q < grids.size(); q += 3) mg_grid = mg_grid || (long long)grids[q + 1] * grids[q + 2] > kMGCoarse;
    const bool use_grid = !grids.empty() && (o.layout == 2 || o.layout == 0) && std::is_same<T, float>::value &&
                          (P.lpr >= 2 || (P.lpr == 1 && o.precond != 1 && mg_grid)) && P.lpr <= 16;
    SFH_TRY(prepare_operator_on(s, st, use_grid ? grids : std::vector<int>()));
    if (use_grid && s->dia_state == 1) plan_grid(P, grids);
    // declared 3D grid: 27-point block stencil with its multigrid hierarchy
    // (FP32 inner type, one problem)
    bool use_g3 = false;
    G3Plan P3;
    if (!P.grid && ns == 1 && std::is_same<T, float>::value && o.precond != 1 && (o.layout == 0 || o.layout == 4) &&
        P.lpr >= 4 && P.lpr <= 16 && g3_applies(s)) {
        SFH_TRY(g3_prepare(s, st));
        use_g3 = s->mg3.L > 0;
        if (use_g3) g3_plan(P3, s->mg3, kp, P.lpr);
    }
    // vector DOFs of a mesh: 3 x 3 block layout (one problem, FP32)
    if (!P.grid && !use_g3 && ns == 1 && s->dofpn == 3 && s->kind != 0 && (o.layout == 3 || o.layout == 0) &&
        std::is_same<T, float>::value && P.ntile == 1)
        plan_bsr3(P, s->nnodes, s->ctx->num_cu);
    // multigrid preconditioner: stencil layout of one grid with a hierarchy
    const MGHier& mg = s->mg;
    bool use_mg = P.grid && o.precond != 1 && mg.built && mg.L > 0 && mg.nseg == (int)grids.size() / 3 &&
                  P.lpr >= 1;
    for (int q = 0; use_mg && q < mg.nseg; ++q)
        use_mg = mg.gbase[q] == grids[3 * q] && mg.gW[q] == grids[3 * q + 1] && mg.gH[q] == grids[3 * q + 2];
    if (use_mg) plan_mg(P, mg);
    if (use_mg && !o.mg_split) plan_mg_tail(P, mg);
    const bool fp64_rec = use_mg || use_g3;   // multigrid: FP64 iterate and residual
    const int nbi = std::max(std::max(P.grid ? P.gblk : P.nblk, use_mg ? P.mg_pn[0] : 0),
                             use_g3 ? P3.nsmooth : 0);   // partial rows per launch
    SFH_TRY(w.b64.ensure(nk * 8));
    SFH_TRY(w.x64.ensure(nk * 8));
    SFH_TRY(w.r64.ensure(nk * 8));
    SFH_TRY(w.x32.ensure(nk * sizeof(T)));
    SFH_TRY(w.r32.ensure(nk * sizeof(T)));
    SFH_TRY(w.p32.ensure(nk * sizeof(T)));
    SFH_TRY(w.q32.ensure(nk * sizeof(T)));
    SFH_TRY(w.part0.ensure((size_t)std::max(P.nblk, nbi) * kp * 8));
    SFH_TRY(w.part1.ensure((size_t)std::max(P.nblk, nbi) * kp * 8));
    SFH_TRY(w.scal.ensure((size_t)S_NSLOT * ks * 8));
    SFH_TRY(w.acc.ensure((size_t)A_N * ks * 8));
    if (use_mg) SFH_TRY(w.mgv.ensure(P.mgv_total * sizeof(T)));
    if (use_g3) SFH_TRY(w.mgv.ensure(P3.total * sizeof(float)));
    // partition tables: brow (nblk + 1), bseg (nblk), sb (nseg + 1), srow (nseg + 1), gtab, multigrid tiles,
    // and the partial-row bounds {0, level-0 prolongation tiles} of a multigrid residual replacement
    // (3D multigrid: its coarse-solve table and the partial-row bounds {0, smoothing work-groups})
    const int sb3[2] = {0, use_g3 ? P3.nsmooth : 0};
    const size_t ntab = P.brow.size() + P.bseg.size() + P.sb.size() + P.srow.size() + P.gtab.size() +
                        P.mgtab.size() + P.mgctab.size() + P.mgsb.size() + 4 + 2 + P.mgttab.size();
    SFH_TRY(w.tab.ensure(sizeof(int) * ntab));
    {
        std::vector<int> h;
        h.reserve(ntab);
        h.insert(h.end(), P.brow.begin(), P.brow.end());
        h.insert(h.end(), P.bseg.begin(), P.bseg.end());
        h.insert(h.end(), P.sb.begin(), P.sb.end());
        h.insert(h.end(), P.srow.begin(), P.srow.end());
        h.insert(h.end(), P.gtab.begin(), P.gtab.end());
        h.insert(h.end(), P.mgtab.begin(), P.mgtab.end());
        h.insert(h.end(), P.mgctab.begin(), P.mgctab.end());
        h.insert(h.end(), P.mgsb.begin(), P.mgsb.end());
        h.insert(h.end(), P3.ctab, P3.ctab + 4);
        h.insert(h.end(), sb3, sb3 + 2);
        h.insert(h.end(), P.mgttab.begin(), P.mgttab.end());
        // a solve of the same shape as the last one on this workspace finds
        // its tables in place (no upload, no wait for the queued work)
        if (h != w.tab_host || w.tab.bytes < sizeof(int) * ntab) {
            SFH_CHECK_HIP(hipMemcpyAsync(w.tab.p, h.data(), sizeof(int) * ntab, hipMemcpyHostToDevice, st));
            SFH_CHECK_HIP(hipStreamSynchronize(st));
            w.tab_host.swap(h);
        }
    }
    const int* d_brow = w.tab.as<int>();
    const int* d_bseg = d_brow + P.brow.size();
    const int* d_sb = d_bseg + P.bseg.size();
    const int* d_srow = d_sb + P.sb.size();
    P.d_gtab = d_srow + P.srow.size();
    P.d_mgtab = P.d_gtab + P.gtab.size();
    P.d_mgctab = P.d_mgtab + P.mgtab.size();
    const int* d_sbmg = P.d_mgctab + P.mgctab.size();
    P3.d_ctab = d_sbmg + P.mgsb.size();
    const int* d_sb3 = P3.d_ctab + 4;
    P.d_mgttab = d_sb3 + 2;
    const Seg G = {d_brow, d_bseg, ks};
    const dim3 gred((kp + 63) / 64, ns);
    double* S = w.scal.as<double>();
    const T* coef = std::is_same<T, float>::value ? (const T*)s->dia32.p : (const T*)s->dia64.p;
    const int* rp = s->rowptr.as<int>();
    const int* ci = s->colind.as<int>();
    const double* va = s->vals.as<double>();
    const T* vaT = std::is_same<T, float>::value ? (const T*)s->vals32.p : (const T*)s->vals.p;
    const T* dinv = std::is_same<T, float>::value ? (const T*)s->dinv32.p : (const T*)s->dinv64.p;
    const T* diag = std::is_same<T, float>::value ? (const T*)s->diag32.p : (const T*)s->diag64.p;
    const Entries<T> M = {s->cv32.as<int2>(), ci, vaT};
    T* x = w.x32.as<T>();
    T* z = w.r32.as<T>();   // Jacobi: z = D^-1 r (r itself is not kept); multigrid: z = M r
    // multigrid: the iteration updates the FP64 iterate itself and, in place
    // of the true residual after each replacement, the FP64 recursive residual
    double* rmg = w.r64.as<double>();
    // 3D: the FP32 copy of r that the gather-bound 3D V-cycle reads (x32 is free here)
    float* rfm = use_g3 ? w.x32.as<float>() : nullptr;
    T* mgv = use_mg ? w.mgv.as<T>() : nullptr;
    const double* coef64 = s->dia64.as<double>();
    const unsigned char* cmask = s->cmask.as<unsigned char>();
    T* pa = w.p32.as<T>();
    T* pb = w.q32.as<T>();
    double* p0 = w.part0.as<double>();
    double* p1 = w.part1.as<double>();

    hipEvent_t ev0, ev1;
    SFH_CHECK_HIP(hipEventCreate(&ev0));
    SFH_CHECK_HIP(hipEventCreate(&ev1));
    SFH_CHECK_HIP(hipEventRecord(ev0, st));
    const unsigned char* cnb = nullptr;
    if (!trans && s->nc > 0) {
        SFH_TRY(w.cnb.ensure((size_t)n));
        if (s->ncoo == 0) {   // whole-mesh pattern: element couplings, structurally symmetric
            SFH_CHECK_HIP(hipMemsetAsync(w.cnb.p, 0, (size_t)n, st));
            k_cnbr_scatter<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, rp, ci, s->cmask.as<unsigned char>(),
                                                                        w.cnb.as<unsigned char>());
        } else {
            k_cnbr<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, rp, ci, s->cmask.as<unsigned char>(),
                                                                w.cnb.as<unsigned char>());
        }
        SFH_LAUNCH_CHECK();
        cnb = w.cnb.as<unsigned char>();
    }
    // b and |b|^2 per column in one pass (the initial residual r = b is scaled
    // into r by the first replacement)
    k_prepare_norm<LPR><<<dim3(P.nblk, P.ntile), kB, 0, st>>>(k, kp, G, trans, B, ldb, rp, ci, va,
                                                           s->diagpos.as<int>(), s->cmask.as<unsigned char>(), cnb,
                                                           w.b64.as<double>(), p0);
    SFH_LAUNCH_CHECK();
    SFH_CHECK_HIP(hipMemsetAsync(w.x64.p, 0, nk * 8, st));
    if (!fp64_rec) SFH_CHECK_HIP(hipMemsetAsync(w.x32.p, 0, nk * sizeof(T), st));
    // (p needs no clearing: the first replacement sets it to 0, its ratio being 0)
    SFH_CHECK_HIP(hipMemsetAsync(w.scal.p, 0, (size_t)S_NSLOT * ks * 8, st));
    SFH_CHECK_HIP(hipMemsetAsync(w.acc.p, 0, (size_t)A_N * ks * 8, st));
    double* acc = w.acc.as<double>();
    sfh_ctx* cx = s->ctx;
    const int nsmp = (!tl_env && cx->ksample > 0 && k >= cx->kmin_rhs) ? (int)cx->kev.size() / 4 : 0;
    int ismp = 0;
    k_red_norm<<<gred, kRedB, 0, st>>>(kp, d_sb, p0, S + S_BN2 * ks);
    SFH_LAUNCH_CHECK();
    // per (segment, column) state, index seg * kp + col
    std::vector<double> bn2(ks), rn2(ks), scale(ks, 0.0), ratio(ks, 0.0), tgt(ks, 0.0);
    // page-locked read-back area: the five check rows, then one row for the norms
    SFH_TRY(w.host.ensure(9 * (size_t)ks * 8));
    double* hs = w.host.as<double>();   // ACTIVE, FROZEN, SCALE, TRIG and the latest r.r per column
    double* hn = hs + 5 * (size_t)ks;   // the norms
    double* hu = hs + 6 * (size_t)ks;   // scale, ratio and target of the next cycle, on their way to the device
    SFH_CHECK_HIP(hipMemcpyAsync(hn, S + S_BN2 * ks, ks * 8, hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    std::copy(hn, hn + ks, bn2.begin());
    rn2 = bn2;
    const double tol2 = o.tol * o.tol;
    // FP32 reduction per refinement cycle and host check interval: a multigrid
    // iteration reduces the residual by about an order of magnitude, a Jacobi
    // one by far less
    // (the multigrid recurrences are FP64: one cycle down to the target)
    const double inner = (o.inner_tol > 0.0 && o.inner_tol < 1.0) ? o.inner_tol : (fp64_rec ? 1e-14 : 0.01);
    int updates = 0, iters = 0;
    double maxrel = 0.0;
    bool done = false;
    const int check = o.check_every > 0 ? o.check_every : (fp64_rec ? 2 : 32);
    int jlast = -1;                           // last iteration of the previous cycle
    std::vector<double> rrchk(ks);            // r.r of every column at the previous check of the cycle
    // per-column FP32 reduction target of a cycle: starts at `inner`; a column
    // whose last cycle reached its target while the true (FP64) residual fell
    // by much less has met the FP32 rounding floor of its right-hand side
    // (about eps32 times the condition number for smooth ones), and its next
    // cycles aim at what the floor allows instead of iterating below it
    std::vector<double> dlt(ks, inner), rn2start(ks, 0.0);
    std::vector<char> reached(ks, 0);
    // largest relative FP64 residual at each replacement, and a non-finite one
    std::vector<double> hist;
    int bad_col = -1;
    // HIP graphs of the multigrid iteration pairs (2D grids): valid while
    // every launch argument and shape equals the key they were captured with
    const bool graphs = use_mg && o.graph != 0 && check == 2;
    int graph_iters = 0;
    if (graphs) {
        std::vector<unsigned long long> key = {
            (unsigned long long)LPR, sizeof(T), (unsigned long long)kp, (unsigned long long)ks, (unsigned long long)ns,
            (unsigned long long)P.gblk, (unsigned long long)P.ntile, (unsigned long long)P.mgL,
            (unsigned long long)mg.nseg, (unsigned long long)mg.cmax, (unsigned long long)(uintptr_t)st,
            (unsigned long long)(uintptr_t)P.d_gtab, (unsigned long long)(uintptr_t)P.d_mgtab,
            (unsigned long long)(uintptr_t)P.d_mgctab, (unsigned long long)(uintptr_t)coef64,
            (unsigned long long)(uintptr_t)coef, (unsigned long long)(uintptr_t)cmask,
            (unsigned long long)(uintptr_t)rmg, (unsigned long long)(uintptr_t)z, (unsigned long long)(uintptr_t)mgv,
            (unsigned long long)(uintptr_t)acc, (unsigned long long)(uintptr_t)p0, (unsigned long long)(uintptr_t)p1,
            (unsigned long long)(uintptr_t)S, (unsigned long long)(uintptr_t)pa, (unsigned long long)(uintptr_t)pb,
            (unsigned long long)(uintptr_t)w.x64.p, (unsigned long long)(uintptr_t)mg.cinv.p,
            (unsigned long long)P.mg_tail, (unsigned long long)P.mgt_lds, (unsigned long long)(uintptr_t)P.d_mgttab};
        for (int l = 0; l <= P.mgL; ++l) {
            key.push_back((unsigned long long)P.mg_rn[l]);
            key.push_back((unsigned long long)P.mg_pn[l]);
            key.push_back((unsigned long long)P.mg_roff[l]);
            key.push_back((unsigned long long)P.mg_poff[l]);
            key.push_back((unsigned long long)P.mgv_off[l]);
            key.push_back((unsigned long long)mg.N[l]);
            key.push_back((unsigned long long)(uintptr_t)mg.coef32[l].p);
            key.push_back((unsigned long long)(uintptr_t)mg.coefT32[l].p);
            key.push_back((unsigned long long)(uintptr_t)mg.dsm[l].p);
        }
        if (key != w.graphs.key) {
            w.graphs.clear();
            w.graphs.key = key;
        }
    }
    // the graph of the iteration pair starting at j (captured on first use):
    // j = 0, then j = 2, 4, 0 mod 6 (j >= 6 repeats those slot patterns)
    auto pair_graph = [&](int j, hipGraphExec_t* out) -> int {
        const int g = j == 0 ? 0 : (j % 6 == 2 ? 1 : (j % 6 == 4 ? 2 : 3));
        const int j0 = g == 0 ? 0 : (g == 1 ? 2 : (g == 2 ? 4 : 6));
        if (!w.graphs.exec[g]) {
            SFH_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            T* a = pa;
            T* b = pb;
            for (int jj = j0; jj < j0 + 2; ++jj) {
                Kern<T, LPR>::dir_spmv_mg(P, st, jj, coef64, S, z, a, b, acc);
                Kern<T, LPR>::update_mg(P, st, jj, coef64, S, b, a, w.x64.as<double>(), rmg, acc);
                Kern<T, LPR>::vcycle(P, mg, st, jj, 0, coef, cmask, rmg, z, mgv, acc, p0, p1, S);
                std::swap(a, b);
            }
            hipGraph_t graph = nullptr;
            const hipError_t e1 = hipGetLastError();
            const hipError_t e2 = hipStreamEndCapture(st, &graph);
            if (e1 != hipSuccess || e2 != hipSuccess || !graph) {
                if (graph) (void)hipGraphDestroy(graph);
                set_error("graph capture of the multigrid iterations failed: %s",
                          hipGetErrorString(e1 != hipSuccess ? e1 : e2));
                return SFH_EDEVICE;
            }
            const hipError_t e3 = hipGraphInstantiate(&w.graphs.exec[g], graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e3 != hipSuccess) {
                w.graphs.exec[g] = nullptr;
                set_error("graph instantiation failed: %s", hipGetErrorString(e3));
                return SFH_EDEVICE;
            }
        }
        *out = w.graphs.exec[g];
        return SFH_OK;
    };
    for (;;) {
        // residual replacement: decide per column, rescale, keep directions
        int nact = 0;
        maxrel = 0.0;
        for (int c = 0; c < ks && bad_col < 0; ++c)
            if (!std::isfinite(rn2[c])) bad_col = c;
        if (bad_col >= 0) break;
        for (int c = 0; c < ks; ++c) {
            const double rel = bn2[c] > 0.0 ? std::sqrt(rn2[c] / bn2[c]) : 0.0;
            maxrel = std::max(maxrel, rel);
            const bool act = bn2[c] > 0.0 && rn2[c] > tol2 * bn2[c];
            const double sn = act ? 1.0 / std::sqrt(rn2[c]) : 0.0;
            ratio[c] = (act && scale[c] > 0.0) ? sn / scale[c] : 0.0;
            if (act && reached[c] && rn2start[c] > 0.0) {
                const double got = std::sqrt(rn2[c] / rn2start[c]);
                if (got > 2.0 * dlt[c]) dlt[c] = std::min(0.1, got);
            }
            rn2start[c] = rn2[c];
            scale[c] = sn;
            // stop the FP32 recurrence at the cycle's target (the residual is
            // rescaled to unit norm), or a little below the final target, then
            // verify in FP64
            tgt[c] = act ? std::max(dlt[c] * dlt[c], 0.09 * tol2 * bn2[c] * sn * sn) : 0.0;
            nact += act;
        }
        std::fill(reached.begin(), reached.end(), 0);
        hist.push_back(maxrel);
        if (nact == 0) { done = true; break; }
        if (updates > o.max_outer || iters >= o.max_inner) break;
        std::copy(scale.begin(), scale.end(), hu);
        std::copy(ratio.begin(), ratio.end(), hu + ks);
        std::copy(tgt.begin(), tgt.end(), hu + 2 * (size_t)ks);
        SFH_CHECK_HIP(hipMemcpyAsync(S + S_SCALE * ks, hu, ks * 8, hipMemcpyHostToDevice, st));
        SFH_CHECK_HIP(hipMemcpyAsync(S + S_RATIO * ks, hu + ks, ks * 8, hipMemcpyHostToDevice, st));
        SFH_CHECK_HIP(hipMemcpyAsync(S + S_TGT * ks, hu + 2 * (size_t)ks, ks * 8, hipMemcpyHostToDevice, st));
        // at the first replacement the residual is b itself (never written to r)
        const double* rsrc = updates == 0 ? w.b64.as<double>() : w.r64.as<double>();
        if (use_g3) {
            k_mg_rescale<<<(unsigned)((nk / 4 + 255) / 256), 256, 0, st>>>((long long)nk / 4, kp, 1, d_srow,
                                                                      S + S_SCALE * ks, S + S_RATIO * ks, rsrc, rmg,
                                                                      rfm, (float*)pa);
            g3_vcycle(P3, s->mg3, st, 0, 1, cmask, rfm, (float*)z, (float*)w.mgv.p, acc, p0, p1);
            k_red_reset<<<gred, kRedB, 0, st>>>(kp, ks, d_sb3, p0, p1, S, acc, 0.0, jlast);
        } else if (use_mg) {
            k_mg_rescale<<<(unsigned)((nk / 4 + 255) / 256), 256, 0, st>>>((long long)nk / 4, kp, ns, d_srow,
                                                                      S + S_SCALE * ks, S + S_RATIO * ks, rsrc, rmg,
                                                                      nullptr, (float*)pa);
            Kern<T, LPR>::vcycle(P, mg, st, 0, 1, coef, cmask, rmg, z, mgv, acc, p0, p1, S);
            k_red_reset<<<gred, kRedB, 0, st>>>(kp, ks, d_sbmg, p0, p1, S, acc, 0.0, jlast);
        } else {
            Kern<T, LPR>::rescale(P, G, st, S + S_SCALE * ks, S + S_RATIO * ks, s->dinv64.as<double>(), diag,
                                  rsrc, z, pa, p0, p1);
            k_red_reset<<<gred, kRedB, 0, st>>>(kp, ks, d_sb, p0, p1, S, acc, 0.0, jlast);
        }
        SFH_LAUNCH_CHECK();
        ++updates;
        std::fill(rrchk.begin(), rrchk.end(), HUGE_VAL);
        // FP32 CG until some column is due for replacement: two kernels per
        // iteration, the step lengths and directions from the rotating slots
        int j = 0;
        for (;; ++j) {
            // an iteration pair as one graph replay, unless a timing sample is due in it
            if (graphs && (j & 1) == 0 && iters + 2 <= o.max_inner &&
                !(ismp < nsmp && (iters % cx->ksample <= 1 || (iters + 1) % cx->ksample == 0))) {
                hipGraphExec_t ex = nullptr;
                SFH_TRY(pair_graph(j, &ex));
                SFH_CHECK_HIP(hipGraphLaunch(ex, st));
                iters += 2;
                graph_iters += 2;
                ++j;
                if (iters >= o.max_inner) break;
                goto checkpoint;
            }
            {
            // timing samples come in iteration pairs (iters = 0 and 1 modulo
            // the sampling period), so both instantiations of the multigrid
            // update (with and without the x step) enter the averages equally
            const bool smp = ismp < nsmp && (iters % cx->ksample <= 1);
            if (smp) SFH_CHECK_HIP(hipEventRecord(cx->kev[4 * ismp], st));
            if (use_g3)
                g3_dir(P3, s->mg3, st, j, S, (const float*)z, (const float*)pa, (float*)pb, acc);
            else if (use_mg)
                Kern<T, LPR>::dir_spmv_mg(P, st, j, coef64, S, z, pa, pb, acc);
            else if (P.grid)
                Kern<T, LPR>::dir_spmv_g(P, st, j, coef, S, z, pa, pb, acc);
            else if (P.bsr)
                Kern<T, LPR>::dir_spmv_b(P, st, j, s->rowptr_node.as<int>(), s->colind_node.as<int>(),
                                         s->bsr_val.as<float>(), S, z, pa, pb, acc);
            else
                Kern<T, LPR>::dir_spmv(P, G, st, j, rp, M, S, z, pa, pb, acc);
            if (smp) SFH_CHECK_HIP(hipEventRecord(cx->kev[4 * ismp + 1], st));
            if (use_g3)
                g3_update(P3, s->mg3, st, j, S, (const float*)pb, w.x64.as<double>(), rmg, rfm, acc);
            else if (use_mg)
                Kern<T, LPR>::update_mg(P, st, j, coef64, S, pb, pa, w.x64.as<double>(), rmg, acc);
            else if (P.grid)
                Kern<T, LPR>::update_g(P, st, j, coef, dinv, diag, S, pb, x, z, acc);
            else if (P.bsr)
                Kern<T, LPR>::update_b(P, st, j, s->rowptr_node.as<int>(), s->colind_node.as<int>(),
                                       s->bsr_val.as<float>(), dinv, diag, S, pb, x, z, acc);
            else
                Kern<T, LPR>::update(P, G, st, j, rp, M, dinv, diag, S, pb, x, z, acc);
            if (smp) SFH_CHECK_HIP(hipEventRecord(cx->kev[4 * ismp + 2], st));
            if (use_mg) Kern<T, LPR>::vcycle(P, mg, st, j, 0, coef, cmask, rmg, z, mgv, acc, p0, p1, S);
            if (use_g3) g3_vcycle(P3, s->mg3, st, j, 0, cmask, rfm, (float*)z, (float*)w.mgv.p, acc, p0, p1);
            if (smp) SFH_CHECK_HIP(hipEventRecord(cx->kev[4 * ismp + 3], st));
            ismp += smp;
            std::swap(pa, pb);
            ++iters;
            if (iters >= o.max_inner) break;
            }
        checkpoint:
            if ((j + 1) % check == 0) {
                // live columns left, and columns waiting for replacement (frozen
                // now or by the r.r of this iteration, or inactive and unconverged)
                SFH_LAUNCH_CHECK();
                // ACTIVE, FROZEN, SCALE, TRIG and r.r (copied next to them) in one
                // transfer into page-locked memory
                SFH_CHECK_HIP(hipMemcpyAsync(S + (size_t)S_CHKRR * ks, acc + (size_t)(A_RR + j % 3) * ks, ks * 8,
                                             hipMemcpyDeviceToDevice, st));
                SFH_CHECK_HIP(hipMemcpyAsync(hs, S + (size_t)S_ACTIVE * ks, 5 * (size_t)ks * 8, hipMemcpyDeviceToHost,
                                             st));
                SFH_CHECK_HIP(hipStreamSynchronize(st));
                // a live column whose r.r did not fall over a whole check
                // interval has stalled in FP32: replace its residual now
                int na = 0, nu = 0;
                bool stall = false;
                for (int c = 0; c < ks; ++c) {
                    const bool act = hs[c] > 0.0;
                    const double rrc = hs[4 * (size_t)ks + c];
                    const bool fz = hs[ks + c] > 0.0 || !(rrc > hs[3 * (size_t)ks + c]);
                    reached[c] = act && fz;
                    if (act && !fz) {
                        ++na;
                        stall = stall || !(rrc < rrchk[c]);
                        rrchk[c] = rrc;
                    } else if (act || hs[2 * (size_t)ks + c] > 0.0) {
                        ++nu;
                    }
                }
                // (FP64 recurrences: a column at its target is done, not
                // waiting for a replacement; the cycle runs until none is live)
                if (na == 0 || (nu > 0 && !fp64_rec) || stall) break;
            }
        }
        jlast = j;
        if (use_mg && (jlast & 1) == 0) {   // x's last step, not added yet (steps go in pairs)
            k_x_flush_mg<<<(unsigned)((nk / 4 + 255) / 256), 256, 0, st>>>((long long)nk / 4, kp, ns, d_srow, jlast, ks, S,
                                                                      acc, (const float*)pa, w.x64.as<double>());
            SFH_LAUNCH_CHECK();
        }
        // flush the FP32 correction, true residual in FP64
        if (!fp64_rec)   // (the multigrid iterations update x64 themselves)
            k_refine<T><<<(unsigned)((nk / 4 + 255) / 256), 256, 0, st>>>((long long)nk / 4, kp, ns, d_srow,
                                                                     S + S_SCALE * ks, x, w.x64.as<double>());
        if (P.grid) {
            SFH_CHECK_HIP(hipMemsetAsync(S + S_RN2 * ks, 0, (size_t)ks * 8, st));
            k_resid_grid64<LPR><<<dim3(P.gblk, P.ntile), GB<LPR>::v, grid_lds_bytes(LPR), st>>>(
                P.d_gtab, kp, coef64, w.x64.as<double>(), w.b64.as<double>(), w.r64.as<double>(), S + S_RN2 * ks);
        } else {
            Kern<T, LPR>::resid(P, G, st, rp, ci, va, w.x64.as<double>(), w.b64.as<double>(), w.r64.as<double>(),
                                p0);
            k_red_norm<<<gred, kRedB, 0, st>>>(kp, d_sb, p0, S + S_RN2 * ks);
        }
        SFH_LAUNCH_CHECK();
        SFH_CHECK_HIP(hipMemcpyAsync(hn, S + S_RN2 * ks, ks * 8, hipMemcpyDeviceToHost, st));
        SFH_CHECK_HIP(hipStreamSynchronize(st));
        std::copy(hn, hn + ks, rn2.begin());
    }
    if ((size_t)n * k < (size_t)INT32_MAX)
        k_finalize<unsigned><<<(unsigned)(((size_t)n * ((k + 3) / 4) + 255) / 256), 256, 0, st>>>(
            n, k, kp, trans, B, ldb, w.x64.as<double>(), rp, ci, va, s->diagpos.as<int>(),
            s->cmask.as<unsigned char>(), X, ldx);
    else
        k_finalize<long long><<<(unsigned)(((size_t)n * ((k + 3) / 4) + 255) / 256), 256, 0, st>>>(
            n, k, kp, trans, B, ldb, w.x64.as<double>(), rp, ci, va, s->diagpos.as<int>(),
            s->cmask.as<unsigned char>(), X, ldx);
    SFH_LAUNCH_CHECK();
    SFH_CHECK_HIP(hipEventRecord(ev1, st));
    SFH_CHECK_HIP(hipEventSynchronize(ev1));
    float ms = 0.f;
    SFH_CHECK_HIP(hipEventElapsedTime(&ms, ev0, ev1));
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    for (int i = 0; i < ismp; ++i) {
        for (int jj = 0; jj < 3; ++jj) {
            float f = 0.f;
            SFH_CHECK_HIP(hipEventElapsedTime(&f, cx->kev[4 * i + jj], cx->kev[4 * i + jj + 1]));
            cx->ksum[jj] += f;
        }
        ++cx->kcount;
    }
    if (stats) {
        stats->outer = updates;
        stats->inner_iters = iters;
        stats->max_relres = maxrel;
        stats->ms = ms;
        stats->converged = done ? 1 : 0;
        stats->layout = use_g3 ? 4 : (P.grid ? 2 : (P.bsr ? 3 : 1));
        stats->precond = fp64_rec ? 2 : 1;
        stats->levels = use_mg ? P.mgL + 1 : (use_g3 ? s->mg3.L + 1 : 1);
        stats->graph_iters = graph_iters;
    }
    if (bad_col >= 0) {
        // (the reference's SparseLU reports the singular factorisation, corlib/SystemSolveEigenSparse.hpp:303-305)
        set_error("non-finite FP64 residual (column %d) after %d iterations: the operator is singular or not "
                  "positive definite", bad_col % kp, iters);
        return SFH_ESINGULAR;
    }
    if (!done) {
        // a residual whose best value over the last replacements is not below
        // half the best before them (stagnating, or growing as the iterate
        // drifts along a null space) marks an operator that cannot be inverted
        // on the right-hand side (singular, or a right-hand side outside its
        // range); a falling one, the iteration cap
        constexpr int kWin = 10;
        const size_t nh = hist.size();
        if (nh > (size_t)kWin &&
            !(*std::min_element(hist.end() - kWin, hist.end()) <
              0.5 * *std::min_element(hist.begin(), hist.end() - kWin))) {
            set_error("the FP64 residual stagnated at %.3g of the right-hand side over the last %d replacements "
                      "(%d iterations): the operator is singular or the right-hand side is not in its range",
                      maxrel, kWin, iters);
            return SFH_ESINGULAR;
        }
        set_error("solve did not reach tol %.3g (max rel. residual %.3g after %d iterations, %d replacements)",
                  o.tol, maxrel, iters, updates);
        return SFH_ENOCONV;
    }
    return SFH_OK;
}

template <class T>
int dispatch_solve(sfh_system* s, const Plan& P, int trans, int k, const double* B, int ldb, double* X, int ldx,
                   const sfh_solve_opts& o, sfh_solve_stats* st) {
    switch (P.lpr) {
        case 1: return run_solve<T, 1>(s, P, trans, k, B, ldb, X, ldx, o, st);
        case 2: return run_solve<T, 2>(s, P, trans, k, B, ldb, X, ldx, o, st);
        case 4: return run_solve<T, 4>(s, P, trans, k, B, ldb, X, ldx, o, st);
        case 8: return run_solve<T, 8>(s, P, trans, k, B, ldb, X, ldx, o, st);
        case 16: return run_solve<T, 16>(s, P, trans, k, B, ldb, X, ldx, o, st);
        case 32: return run_solve<T, 32>(s, P, trans, k, B, ldb, X, ldx, o, st);
        default: return run_solve<T, 64>(s, P, trans, k, B, ldb, X, ldx, o, st);
    }
}

}  // namespace

int solve(sfh_system* s, int trans, int k, const double* B, int ldb, double* X, int ldx,
          const sfh_solve_opts* opts, sfh_solve_stats* stats, int nseg, const int* seg_rows) {
    if (!s->assembled) { set_error("solve before assembly"); return SFH_ESTATE; }
    if (k <= 0 || ldb < k || ldx < k || !B || !X) { set_error("bad solve arguments"); return SFH_EARG; }
    if (nseg < 1 || (nseg > 1 && !seg_rows)) { set_error("bad segments"); return SFH_EARG; }
    if (seg_rows) {
        if (seg_rows[0] != 0 || seg_rows[nseg] != s->ndof) { set_error("segments must cover all rows"); return SFH_EARG; }
        for (int q = 0; q < nseg; ++q)
            if (seg_rows[q + 1] < seg_rows[q]) { set_error("segment bounds must not decrease"); return SFH_EARG; }
    }
    sfh_solve_opts o;
    sfh_solve_opts_default(&o);
    if (opts) o = *opts;
    if (o.tol <= 0) o.tol = 1e-10;
    if (o.inner_tol < 0 || o.inner_tol >= 1) o.inner_tol = 0;   // automatic
    if (o.max_outer <= 0) o.max_outer = 500;
    if (o.max_inner <= 0) o.max_inner = 50000;
    if (o.check_every < 0) o.check_every = 0;
    if (o.layout < 0 || o.layout > 4) o.layout = 0;
    if (o.precond < 0 || o.precond > 2) o.precond = 0;
    // column tiles: each tile is an independent multi-RHS solve whose working
    // set can stay in the last-level cache
    if (o.rhs_tile > 0 && o.rhs_tile < k) {
        sfh_solve_stats acc = {0, 0, 0.0, 0.0, 1, 0, 1, 1, 0};
        for (int c0 = 0; c0 < k; c0 += o.rhs_tile) {
            const int kc = std::min(o.rhs_tile, k - c0);
            sfh_solve_stats st1;
            sfh_solve_opts o1 = o;
            o1.rhs_tile = 0;
            const int rc = solve(s, trans, kc, B + c0, ldb, X + c0, ldx, &o1, &st1, nseg, seg_rows);
            acc.outer += st1.outer;
            acc.inner_iters += st1.inner_iters;
            acc.max_relres = std::max(acc.max_relres, st1.max_relres);
            acc.ms += st1.ms;
            acc.converged = acc.converged && st1.converged;
            acc.layout = st1.layout;
            acc.precond = st1.precond;
            acc.levels = st1.levels;
            acc.graph_iters += st1.graph_iters;
            if (rc != SFH_OK) {
                if (stats) *stats = acc;
                return rc;
            }
        }
        if (stats) *stats = acc;
        return SFH_OK;
    }
    // Declared 2D grids (stencil kernels): at most 16 lanes (64 columns) per
    // column tile, wider solves run their column tiles side by side in the
    // same launches; narrow solves on a grid with a multigrid hierarchy run as
    // 4-, 8- or 16-column tiles (1, 2 or 4 lanes per node): a dozen multigrid
    // iterations on padded columns beat thousands of Jacobi ones on the CSR
    // kernels
    int min_kp = 4, max_lpr = 64;
    if (o.precision == 0 && (o.layout == 0 || o.layout == 2)) {
        std::vector<int> srow(nseg + 1);
        for (int q = 0; q <= nseg; ++q) srow[q] = seg_rows ? seg_rows[q] : (q == 0 ? 0 : s->ndof);
        const std::vector<int> g = grids_of(s, nseg, srow);
        if (!g.empty()) max_lpr = 16;
        bool big = false;   // a grid with a multigrid hierarchy (one grid or a batch)
        for (size_t q = 0; q < g.size(); q += 3) big = big || (long long)g[q + 1] * g[q + 2] > kMGCoarse;
        if (big && o.precond != 1 && k < 16) min_kp = k <= 4 ? 4 : (k <= 8 ? 8 : 16);
    }
    if (o.precision == 0 && (o.layout == 0 || o.layout == 4) && nseg == 1 && g3_applies(s)) {
        max_lpr = 16;
        if (o.precond != 1 && k < 16) min_kp = 16;
    }
    if (o.lanes > 0) {   // caller's cap on the column tile
        int l = 1;
        while (2 * l <= std::min(o.lanes, 64)) l *= 2;
        max_lpr = std::min(max_lpr, l);
    }
    const Plan P = make_plan(s->ndof, k, s->ctx->num_cu, nseg, seg_rows, min_kp, max_lpr);
    if (o.precision == 1) return dispatch_solve<double>(s, P, trans, k, B, ldb, X, ldx, o, stats);
    return dispatch_solve<float>(s, P, trans, k, B, ldb, X, ldx, o, stats);
}

int prepare_operator(sfh_system* s, int nseg, const int* seg_rows) {
    std::vector<int> srow(nseg + 1);
    for (int q = 0; q <= nseg; ++q) srow[q] = seg_rows ? seg_rows[q] : (q == 0 ? 0 : s->ndof);
    SFH_TRY(prepare_operator_on(s, s->ctx->stream, grids_of(s, nseg, srow)));
    if (nseg == 1 && g3_applies(s)) SFH_TRY(g3_prepare(s, s->ctx->stream));
    SFH_CHECK_HIP(hipStreamSynchronize(s->ctx->stream));
    return SFH_OK;
}

int solve_on(sfh_system* s, hipStream_t st, SolveWork& w, int trans, int k, const double* B, int ldb, double* X,
             int ldx, const sfh_solve_opts* opts, sfh_solve_stats* stats) {
    const SolveEnv env = {st, &w};
    const SolveEnv* prev = tl_env;
    tl_env = &env;
    const int rc = solve(s, trans, k, B, ldb, X, ldx, opts, stats);
    tl_env = prev;
    return rc;
}

}  // namespace sfh

extern "C" void sfh_solve_opts_default(sfh_solve_opts* o) {
    o->tol = 1e-10;
    o->inner_tol = 0.0;
    o->max_outer = 500;
    o->max_inner = 50000;
    o->check_every = 0;
    o->precision = 0;
    o->rhs_tile = 0;
    o->layout = 0;
    o->precond = 0;
    o->graph = 1;
    o->lanes = 0;
    o->mg_split = 0;
}
