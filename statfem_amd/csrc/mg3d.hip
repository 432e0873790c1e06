// 3D tensor-grid path of the solver: the operator as a 27-point block stencil
// (1 or 3 DOFs per node), its geometric multigrid hierarchy, one V(1,1)-cycle
// as the preconditioner, and the PCG iteration with FP64 recurrences.
// No reference counterpart: the reference factorises the assembled operator
// (corlib/SystemSolveEigenSparse.hpp:273-311); on a declared 3D grid (hex8
// elasticity, hex8 scalar operators) this library iterates instead.
//
// Layout: node (x, y, z) of a W x H x D grid is n = (z H + y) W + x (the mesh's
// node order), DOF (n, a) is row n DPN + a (interleaved components, the
// reference's numbering). Stencil coefficient (k, a, b) of node n couples DOF
// (n, a) with DOF (n + off(k), b), off(k) = (k % 3 - 1, k / 3 % 3 - 1,
// k / 9 - 1); stored at ((n 27 + k) DPN + a) DPN + b, FP64 and FP32.
//
// Hierarchy: level l + 1 keeps every other node of level l in each direction;
// trilinear prolongation per component (odd last nodes copy their single
// parent), restricted to free DOFs on the finest level; Galerkin coarse
// operators P^T A P in FP64, 27-point again; l1-Jacobi smoothing (omega 4/3,
// D_l1 = row sums of |a|) as in the 2D hierarchy; dense FP32 inverse of the
// coarsest level (at most kMGCoarse DOFs).
//
// Kernels gather their 27 neighbours from global memory (the node-major
// layout makes each neighbour's kp columns one contiguous run); a lane group
// of LPR lanes owns one node and 4 columns per lane of each of its DPN rows.
#include <cmath>
#include <utility>
#include <vector>

#include "pcg_dev.hpp"

namespace sfh {
namespace {

constexpr int kB3 = 256;
constexpr double kOmega3 = 4.0 / 3.0;

// Node kernels step through their nodes in batches of one node per lane group
// (kB3 / LPR nodes). With 16 lane groups (LPR 16) each batch's 27 DPN^2
// stencil coefficients are first staged in LDS with coalesced loads (every
// lane of a group needs all of them); with more groups the staging would cost
// occupancy, and the lanes read the coefficients directly.
template <int LPR>
constexpr bool kStage3 = kB3 / LPR == 16;

template <int LPR, int NC, class TC>
struct CoefStage {
    TC s[kStage3<LPR> ? (kB3 / LPR) * NC : 1];
    // coefficients of batch node b0 + grp (cnt nodes in the batch); contains
    // the barriers of the staged form, so every thread of the block calls it
    __device__ const TC* load(const TC* __restrict__ coef, int b0, int cnt, int grp) {
        if constexpr (kStage3<LPR>) {
            __syncthreads();
            const TC* src = coef + (size_t)b0 * NC;
            for (int i = threadIdx.x; i < cnt * NC; i += blockDim.x) s[i] = src[i];
            __syncthreads();
            return s + grp * NC;
        } else {
            return coef + (size_t)(b0 + grp) * NC;
        }
    }
};

struct G3 {
    int W, H, D;
};

__host__ __device__ inline int g3n(const G3& g) { return g.W * g.H * g.D; }

// node index of neighbour k of (x, y, z), -1 off the grid
__device__ inline int nb27(int k, int x, int y, int z, const G3& g) {
    const int a = x + k % 3 - 1, b = y + (k / 3) % 3 - 1, c = z + k / 9 - 1;
    if (a < 0 || a >= g.W || b < 0 || b >= g.H || c < 0 || c >= g.D) return -1;
    return (c * g.H + b) * g.W + a;
}

// 1D trilinear weight of coarse index X in fine index x (n fine nodes)
__host__ __device__ inline float w1(int x, int X, int n) { return mg_axis_w(x, X, n); }

// coarse parents of fine index x: first parent, count, weight of each
__host__ __device__ inline int parents1(int x, int n, int& p0, float& w) { return mg_axis_parents(x, n, p0, w); }

// node tile of work-group (logical) b: nodes [b npb, min(n, (b + 1) npb))
template <int LPR>
__device__ inline Tile tile3(int npb, int n) {
    Tile t;
    t.lane = threadIdx.x % LPR;
    t.grp = threadIdx.x / LPR;
    t.ngrp = blockDim.x / LPR;
    t.col0 = (blockIdx.y * LPR + t.lane) * 4;
    t.bid = xcd_block(blockIdx.x, gridDim.x);
    t.r0 = min(n, t.bid * npb);
    t.r1 = min(n, t.r0 + npb);
    t.so = 0;
    return t;
}

// ---------------------------------------------------------------- set-up
// CSR -> 27-point block stencil; an entry that couples no grid neighbour sets err
template <int DPN>
__global__ void k_build_dia27(G3 g, const int* __restrict__ rp, const int* __restrict__ ci,
                              const double* __restrict__ va, double* __restrict__ d64, float* __restrict__ d32,
                              int* err) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= g3n(g) * DPN) return;
    const int n = row / DPN, a = row % DPN;
    const int x = n % g.W, y = (n / g.W) % g.H, z = n / (g.W * g.H);
    for (int k = 0; k < 27; ++k)
        for (int b = 0; b < DPN; ++b) {
            const size_t o = (((size_t)n * 27 + k) * DPN + a) * DPN + b;
            d64[o] = 0.0;
            d32[o] = 0.f;
        }
    for (int e = rp[row]; e < rp[row + 1]; ++e) {
        const int c = ci[e], m = c / DPN, b = c % DPN;
        const int dx = m % g.W - x, dy = (m / g.W) % g.H - y, dz = m / (g.W * g.H) - z;
        if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) { atomicOr(err, 1); continue; }
        const int k = (dz + 1) * 9 + (dy + 1) * 3 + dx + 1;
        const size_t o = (((size_t)n * 27 + k) * DPN + a) * DPN + b;
        d64[o] = va[e];
        d32[o] = (float)va[e];
    }
}

// A_c = P^T A P for coarse row (I, a): every fine node i of I's support, every
// stencil neighbour j of i and every coarse parent J of j. cmask (finest level,
// per DOF) drops constrained DOFs from P. A row without free support becomes
// a decoupled unit row.
template <int DPN>
__global__ __launch_bounds__(64) void k_g3_galerkin(G3 gf, G3 gc, const double* __restrict__ Af,
                                                    const unsigned char* __restrict__ cmask,
                                                    double* __restrict__ Ac) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= g3n(gc) * DPN) return;
    const int I = row / DPN, a = row % DPN;
    const int IX = I % gc.W, IY = (I / gc.W) % gc.H, IZ = I / (gc.W * gc.H);
    double acc[27 * DPN];
    for (int q = 0; q < 27 * DPN; ++q) acc[q] = 0.0;
    for (int di = 0; di < 27; ++di) {
        const int ix = 2 * IX + di % 3 - 1, iy = 2 * IY + (di / 3) % 3 - 1, iz = 2 * IZ + di / 9 - 1;
        if (ix < 0 || ix >= gf.W || iy < 0 || iy >= gf.H || iz < 0 || iz >= gf.D) continue;
        const double wi = (double)w1(ix, IX, gf.W) * w1(iy, IY, gf.H) * w1(iz, IZ, gf.D);
        const int i = (iz * gf.H + iy) * gf.W + ix;
        if (wi == 0.0 || (cmask && cmask[(size_t)i * DPN + a])) continue;
        for (int k = 0; k < 27; ++k) {
            const int j = nb27(k, ix, iy, iz, gf);
            if (j < 0) continue;
            const int jx = j % gf.W, jy = (j / gf.W) % gf.H, jz = j / (gf.W * gf.H);
            int px, py, pz;
            float fx, fy, fz;
            const int nx = parents1(jx, gf.W, px, fx), ny = parents1(jy, gf.H, py, fy),
                      nz = parents1(jz, gf.D, pz, fz);
            for (int b = 0; b < DPN; ++b) {
                const double v = Af[(((size_t)i * 27 + k) * DPN + a) * DPN + b];
                if (v == 0.0 || (cmask && cmask[(size_t)j * DPN + b])) continue;
                for (int uz = 0; uz < nz; ++uz)
                    for (int uy = 0; uy < ny; ++uy)
                        for (int ux = 0; ux < nx; ++ux) {
                            const int K = (pz + uz - IZ + 1) * 9 + (py + uy - IY + 1) * 3 + (px + ux - IX + 1);
                            acc[K * DPN + b] += wi * v * ((double)fx * fy * fz);
                        }
            }
        }
    }
    if (!(acc[13 * DPN + a] > 0.0)) {
        for (int q = 0; q < 27 * DPN; ++q) acc[q] = 0.0;
        acc[13 * DPN + a] = 1.0;
    }
    for (int K = 0; K < 27; ++K)
        for (int b = 0; b < DPN; ++b) Ac[(((size_t)I * 27 + K) * DPN + a) * DPN + b] = acc[K * DPN + b];
}

// Smoother weights of one node, DPN^2 floats: for DPN 1 the l1-Jacobi weight
// omega / sum |a| of the row; for DPN 3 the inverse of the node's 3x3 diagonal
// block, with constrained DOFs (cmask, finest level) decoupled onto unit rows.
// The block inverses are scaled afterwards by the level's omega
// (g3_block_weights). Also the FP32 copy of the node's stencil (A32).
template <int DPN>
__global__ void k_g3_level(int nn, const double* __restrict__ A, const unsigned char* __restrict__ cmask,
                           float* __restrict__ dsm, float* __restrict__ A32) {
    constexpr int NC = 27 * DPN * DPN;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn) return;
    const double* An = A + (size_t)n * NC;
    if (A32)
        for (int o = 0; o < NC; ++o) A32[(size_t)n * NC + o] = (float)An[o];
    if constexpr (DPN == 1) {
        double s = 0.0;
        for (int k = 0; k < 27; ++k) s += fabs(An[k]);
        dsm[n] = s > 0.0 ? (float)(kOmega3 / s) : 0.f;
    } else {
        double B[3][3];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) B[a][b] = An[13 * 9 + a * 3 + b];
        for (int a = 0; a < 3; ++a)
            if (cmask && cmask[(size_t)n * 3 + a]) {
                for (int b = 0; b < 3; ++b) B[a][b] = B[b][a] = 0.0;
                B[a][a] = 1.0;
            }
        // cofactor inverse; the block of an SPD operator is SPD, so det > 0
        const double c00 = B[1][1] * B[2][2] - B[1][2] * B[2][1], c01 = B[1][2] * B[2][0] - B[1][0] * B[2][2],
                     c02 = B[1][0] * B[2][1] - B[1][1] * B[2][0];
        const double det = B[0][0] * c00 + B[0][1] * c01 + B[0][2] * c02;
        float* w = dsm + (size_t)n * 9;
        if (!(det > 0.0)) {   // not SPD: diagonal weights, zero where the diagonal is not positive
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) w[a * 3 + b] = a == b && B[a][a] > 0.0 ? (float)(1.0 / B[a][a]) : 0.f;
            return;
        }
        const double id = 1.0 / det;
        const double inv[9] = {c00 * id, (B[0][2] * B[2][1] - B[0][1] * B[2][2]) * id,
                               (B[0][1] * B[1][2] - B[0][2] * B[1][1]) * id, c01 * id,
                               (B[0][0] * B[2][2] - B[0][2] * B[2][0]) * id,
                               (B[0][2] * B[1][0] - B[0][0] * B[1][2]) * id, c02 * id,
                               (B[0][1] * B[2][0] - B[0][0] * B[2][1]) * id,
                               (B[0][0] * B[1][1] - B[0][1] * B[1][0]) * id};
        for (int o = 0; o < 9; ++o) w[o] = (float)inv[o];
    }
}

// Start vector of the power iteration: hashed values in [-1, 1), zero on
// constrained DOFs
__global__ void k_g3_pinit(int nrow, const unsigned char* __restrict__ cmask, double* __restrict__ v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrow) return;
    unsigned h = (unsigned)i * 2654435761u;
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    v[i] = cmask && cmask[i] ? 0.0 : (double)(h & 0xffffu) / 32768.0 - 1.0;
}

// One power-iteration step w = B^-1 A v on a DPN-3 level (B: the nodal
// diagonal blocks, binv their inverses from k_g3_level; w zero on
// constrained DOFs), with the Rayleigh-quotient sums of v per work-group:
// part[2 blk] = v.Av, part[2 blk + 1] = v.Bv, in a fixed summation order
__global__ __launch_bounds__(256) void k_g3_power(G3 g, const double* __restrict__ A, const float* __restrict__ binv,
                                                  const unsigned char* __restrict__ cmask,
                                                  const double* __restrict__ v, double* __restrict__ w,
                                                  double* __restrict__ part) {
    __shared__ double red[2][256];
    const int n = blockIdx.x * 256 + threadIdx.x;
    double vav = 0.0, vbv = 0.0;
    if (n < g3n(g)) {
        const int x = n % g.W, y = (n / g.W) % g.H, z = n / (g.W * g.H);
        double av[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < 27; ++k) {
            const int j = nb27(k, x, y, z, g);
            if (j < 0) continue;
            const double* c = A + ((size_t)n * 27 + k) * 9;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) av[a] += c[a * 3 + b] * v[(size_t)j * 3 + b];
        }
        const double* c = A + ((size_t)n * 27 + 13) * 9;
        const float* bi = binv + (size_t)n * 9;
        for (int a = 0; a < 3; ++a) {
            const double va = v[(size_t)n * 3 + a];
            vav += va * av[a];
            for (int b = 0; b < 3; ++b) vbv += va * c[a * 3 + b] * v[(size_t)n * 3 + b];
        }
        for (int a = 0; a < 3; ++a) {
            double s = 0.0;
            for (int b = 0; b < 3; ++b) s += (double)bi[a * 3 + b] * av[b];
            w[(size_t)n * 3 + a] = cmask && cmask[(size_t)n * 3 + a] ? 0.0 : s;
        }
    }
    red[0][threadIdx.x] = vav;
    red[1][threadIdx.x] = vbv;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            red[0][threadIdx.x] += red[0][threadIdx.x + h];
            red[1][threadIdx.x] += red[1][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = red[0][0];
        part[2 * blockIdx.x + 1] = red[1][0];
    }
}

__global__ void k_g3_scale(int n, float s, float* __restrict__ a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] *= s;
}

// dense FP64 Cholesky inverse of the coarsest level's stencil, FP32 result
bool coarse_inverse3(const G3& g, int dpn, const std::vector<double>& a, std::vector<float>& inv) {
    const int nn = g3n(g), n = nn * dpn;
    std::vector<double> A((size_t)n * n, 0.0);
    for (int m = 0; m < nn; ++m) {
        const int x = m % g.W, y = (m / g.W) % g.H, z = m / (g.W * g.H);
        for (int k = 0; k < 27; ++k) {
            const int jx = x + k % 3 - 1, jy = y + (k / 3) % 3 - 1, jz = z + k / 9 - 1;
            if (jx < 0 || jx >= g.W || jy < 0 || jy >= g.H || jz < 0 || jz >= g.D) continue;
            const int j = (jz * g.H + jy) * g.W + jx;
            for (int ca = 0; ca < dpn; ++ca)
                for (int cb = 0; cb < dpn; ++cb)
                    A[(size_t)(m * dpn + ca) * n + j * dpn + cb] = a[(((size_t)m * 27 + k) * dpn + ca) * dpn + cb];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j)
            A[(size_t)i * n + j] = A[(size_t)j * n + i] = 0.5 * (A[(size_t)i * n + j] + A[(size_t)j * n + i]);
    for (int j = 0; j < n; ++j) {
        double d = A[(size_t)j * n + j];
        const double d0 = d;
        for (int k = 0; k < j; ++k) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
        if (!(d > kPivotRel * d0)) return false;   // not SPD, or singular to rounding
        d = std::sqrt(d);
        A[(size_t)j * n + j] = d;
        for (int i = j + 1; i < n; ++i) {
            double s = A[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) s -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
            A[(size_t)i * n + j] = s / d;
        }
    }
    inv.assign((size_t)n * n, 0.f);
    std::vector<double> col(n);
    for (int c = 0; c < n; ++c) {
        for (int i = 0; i < n; ++i) {
            double s = i == c ? 1.0 : 0.0;
            for (int k = 0; k < i; ++k) s -= A[(size_t)i * n + k] * col[k];
            col[i] = s / A[(size_t)i * n + i];
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = col[i];
            for (int k = i + 1; k < n; ++k) s -= A[(size_t)k * n + i] * col[k];
            col[i] = s / A[(size_t)i * n + i];
        }
        for (int i = 0; i < n; ++i) inv[(size_t)i * n + c] = (float)col[i];
    }
    return true;
}

// ------------------------------------------------------------ iteration
// p_new = m z + b p_old for every row (streaming); clears the r.z and r.r
// slots of this iteration and takes the freeze decisions
template <int LPR>
__global__ __launch_bounds__(kB3) void k_g3_pdir(int kp, int nrow, int rpb, int it, double* __restrict__ S,
                                                 const float* __restrict__ z, const float* __restrict__ pold,
                                                 float* __restrict__ pnew, double* __restrict__ acc) {
    Tile t = tile3<LPR>(rpb, nrow);
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

// q = A p at node n (FP64 stencil and sums, FP32 p); cn: the node's 27 DPN^2
// coefficients (staged in LDS)
template <int DPN>
__device__ inline void apply_p64(const G3& g, int n, const double* cn, const float* __restrict__ p,
                                 int kp, int col, V4<double> q[DPN]) {
    const int x = n % g.W, y = (n / g.W) % g.H, z = n / (g.W * g.H);
#pragma unroll
    for (int a = 0; a < DPN; ++a) q[a] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < 27; ++k) {
        const int j = nb27(k, x, y, z, g);
        if (j < 0) continue;
        const double* c = cn + k * DPN * DPN;
        V4<float> v[DPN];
#pragma unroll
        for (int b = 0; b < DPN; ++b) v[b] = ld4(p + ((size_t)j * DPN + b) * kp + col);
#pragma unroll
        for (int a = 0; a < DPN; ++a)
#pragma unroll
            for (int b = 0; b < DPN; ++b) {
                const double w = c[a * DPN + b];
                q[a].x += w * v[b].x; q[a].y += w * v[b].y; q[a].z += w * v[b].z; q[a].w += w * v[b].w;
            }
    }
}

// p.Ap into the iteration's accumulator slot
template <int DPN, int LPR>
__global__ __launch_bounds__(kB3) void k_g3_pq(G3 g, int kp, int npb, int it, const double* __restrict__ coef,
                                               const double* __restrict__ S, const float* __restrict__ p,
                                               double* __restrict__ acc) {
    constexpr int NC = 27 * DPN * DPN;
    __shared__ double red[kB3 * 4];
    __shared__ CoefStage<LPR, NC, double> cs;
    const Tile t = tile3<LPR>(npb, g3n(g));
    if (block_done(t, kp, kp, S)) return;
    V4<double> pq = {0.0, 0.0, 0.0, 0.0};
    for (int b0 = t.r0; b0 < t.r1; b0 += t.ngrp) {
        const int cnt = min(t.ngrp, t.r1 - b0);
        const double* cn = cs.load(coef, b0, cnt, t.grp);
        const int n = b0 + t.grp;
        if (t.grp < cnt && t.col0 < kp) {
            V4<double> q[DPN];
            apply_p64<DPN>(g, n, cn, p, kp, t.col0, q);
#pragma unroll
            for (int a = 0; a < DPN; ++a) {
                const V4<float> pi = ld4(p + ((size_t)n * DPN + a) * kp + t.col0);
                pq.x += pi.x * q[a].x; pq.y += pi.y * q[a].y; pq.z += pi.z * q[a].z; pq.w += pi.w * q[a].w;
            }
        }
    }
    block_accumulate<LPR>(t, kp, pq, red, acc + (A_PQ + (it & 1)) * kp);
}

// x += (alpha / s) p, r -= alpha A p (FP64 x, r, stencil and sums), r.r; rf is
// the FP32 copy of r that the V-cycle reads
template <int DPN, int LPR>
__global__ __launch_bounds__(kB3) void k_g3_update(G3 g, int kp, int npb, int it, const double* __restrict__ coef,
                                                   const double* __restrict__ S, const float* __restrict__ p,
                                                   double* __restrict__ x, double* __restrict__ r,
                                                   float* __restrict__ rf, double* __restrict__ acc) {
    constexpr int NC = 27 * DPN * DPN;
    __shared__ double red[kB3 * 4];
    __shared__ CoefStage<LPR, NC, double> cs;
    __shared__ double ca[4 * LPR], cx[4 * LPR];
    clear_slots(acc, A_PQ + ((it + 1) & 1), 1, kp);
    const Tile t = tile3<LPR>(npb, g3n(g));
    if (block_done(t, kp, kp, S)) return;
    shared_alpha<double, LPR>(it, kp, kp, 0, S, acc, ca);
    for (int c = threadIdx.x; c < 4 * LPR; c += blockDim.x) {
        const int col = blockIdx.y * 4 * LPR + c;
        const double s = col < kp ? S[S_SCALE * kp + col] : 0.0;
        cx[c] = s > 0.0 ? ca[c] / s : 0.0;
    }
    __syncthreads();
    V4<double> rr = {0.0, 0.0, 0.0, 0.0};
    const int l4 = t.lane * 4;
    const V4<double> al = {ca[l4], ca[l4 + 1], ca[l4 + 2], ca[l4 + 3]};
    const V4<double> ax = {cx[l4], cx[l4 + 1], cx[l4 + 2], cx[l4 + 3]};
    for (int b0 = t.r0; b0 < t.r1; b0 += t.ngrp) {
        const int cnt = min(t.ngrp, t.r1 - b0);
        const double* cn = cs.load(coef, b0, cnt, t.grp);
        const int n = b0 + t.grp;
        if (t.grp < cnt && t.col0 < kp) {
            V4<double> q[DPN];
            apply_p64<DPN>(g, n, cn, p, kp, t.col0, q);
#pragma unroll
            for (int a = 0; a < DPN; ++a) {
                const size_t o = ((size_t)n * DPN + a) * kp + t.col0;
                const V4<float> pi = ld4(p + o);
                V4<double> xv = ld4(x + o), rv = ld4(r + o);
                xv.x += ax.x * pi.x; xv.y += ax.y * pi.y; xv.z += ax.z * pi.z; xv.w += ax.w * pi.w;
                rv.x -= al.x * q[a].x; rv.y -= al.y * q[a].y; rv.z -= al.z * q[a].z; rv.w -= al.w * q[a].w;
                st4(x + o, xv);
                st4(r + o, rv);
                *reinterpret_cast<float4*>(rf + o) = make_float4((float)rv.x, (float)rv.y, (float)rv.z, (float)rv.w);
                rr.x += rv.x * rv.x; rr.y += rv.y * rv.y; rr.z += rv.z * rv.z; rr.w += rv.w * rv.w;
            }
        }
    }
    block_accumulate<LPR>(t, kp, rr, red, acc + (A_RR + it % 3) * kp);
}

// -------------------------------------------------------------- V-cycle
// res = r - A (S r) on free DOFs (zero on constrained ones)
template <int DPN, int LPR, class TR>
__global__ __launch_bounds__(kB3) void k_g3_resid(G3 g, int kp, int npb, const float* __restrict__ coef,
                                                  const float* __restrict__ dsm, const unsigned char* __restrict__ cmask,
                                                  const TR* __restrict__ r, float* __restrict__ res) {
    constexpr int NC = 27 * DPN * DPN;
    __shared__ CoefStage<LPR, NC, float> cs;
    const Tile t = tile3<LPR>(npb, g3n(g));
    for (int b0 = t.r0; b0 < t.r1; b0 += t.ngrp) {
        const int cnt = min(t.ngrp, t.r1 - b0);
        const float* cn = cs.load(coef, b0, cnt, t.grp);
        const int n = b0 + t.grp;
        if (!(t.grp < cnt && t.col0 < kp)) continue;
        const int x = n % g.W, y = (n / g.W) % g.H, z = n / (g.W * g.H);
        float4 q[DPN];
#pragma unroll
        for (int a = 0; a < DPN; ++a) q[a] = ldf4(r + ((size_t)n * DPN + a) * kp + t.col0);
        for (int k = 0; k < 27; ++k) {
            const int j = nb27(k, x, y, z, g);
            if (j < 0) continue;
            const float* c = cn + k * DPN * DPN;
            float4 v[DPN], e[DPN];
#pragma unroll
            for (int b = 0; b < DPN; ++b) v[b] = ldf4(r + ((size_t)j * DPN + b) * kp + t.col0);
            const float* wj = dsm + (size_t)j * DPN * DPN;
#pragma unroll
            for (int a = 0; a < DPN; ++a) {
                e[a] = zero4();
#pragma unroll
                for (int b = 0; b < DPN; ++b) fma4(wj[a * DPN + b], v[b], e[a]);
            }
#pragma unroll
            for (int a = 0; a < DPN; ++a)
#pragma unroll
                for (int b = 0; b < DPN; ++b) fma4(-c[a * DPN + b], e[b], q[a]);
        }
#pragma unroll
        for (int a = 0; a < DPN; ++a) {
            const size_t row = (size_t)n * DPN + a;
            *reinterpret_cast<float4*>(res + row * kp + t.col0) = (cmask && cmask[row]) ? zero4() : q[a];
        }
    }
}

// r_c = P^T res on coarse nodes
template <int DPN, int LPR>
__global__ __launch_bounds__(kB3) void k_g3_restrict(G3 gf, G3 gc, int kp, int npb, const float* __restrict__ res,
                                                     float* __restrict__ rc) {
    const Tile t = tile3<LPR>(npb, g3n(gc));
    if (t.col0 >= kp) return;
    for (int I = t.r0 + t.grp; I < t.r1; I += t.ngrp) {
        const int X = I % gc.W, Y = (I / gc.W) % gc.H, Z = I / (gc.W * gc.H);
        float4 s[DPN];
#pragma unroll
        for (int a = 0; a < DPN; ++a) s[a] = zero4();
        for (int d = 0; d < 27; ++d) {
            const int fx = 2 * X + d % 3 - 1, fy = 2 * Y + (d / 3) % 3 - 1, fz = 2 * Z + d / 9 - 1;
            if (fx < 0 || fx >= gf.W || fy < 0 || fy >= gf.H || fz < 0 || fz >= gf.D) continue;
            const float w = w1(fx, X, gf.W) * w1(fy, Y, gf.H) * w1(fz, Z, gf.D);
            if (w == 0.f) continue;
            const int i = (fz * gf.H + fy) * gf.W + fx;
#pragma unroll
            for (int a = 0; a < DPN; ++a) fma4(w, ldf4(res + ((size_t)i * DPN + a) * kp + t.col0), s[a]);
        }
#pragma unroll
        for (int a = 0; a < DPN; ++a)
            *reinterpret_cast<float4*>(rc + ((size_t)I * DPN + a) * kp + t.col0) = s[a];
    }
}

// e0 = S r + P e_c on fine nodes (P on free DOFs only on the finest level)
template <int DPN, int LPR, class TR>
__global__ __launch_bounds__(kB3) void k_g3_prolong(G3 gf, G3 gc, int kp, int npb, const float* __restrict__ dsm,
                                                    const unsigned char* __restrict__ cmask,
                                                    const TR* __restrict__ r, const float* __restrict__ ec,
                                                    float* __restrict__ e0) {
    const Tile t = tile3<LPR>(npb, g3n(gf));
    if (t.col0 >= kp) return;
    for (int n = t.r0 + t.grp; n < t.r1; n += t.ngrp) {
        const int x = n % gf.W, y = (n / gf.W) % gf.H, z = n / (gf.W * gf.H);
        int px, py, pz;
        float fx, fy, fz;
        const int nx = parents1(x, gf.W, px, fx), ny = parents1(y, gf.H, py, fy), nz = parents1(z, gf.D, pz, fz);
        float4 rv[DPN], v[DPN];
#pragma unroll
        for (int b = 0; b < DPN; ++b) rv[b] = ldf4(r + ((size_t)n * DPN + b) * kp + t.col0);
        const float* wn = dsm + (size_t)n * DPN * DPN;
#pragma unroll
        for (int a = 0; a < DPN; ++a) {
            v[a] = zero4();
#pragma unroll
            for (int b = 0; b < DPN; ++b) fma4(wn[a * DPN + b], rv[b], v[a]);
        }
        for (int uz = 0; uz < nz; ++uz)
            for (int uy = 0; uy < ny; ++uy)
                for (int ux = 0; ux < nx; ++ux) {
                    const int J = ((pz + uz) * gc.H + py + uy) * gc.W + px + ux;
                    const float w = fx * fy * fz;
#pragma unroll
                    for (int a = 0; a < DPN; ++a) fma4(w, ldf4(ec + ((size_t)J * DPN + a) * kp + t.col0), v[a]);
                }
#pragma unroll
        for (int a = 0; a < DPN; ++a) {
            const size_t row = (size_t)n * DPN + a;
            *reinterpret_cast<float4*>(e0 + row * kp + t.col0) = (cmask && cmask[row]) ? zero4() : v[a];
        }
    }
}

// e = e0 + S (r - A e0); TOP (finest level): e is zero on constrained DOFs and
// r.e (mode 0: into the iteration's slot; mode 1: r.e and r.r as per-work-group
// partial rows) is summed in FP64
template <int DPN, int LPR, class TR, bool TOP>
__global__ __launch_bounds__(kB3) void k_g3_smooth(G3 g, int kp, int npb, int it, int mode,
                                                   const float* __restrict__ coef, const float* __restrict__ dsm,
                                                   const unsigned char* __restrict__ cmask,
                                                   const TR* __restrict__ r, const float* __restrict__ e0,
                                                   float* __restrict__ e, double* __restrict__ acc,
                                                   double* __restrict__ part0, double* __restrict__ part1) {
    constexpr int NC = 27 * DPN * DPN;
    __shared__ double red[kB3 * 4];
    __shared__ CoefStage<LPR, NC, float> cs;
    const Tile t = tile3<LPR>(npb, g3n(g));
    V4<double> rz = {0.0, 0.0, 0.0, 0.0}, rr = {0.0, 0.0, 0.0, 0.0};
    for (int b0 = t.r0; b0 < t.r1; b0 += t.ngrp) {
        const int cnt = min(t.ngrp, t.r1 - b0);
        const float* cn = cs.load(coef, b0, cnt, t.grp);
        const int n = b0 + t.grp;
        if (t.grp < cnt && t.col0 < kp) {
            const int x = n % g.W, y = (n / g.W) % g.H, z = n / (g.W * g.H);
            float4 q[DPN];
#pragma unroll
            for (int a = 0; a < DPN; ++a) q[a] = ldf4(r + ((size_t)n * DPN + a) * kp + t.col0);
            for (int k = 0; k < 27; ++k) {
                const int j = nb27(k, x, y, z, g);
                if (j < 0) continue;
                const float* c = cn + k * DPN * DPN;
                float4 v[DPN];
#pragma unroll
                for (int b = 0; b < DPN; ++b) v[b] = ldf4(e0 + ((size_t)j * DPN + b) * kp + t.col0);
#pragma unroll
                for (int a = 0; a < DPN; ++a)
#pragma unroll
                    for (int b = 0; b < DPN; ++b) fma4(-c[a * DPN + b], v[b], q[a]);
            }
            const float* wn = dsm + (size_t)n * DPN * DPN;
#pragma unroll
            for (int a = 0; a < DPN; ++a) {
                const size_t row = (size_t)n * DPN + a;
                float4 zz = ldf4(e0 + row * kp + t.col0);
#pragma unroll
                for (int b = 0; b < DPN; ++b) fma4(wn[a * DPN + b], q[b], zz);
                if (TOP && cmask[row]) zz = zero4();
                *reinterpret_cast<float4*>(e + row * kp + t.col0) = zz;
                if constexpr (TOP) {
                    const float4 rd = ldf4(r + row * kp + t.col0);
                    rz.x += (double)rd.x * zz.x; rz.y += (double)rd.y * zz.y; rz.z += (double)rd.z * zz.z;
                    rz.w += (double)rd.w * zz.w;
                    rr.x += (double)rd.x * rd.x; rr.y += (double)rd.y * rd.y; rr.z += (double)rd.z * rd.z;
                    rr.w += (double)rd.w * rd.w;
                }
            }
        }
    }
    if constexpr (TOP) {
        if (mode == 0) {
            block_accumulate<LPR>(t, kp, rz, red, acc + (A_RZ + it % 3) * kp);
        } else {
            block_partial<LPR>(t, kp, rz, red, part0);
            __syncthreads();
            block_partial<LPR>(t, kp, rr, red, part1);
        }
    }
}

inline unsigned blocks8(long long n, int per) {
    const long long b = (n + per - 1) / per;
    return (unsigned)((b + 7) / 8 * 8);
}

// nodes per work-group: 4 per lane group
inline int npb_of(int lpr) { return 4 * (kB3 / lpr); }

template <int DPN, int LPR>
struct K3 {
    static void dir(const G3Plan& P, const MG3Hier& m, hipStream_t st, int it, double* S, const float* z,
                    const float* pold, float* pnew, double* acc) {
        const G3 g = {m.W[0], m.H[0], m.D[0]};
        const int nrow = g3n(g) * DPN, rpb = 4 * (kB3 / LPR);
        k_g3_pdir<LPR><<<dim3(blocks8(nrow, rpb), P.ntile), kB3, 0, st>>>(P.kp, nrow, rpb, it, S, z, pold, pnew, acc);
        const int npb = npb_of(LPR);
        k_g3_pq<DPN, LPR><<<dim3(blocks8(g3n(g), npb), P.ntile), kB3, 0, st>>>(g, P.kp, npb, it, m.dia64.as<double>(),
                                                                               S, pnew, acc);
    }
    static void update(const G3Plan& P, const MG3Hier& m, hipStream_t st, int it, const double* S, const float* p,
                       double* x, double* r, float* rf, double* acc) {
        const G3 g = {m.W[0], m.H[0], m.D[0]};
        const int npb = npb_of(LPR);
        k_g3_update<DPN, LPR><<<dim3(blocks8(g3n(g), npb), P.ntile), kB3, 0, st>>>(g, P.kp, npb, it,
                                                                                   m.dia64.as<double>(), S, p, x, r,
                                                                                   rf, acc);
    }
    static void vcycle(const G3Plan& P, const MG3Hier& m, hipStream_t st, int it, int mode,
                       const unsigned char* cmask, const float* r0, float* z, float* mgv, double* acc, double* p0,
                       double* p1) {
        const int L = m.L, kp = P.kp, npb = npb_of(LPR);
        auto g = [&](int l) { return G3{m.W[l], m.H[l], m.D[l]}; };
        auto rv = [&](int l) { return mgv + P.off[l]; };
        auto ev = [&](int l) { return mgv + P.off[l] + (size_t)g3n(g(l)) * DPN * kp; };
        auto tv = [&](int l) { return l == 0 ? mgv : mgv + P.off[l] + 2 * (size_t)g3n(g(l)) * DPN * kp; };
        const dim3 b0(blocks8(g3n(g(0)), npb), P.ntile);
        k_g3_resid<DPN, LPR, float><<<b0, kB3, 0, st>>>(g(0), kp, npb, m.dia32.as<float>(), m.dsm[0].as<float>(),
                                                          cmask, r0, tv(0));
        for (int l = 0; l < L; ++l) {
            if (l > 0)
                k_g3_resid<DPN, LPR, float><<<dim3(blocks8(g3n(g(l)), npb), P.ntile), kB3, 0, st>>>(
                    g(l), kp, npb, m.coef32[l].as<float>(), m.dsm[l].as<float>(), nullptr, rv(l), tv(l));
            k_g3_restrict<DPN, LPR><<<dim3(blocks8(g3n(g(l + 1)), npb), P.ntile), kB3, 0, st>>>(
                g(l), g(l + 1), kp, npb, tv(l), rv(l + 1));
        }
        k_mg_coarse<<<dim3((kp + kCoarseCols - 1) / kCoarseCols, 1), 256, mg_coarse_lds(m.nc), st>>>(
            P.d_ctab, kp, m.cinv.as<float>(), mgv);
        for (int l = L - 1; l >= 0; --l) {
            const dim3 bl(blocks8(g3n(g(l)), npb), P.ntile);
            if (l > 0) {
                k_g3_prolong<DPN, LPR, float><<<bl, kB3, 0, st>>>(g(l), g(l + 1), kp, npb, m.dsm[l].as<float>(),
                                                                   nullptr, rv(l), ev(l + 1), tv(l));
                k_g3_smooth<DPN, LPR, float, false><<<bl, kB3, 0, st>>>(g(l), kp, npb, it, mode,
                                                                         m.coef32[l].as<float>(),
                                                                         m.dsm[l].as<float>(), nullptr, rv(l), tv(l),
                                                                         ev(l), nullptr, nullptr, nullptr);
            } else {
                k_g3_prolong<DPN, LPR, float><<<bl, kB3, 0, st>>>(g(0), g(1), kp, npb, m.dsm[0].as<float>(), cmask,
                                                                    r0, ev(1), tv(0));
                k_g3_smooth<DPN, LPR, float, true><<<bl, kB3, 0, st>>>(g(0), kp, npb, it, mode,
                                                                         m.dia32.as<float>(), m.dsm[0].as<float>(),
                                                                         cmask, r0, tv(0), z, acc, p0, p1);
            }
        }
    }
};

template <int DPN>
void dispatch_lpr(int lpr, int what, const G3Plan& P, const MG3Hier& m, hipStream_t st, int it, int mode,
                  double* S, const float* z, const float* pold, float* pnew, double* x, double* r, float* rf,
                  const unsigned char* cmask, float* mgv, double* acc, double* p0, double* p1) {
#define SFH_G3(L)                                                                                \
    do {                                                                                         \
        if (what == 0) K3<DPN, L>::dir(P, m, st, it, S, z, pold, pnew, acc);                     \
        else if (what == 1) K3<DPN, L>::update(P, m, st, it, S, pold, x, r, rf, acc);            \
        else K3<DPN, L>::vcycle(P, m, st, it, mode, cmask, rf, pnew, mgv, acc, p0, p1);          \
    } while (0)
    if (lpr == 4) SFH_G3(4);
    else if (lpr == 8) SFH_G3(8);
    else SFH_G3(16);
#undef SFH_G3
}

void g3_launch(int what, const G3Plan& P, const MG3Hier& m, hipStream_t st, int it, int mode, double* S,
               const float* z, const float* pold, float* pnew, double* x, double* r, float* rf,
               const unsigned char* cmask, float* mgv, double* acc, double* p0, double* p1) {
    if (m.dpn == 3)
        dispatch_lpr<3>(P.lpr, what, P, m, st, it, mode, S, z, pold, pnew, x, r, rf, cmask, mgv, acc, p0, p1);
    else
        dispatch_lpr<1>(P.lpr, what, P, m, st, it, mode, S, z, pold, pnew, x, r, rf, cmask, mgv, acc, p0, p1);
}

}  // namespace

// ------------------------------------------------------------------ host
void level_weights(int dpn, int nn, const double* A, const unsigned char* cmask, float* dsm, float* A32,
                   hipStream_t st) {
    const unsigned b = (unsigned)((nn + 255) / 256);
    if (dpn == 3) k_g3_level<3><<<b, 256, 0, st>>>(nn, A, cmask, dsm, A32);
    else k_g3_level<1><<<b, 256, 0, st>>>(nn, A, cmask, dsm, A32);
}

// Block-Jacobi weight of level l (DPN 3): omega = kBlockOmegaLambda / lambda,
// lambda the Rayleigh quotient of kPowerSteps power-iteration steps on B^-1 A,
// a lower estimate of lambda_max(B^-1 A) (the level's 3x3 inverses are scaled
// by omega in place). The smoother contracts in the A-norm for
// omega lambda_max < 2; the estimate falls short of lambda_max by a few per
// cent, so omega lambda_max stays near 1.7, where block Jacobi damps best as
// a V-cycle smoother of hex8 elasticity.
constexpr int kPowerSteps = 15;
constexpr double kBlockOmegaLambda = 1.6;
int g3_block_weights(MG3Hier& m, int l, const unsigned char* cmask, hipStream_t st) {
    const G3 g = {m.W[l], m.H[l], m.D[l]};
    const int nn = g3n(g), nrow = 3 * nn, nb = (nn + 255) / 256;
    const double* A = l == 0 ? m.dia64.as<double>() : m.coef64[l].as<double>();
    DevBuf v, w, part;
    SFH_TRY(v.ensure(sizeof(double) * nrow));
    SFH_TRY(w.ensure(sizeof(double) * nrow));
    SFH_TRY(part.ensure(sizeof(double) * 2 * nb));
    std::vector<double> hp(2 * (size_t)nb);
    k_g3_pinit<<<(unsigned)((nrow + 255) / 256), 256, 0, st>>>(nrow, cmask, v.as<double>());
    SFH_LAUNCH_CHECK();
    double lam = 0.0;
    for (int k = 0; k < kPowerSteps; ++k) {
        k_g3_power<<<(unsigned)nb, 256, 0, st>>>(g, A, m.dsm[l].as<float>(), cmask, v.as<double>(), w.as<double>(),
                                                 part.as<double>());
        SFH_LAUNCH_CHECK();
        SFH_CHECK_HIP(hipMemcpyAsync(hp.data(), part.p, hp.size() * 8, hipMemcpyDeviceToHost, st));
        SFH_CHECK_HIP(hipStreamSynchronize(st));
        double vav = 0.0, vbv = 0.0;
        for (int b = 0; b < nb; ++b) {
            vav += hp[2 * b];
            vbv += hp[2 * b + 1];
        }
        if (vbv > 0.0 && vav / vbv > lam) lam = vav / vbv;
        std::swap(v.p, w.p);
    }
    if (!(lam > 0.0) || !std::isfinite(lam)) return 1;   // no usable estimate: no hierarchy
    k_g3_scale<<<(unsigned)((9 * nn + 255) / 256), 256, 0, st>>>(9 * nn, (float)(kBlockOmegaLambda / lam),
                                                                   m.dsm[l].as<float>());
    SFH_LAUNCH_CHECK();
    m.lam[l] = lam;
    return SFH_OK;
}

bool g3_applies(const sfh_system* s) {
    return s->dim == 3 && (s->dofpn == 1 || s->dofpn == 3) && s->gn[0] >= 2 && s->gn[1] >= 2 && s->gn[2] >= 2 &&
           (long long)s->gn[0] * s->gn[1] * s->gn[2] == s->nnodes &&
           (long long)s->nnodes * s->dofpn > kMGCoarse;
}

int g3_prepare(sfh_system* s, hipStream_t st) {
    MG3Hier& m = s->mg3;
    if (m.built) return SFH_OK;
    m.L = 0;
    m.built = true;
    if (!g3_applies(s)) return SFH_OK;
    const int dpn = s->dofpn;
    m.dpn = dpn;
    m.W[0] = s->gn[0];
    m.H[0] = s->gn[1];
    m.D[0] = s->gn[2];
    const G3 g0 = {m.W[0], m.H[0], m.D[0]};
    const long long n0 = (long long)g3n(g0), nrow0 = n0 * dpn;
    const size_t ncoef0 = (size_t)n0 * 27 * dpn * dpn;
    SFH_TRY(m.dia64.ensure(ncoef0 * 8));
    SFH_TRY(m.dia32.ensure(ncoef0 * 4));
    DevBuf err;
    SFH_TRY(err.ensure(sizeof(int)));
    SFH_CHECK_HIP(hipMemsetAsync(err.p, 0, sizeof(int), st));
    if (dpn == 3)
        k_build_dia27<3><<<(unsigned)((nrow0 + 255) / 256), 256, 0, st>>>(g0, s->rowptr.as<int>(), s->colind.as<int>(),
                                                                        s->vals.as<double>(), m.dia64.as<double>(),
                                                                        m.dia32.as<float>(), err.as<int>());
    else
        k_build_dia27<1><<<(unsigned)((nrow0 + 255) / 256), 256, 0, st>>>(g0, s->rowptr.as<int>(), s->colind.as<int>(),
                                                                        s->vals.as<double>(), m.dia64.as<double>(),
                                                                        m.dia32.as<float>(), err.as<int>());
    SFH_LAUNCH_CHECK();
    int he = 0;
    SFH_CHECK_HIP(hipMemcpyAsync(&he, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    if (he) return SFH_OK;   // not a 27-point operator on this grid
    SFH_TRY(m.dsm[0].ensure(sizeof(float) * (size_t)n0 * dpn * dpn));
    level_weights(dpn, (int)n0, m.dia64.as<double>(), s->cmask.as<unsigned char>(), m.dsm[0].as<float>(), nullptr,
                  st);
    SFH_LAUNCH_CHECK();
    int l = 0;
    while ((long long)m.W[l] * m.H[l] * m.D[l] * dpn > kMGCoarse && l < kMGMaxLev &&
           (m.W[l] > 2 || m.H[l] > 2 || m.D[l] > 2)) {
        const G3 gf = {m.W[l], m.H[l], m.D[l]};
        const G3 gc = {mg_coarse_n(gf.W), mg_coarse_n(gf.H), mg_coarse_n(gf.D)};
        m.W[l + 1] = gc.W;
        m.H[l + 1] = gc.H;
        m.D[l + 1] = gc.D;
        const long long nrc = (long long)g3n(gc) * dpn;
        SFH_TRY(m.coef64[l + 1].ensure((size_t)g3n(gc) * 27 * dpn * dpn * 8));
        SFH_TRY(m.coef32[l + 1].ensure((size_t)g3n(gc) * 27 * dpn * dpn * 4));
        SFH_TRY(m.dsm[l + 1].ensure(sizeof(float) * (size_t)g3n(gc) * dpn * dpn));
        const double* Af = l == 0 ? m.dia64.as<double>() : m.coef64[l].as<double>();
        const unsigned char* cm = l == 0 ? s->cmask.as<unsigned char>() : nullptr;
        if (dpn == 3)
            k_g3_galerkin<3><<<(unsigned)((nrc + 63) / 64), 64, 0, st>>>(gf, gc, Af, cm, m.coef64[l + 1].as<double>());
        else
            k_g3_galerkin<1><<<(unsigned)((nrc + 63) / 64), 64, 0, st>>>(gf, gc, Af, cm, m.coef64[l + 1].as<double>());
        SFH_LAUNCH_CHECK();
        level_weights(dpn, g3n(gc), m.coef64[l + 1].as<double>(), nullptr, m.dsm[l + 1].as<float>(),
                      m.coef32[l + 1].as<float>(), st);
        SFH_LAUNCH_CHECK();
        ++l;
    }
    if (l == 0) return SFH_OK;
    if (dpn == 3)
        for (int k = 0; k < l; ++k) {
            const int rc = g3_block_weights(m, k, k == 0 ? s->cmask.as<unsigned char>() : nullptr, st);
            if (rc == 1) return SFH_OK;
            SFH_TRY(rc);
        }
    const G3 gL = {m.W[l], m.H[l], m.D[l]};
    if ((long long)g3n(gL) * dpn > kMGCoarse) return SFH_OK;   // coarsest inverse must fit the LDS-staged solve
    std::vector<double> a((size_t)g3n(gL) * 27 * dpn * dpn);
    SFH_CHECK_HIP(hipMemcpyAsync(a.data(), m.coef64[l].p, a.size() * 8, hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    std::vector<float> inv;
    if (!coarse_inverse3(gL, dpn, a, inv)) return SFH_OK;
    SFH_TRY(m.cinv.ensure(inv.size() * 4));
    SFH_CHECK_HIP(hipMemcpyAsync(m.cinv.p, inv.data(), inv.size() * 4, hipMemcpyHostToDevice, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    m.nc = g3n(gL) * dpn;
    m.L = l;
    return SFH_OK;
}

void g3_plan(G3Plan& P, const MG3Hier& m, int kp, int lpr) {
    P.kp = kp;
    P.lpr = lpr;
    P.ntile = (kp + 4 * lpr - 1) / (4 * lpr);
    size_t off = (size_t)m.W[0] * m.H[0] * m.D[0] * m.dpn * kp;   // level-0 temporary
    for (int l = 1; l <= m.L; ++l) {
        P.off[l] = off;
        off += 3 * (size_t)m.W[l] * m.H[l] * m.D[l] * m.dpn * kp;   // r_l, e_l, t_l
    }
    P.total = off;
    const int L = m.L;
    P.ctab[0] = m.nc;
    P.ctab[1] = 0;
    P.ctab[2] = (int)P.off[L];
    P.ctab[3] = (int)(P.off[L] + (size_t)m.W[L] * m.H[L] * m.D[L] * m.dpn * kp);
    P.nsmooth = (int)blocks8((long long)m.W[0] * m.H[0] * m.D[0], npb_of(lpr));
}

void g3_dir(const G3Plan& P, const MG3Hier& m, hipStream_t st, int it, double* S, const float* z, const float* pold,
            float* pnew, double* acc) {
    g3_launch(0, P, m, st, it, 0, S, z, pold, pnew, nullptr, nullptr, nullptr, nullptr, nullptr, acc, nullptr,
              nullptr);
}

void g3_update(const G3Plan& P, const MG3Hier& m, hipStream_t st, int it, const double* S, const float* p, double* x,
               double* r, float* rf, double* acc) {
    g3_launch(1, P, m, st, it, 0, const_cast<double*>(S), nullptr, p, nullptr, x, r, rf, nullptr, nullptr, acc,
              nullptr, nullptr);
}

void g3_vcycle(const G3Plan& P, const MG3Hier& m, hipStream_t st, int it, int mode, const unsigned char* cmask,
               const float* r, float* z, float* mgv, double* acc, double* p0, double* p1) {
    g3_launch(2, P, m, st, it, mode, nullptr, nullptr, nullptr, z, nullptr, nullptr, const_cast<float*>(r), cmask,
              mgv, acc, p0, p1);
}

}  // namespace sfh
