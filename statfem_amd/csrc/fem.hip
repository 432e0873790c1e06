// GPU pattern build, element-matrix assembly, body-force assembly and row-zero
// Dirichlet constraints.
//
// Assembly is node-row-owner ("gather") and atomic-free: one thread owns one
// DOF row, walks the elements adjacent to its node in ascending element order
// (the node->element adjacency is sorted at setup), integrates that element's
// row block and accumulates it into a per-thread row buffer in LDS through the
// precomputed scatter map (element-local column -> CSR position). Rows are
// written once, in FP64 and as the FP32 shadow used by the inner PCG.
//
// Reference semantics: corlib/MatrixAssembler.hpp:198-232 (compute, *factor,
// insert), corlib/Triplet.hpp:126-170 (pattern = union of element couplings,
// ascending columns, explicit zeros kept), corlib/ComputeElementMatrix.hpp:62-77
// (quadrature loop), corlib/ElementBasic.ipp:177-205 (geometry).
#include "common.hpp"

namespace sfh {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxAdj = 64;      // max elements around a node
constexpr int kMaxCand = 64;     // max distinct nodes coupled to a node

// ------------------------------------------------------------ quadrature
// corlib/GaussLegendre.hpp:58-80 (constants as stored there)
__device__ inline void gl_point(int n, int i, double& w, double& x) {
    if (n == 1) {
        w = 1.0; x = 0.5;
    } else if (n == 2) {
        w = 0.5; x = (i == 0) ? 0.788675134594813 : 0.211324865405187;
    } else {
        w = (i == 1) ? 0.444444444444444 : 0.277777777777777;
        x = (i == 0) ? 0.887298334620741 : ((i == 1) ? 0.5 : 0.112701665379259);
    }
}

// Tensor ordering of corlib/TensorQuadrature.hpp:100-140 (x innermost);
// 1-point triangle rule of corlib/SimplexQuadrature.hpp:162-171.
template <int KIND>
__device__ inline void quad_point(int q, int n, double xi[3], double& w) {
    if (KIND == SFH_TRI3) {
        xi[0] = 0.333333333333333; xi[1] = 0.333333333333333; xi[2] = 0.0; w = 0.5;
        return;
    }
    double w0, w1, w2;
    if (KIND == SFH_LINE2) {
        gl_point(n, q, w0, xi[0]); xi[1] = xi[2] = 0.0; w = w0;
    } else if (KIND == SFH_QUAD4) {
        gl_point(n, q % n, w0, xi[0]); gl_point(n, q / n, w1, xi[1]); xi[2] = 0.0;
        w = w0; w *= w1;
    } else {
        gl_point(n, q % n, w0, xi[0]); gl_point(n, (q / n) % n, w1, xi[1]);
        gl_point(n, q / (n * n), w2, xi[2]);
        w = w0; w *= w1; w *= w2;
    }
}

template <int KIND> struct ET;
template <> struct ET<SFH_LINE2> { static constexpr int nen = 2, dim = 1; };
template <> struct ET<SFH_TRI3>  { static constexpr int nen = 3, dim = 2; };
template <> struct ET<SFH_QUAD4> { static constexpr int nen = 4, dim = 2; };
template <> struct ET<SFH_HEX8>  { static constexpr int nen = 8, dim = 3; };

// shape functions and reference gradients, corlib/Shapefun.ipp
template <int KIND>
__device__ inline void shape(const double xi[3], double* phi, double (*dphi)[ET<KIND>::nen]) {
    const double a = xi[0], b = xi[1], c = xi[2];
    if (KIND == SFH_LINE2) {
        phi[0] = 1. - a; phi[1] = a;
        dphi[0][0] = -1.; dphi[0][1] = 1.;
    } else if (KIND == SFH_TRI3) {
        phi[0] = 1. - (a + b); phi[1] = a; phi[2] = b;
        dphi[0][0] = -1.; dphi[0][1] = 1.; dphi[0][2] = 0.;
        dphi[1][0] = -1.; dphi[1][1] = 0.; dphi[1][2] = 1.;
    } else if (KIND == SFH_QUAD4) {
        phi[0] = (1. - a) * (1. - b); phi[1] = a * (1. - b); phi[2] = a * b; phi[3] = (1. - a) * b;
        dphi[0][0] = -(1. - b); dphi[0][1] = (1. - b); dphi[0][2] = b; dphi[0][3] = -b;
        dphi[1][0] = -(1. - a); dphi[1][1] = -a; dphi[1][2] = a; dphi[1][3] = (1. - a);
    } else {
        const double ma = 1. - a, mb = 1. - b, mc = 1. - c;
        phi[0] = ma * mb * mc; phi[1] = a * mb * mc; phi[2] = a * b * mc; phi[3] = ma * b * mc;
        phi[4] = ma * mb * c;  phi[5] = a * mb * c;  phi[6] = a * b * c;  phi[7] = ma * b * c;
        dphi[0][0] = -mb * mc; dphi[0][1] = mb * mc; dphi[0][2] = b * mc; dphi[0][3] = -b * mc;
        dphi[0][4] = -mb * c;  dphi[0][5] = mb * c;  dphi[0][6] = b * c;  dphi[0][7] = -b * c;
        dphi[1][0] = -ma * mc; dphi[1][1] = -a * mc; dphi[1][2] = a * mc; dphi[1][3] = ma * mc;
        dphi[1][4] = -ma * c;  dphi[1][5] = -a * c;  dphi[1][6] = a * c;  dphi[1][7] = ma * c;
        dphi[2][0] = -ma * mb; dphi[2][1] = -a * mb; dphi[2][2] = -a * b; dphi[2][3] = -ma * b;
        dphi[2][4] = ma * mb;  dphi[2][5] = a * mb;  dphi[2][6] = a * b;  dphi[2][7] = ma * b;
    }
}

// corlib/ElementBasic.ipp:177-205 with the square contravariant basis of
// corlib/linalg.hpp:200-223: g = (J^T)^-1 dphi/dxi, returns det J.
template <int KIND>
__device__ inline double gderiv(const double (*X)[ET<KIND>::dim], const double xi[3], double* phi,
                                double (*g)[ET<KIND>::nen]) {
    constexpr int D = ET<KIND>::dim, NEN = ET<KIND>::nen;
    double dphi[D][NEN];
    shape<KIND>(xi, phi, dphi);
    double m[D][D];   // m[al][i] = sum_K dphi[al][K] X[K][i]  (= J^T)
#pragma unroll
    for (int al = 0; al < D; ++al)
#pragma unroll
        for (int i = 0; i < D; ++i) {
            double s = 0.0;
#pragma unroll
            for (int K = 0; K < NEN; ++K) s += X[K][i] * dphi[al][K];
            m[al][i] = s;
        }
    double inv[D][D];
    double det;
    if (D == 1) {
        det = m[0][0];
        inv[0][0] = 1.0 / m[0][0];
    } else if (D == 2) {
        det = m[0][0] * m[1][1] - m[0][1] * m[1][0];
        inv[0][0] = m[1][1] / det; inv[0][1] = -m[0][1] / det;
        inv[1][0] = -m[1][0] / det; inv[1][1] = m[0][0] / det;
    } else {
        const double* a = &m[0][0];
        det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
              a[2] * (a[3] * a[7] - a[4] * a[6]);
        double* v = &inv[0][0];
        v[0] = (a[4] * a[8] - a[5] * a[7]) / det; v[1] = (a[2] * a[7] - a[1] * a[8]) / det;
        v[2] = (a[1] * a[5] - a[2] * a[4]) / det; v[3] = (a[5] * a[6] - a[3] * a[8]) / det;
        v[4] = (a[0] * a[8] - a[2] * a[6]) / det; v[5] = (a[2] * a[3] - a[0] * a[5]) / det;
        v[6] = (a[3] * a[7] - a[4] * a[6]) / det; v[7] = (a[1] * a[6] - a[0] * a[7]) / det;
        v[8] = (a[0] * a[4] - a[1] * a[3]) / det;
    }
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int K = 0; K < NEN; ++K) {
            double s = 0.0;
#pragma unroll
            for (int al = 0; al < D; ++al) s += inv[i][al] * dphi[al][K];
            g[i][K] = s;
        }
    return det;
}

template <int KIND>
__device__ inline void load_X(const double* coords, const int* cn, double (*X)[ET<KIND>::dim]) {
#pragma unroll
    for (int K = 0; K < ET<KIND>::nen; ++K) {
        const int n = cn[K];
#pragma unroll
        for (int i = 0; i < ET<KIND>::dim; ++i) X[K][i] = coords[(size_t)n * ET<KIND>::dim + i];
    }
}

// ---------------------------------------------------------------- pattern
__global__ void k_adj_count(int nel, int nen, const int* __restrict__ conn, int* cnt) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nel * nen) return;
    atomicAdd(&cnt[conn[t]], 1);
}

__global__ void k_adj_fill(int nel, int nen, const int* __restrict__ conn, const int* ptr, int* fill,
                           int* adj) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nel * nen) return;
    const int n = conn[t];
    const int pos = atomicAdd(&fill[n], 1);
    const int e = (int)(t / nen), a = (int)(t % nen);
    adj[ptr[n] + pos] = (e << 3) | a;
}

// sorts each node's adjacency by (element, local index): element order of the
// reference's mesh.iterateOverElements, independent of atomic arrival order
__global__ void k_adj_sort(int nnodes, const int* ptr, int* adj, int* err) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nnodes) return;
    const int b = ptr[n], e = ptr[n + 1];
    if (e - b > kMaxAdj) { atomicOr(err, 1); return; }
    for (int i = b + 1; i < e; ++i) {
        const int v = adj[i];
        int j = i - 1;
        while (j >= b && adj[j] > v) { adj[j + 1] = adj[j]; --j; }
        adj[j + 1] = v;
    }
}

// distinct coupled nodes of node n in ascending order -> u[0..cnt)
__device__ inline int node_row(int n, int nen, const int* conn, const int* ptr, const int* adj,
                               int* u, int* err) {
    int cnt = 0;
    for (int s = ptr[n]; s < ptr[n + 1]; ++s) {
        const int e = adj[s] >> 3;
        for (int b = 0; b < nen; ++b) {
            const int c = conn[(size_t)e * nen + b];
            int pos = 0;
            while (pos < cnt && u[pos] < c) ++pos;
            if (pos < cnt && u[pos] == c) continue;
            if (cnt == kMaxCand) { atomicOr(err, 2); return cnt; }
            for (int q = cnt; q > pos; --q) u[q] = u[q - 1];
            u[pos] = c;
            ++cnt;
        }
    }
    return cnt;
}

__global__ void k_row_count(int nnodes, int nen, const int* conn, const int* ptr, const int* adj,
                            int* len, int* err) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nnodes) return;
    int u[kMaxCand];
    len[n] = node_row(n, nen, conn, ptr, adj, u, err);
}

__global__ void k_row_fill(int nnodes, int nen, const int* conn, const int* ptr, const int* adj,
                           const int* rowptr, int* colind, unsigned char* emap, int* err) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nnodes) return;
    int u[kMaxCand];
    const int cnt = node_row(n, nen, conn, ptr, adj, u, err);
    const int base = rowptr[n];
    for (int p = 0; p < cnt; ++p) colind[base + p] = u[p];
    for (int s = ptr[n]; s < ptr[n + 1]; ++s) {
        const int e = adj[s] >> 3;
        for (int b = 0; b < nen; ++b) {
            const int c = conn[(size_t)e * nen + b];
            int lo = 0, hi = cnt - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (u[mid] < c) lo = mid + 1; else hi = mid;
            }
            emap[(size_t)s * nen + b] = (unsigned char)lo;
        }
    }
}

// interleaved vector DOFs: DOF row 3i+c holds columns 3j+k for every coupled node j
__global__ void k_expand_pattern(int nnodes, int d, const int* rpn, const int* cin, int* rp, int* ci,
                                 int* diag) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nnodes * d) return;
    const int i = r / d, c = r % d;
    const int b = rpn[i], len = rpn[i + 1] - b;
    const int start = d * d * b + c * d * len;
    rp[r] = start;
    if (r == nnodes * d - 1) rp[r + 1] = d * d * rpn[nnodes];
    for (int p = 0; p < len; ++p) {
        const int j = cin[b + p];
        for (int k = 0; k < d; ++k) ci[start + d * p + k] = d * j + k;
        if (j == i) diag[r] = start + d * p + c;
    }
}

__global__ void k_diagpos(int n, const int* rp, const int* ci, int* diag) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int pos = -1;
    for (int p = rp[r]; p < rp[r + 1]; ++p)
        if (ci[p] == r) pos = p;
    diag[r] = pos;
}

// exclusive scan of n ints in three launches: per-block sums of kScanTile
// entries, a scan of the block sums (one work-group), and the block scans
// with their offsets added; out has n + 1 entries (the total last)
constexpr int kScanTile = 4096;   // entries per work-group (256 threads x 16)
__global__ __launch_bounds__(256) void k_scan_sums(int n, const int* __restrict__ in, long long* __restrict__ sums) {
    __shared__ long long red[256];
    const long long b0 = (long long)blockIdx.x * kScanTile;
    long long s = 0;
    for (int k = threadIdx.x; k < kScanTile; k += 256)
        if (b0 + k < n) s += in[b0 + k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(1024) void k_scan_offsets(int nb, long long* sums) {
    __shared__ long long part[1024];
    const int t = threadIdx.x;
    const int chunk = (nb + 1023) / 1024, b = t * chunk, e = min(nb, b + chunk);
    long long s = 0;
    for (int i = b; i < e; ++i) s += sums[i];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        long long acc = 0;
        for (int i = 0; i < 1024; ++i) { const long long v = part[i]; part[i] = acc; acc += v; }
    }
    __syncthreads();
    long long acc = part[t];
    for (int i = b; i < e; ++i) { const long long v = sums[i]; sums[i] = acc; acc += v; }
}

// block scan: each thread scans 16 consecutive entries, the 256 thread totals
// are scanned in LDS, then written with the block's offset
__global__ __launch_bounds__(256) void k_scan_apply(int n, const int* __restrict__ in, const long long* __restrict__ offs,
                                                   int* __restrict__ out) {
    __shared__ long long part[256];
    constexpr int kPer = kScanTile / 256;
    const long long b0 = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kPer;
    int v[kPer];
    long long s = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        v[k] = b0 + k < n ? in[b0 + k] : 0;
        s += v[k];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long acc = offs[blockIdx.x];
        for (int i = 0; i < 256; ++i) { const long long x = part[i]; part[i] = acc; acc += x; }
    }
    __syncthreads();
    long long acc = part[threadIdx.x];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        if (b0 + k < n) out[b0 + k] = (int)acc;
        acc += v[k];
        if (b0 + k == n - 1) out[n] = (int)acc;
    }
}

int scan_exclusive(int n, const int* in, int* out, hipStream_t st) {
    const int nb = (n + kScanTile - 1) / kScanTile;
    if (nb == 0) return SFH_OK;
    DevBuf sums;
    SFH_TRY(sums.ensure(sizeof(long long) * nb));
    k_scan_sums<<<nb, 256, 0, st>>>(n, in, sums.as<long long>());
    SFH_LAUNCH_CHECK();
    k_scan_offsets<<<1, 1024, 0, st>>>(nb, sums.as<long long>());
    SFH_LAUNCH_CHECK();
    k_scan_apply<<<nb, 256, 0, st>>>(n, in, sums.as<long long>(), out);
    SFH_LAUNCH_CHECK();
    SFH_CHECK_HIP(hipStreamSynchronize(st));   // (sums is freed on return)
    return SFH_OK;
}

// -------------------------------------------------------------- assembly
// Scalar integrands (Poisson, Matern). One thread per node row; the row lives
// in LDS as buf[p * blockDim + tid] so that same-offset accesses of a wave hit
// distinct banks.
template <int KIND, int PHYS>
__global__ __launch_bounds__(kBlock) void k_assemble_scalar(
    int nnodes, int nq1d, const double* __restrict__ coords, const int* __restrict__ conn,
    const int* __restrict__ n2e_ptr, const int* __restrict__ n2e, const unsigned char* __restrict__ emap,
    const int* __restrict__ rowptr, const double* __restrict__ p1, const double* __restrict__ p2,
    double s1, double s2, double factor, int rowmax, double* __restrict__ vals, float* __restrict__ vals32,
    int* err) {
    constexpr int NEN = ET<KIND>::nen, D = ET<KIND>::dim;
    extern __shared__ double buf[];
    const int tid = threadIdx.x;
    const int i = blockIdx.x * blockDim.x + tid;
    if (i >= nnodes) return;
    const int r0 = rowptr[i], len = rowptr[i + 1] - r0;
    for (int p = 0; p < len; ++p) buf[p * kBlock + tid] = 0.0;
    const int nq = (KIND == SFH_TRI3) ? 1 : (D == 1 ? nq1d : (D == 2 ? nq1d * nq1d : nq1d * nq1d * nq1d));
    // kUA adjacent elements at a time: their indices, connectivity, coefficients
    // and coordinates are all requested before any is used (three dependent
    // levels of loads per element otherwise serialise element by element)
    constexpr int kUA = D == 3 ? 2 : 4;
    const int s0 = n2e_ptr[i], s1e = n2e_ptr[i + 1];
    for (int sb = s0; sb < s1e; sb += kUA) {
        int pks[kUA], cns[kUA][NEN];
#pragma unroll
        for (int u = 0; u < kUA; ++u) pks[u] = n2e[sb + u < s1e ? sb + u : s0];
#pragma unroll
        for (int u = 0; u < kUA; ++u)
#pragma unroll
            for (int K = 0; K < NEN; ++K) cns[u][K] = conn[(size_t)(pks[u] >> 3) * NEN + K];
        double Xs[kUA][NEN][D], k1s[kUA], k2s[kUA];
#pragma unroll
        for (int u = 0; u < kUA; ++u) {
            const int e = pks[u] >> 3;
            k1s[u] = p1 ? p1[e] : s1;
            k2s[u] = p2 ? p2[e] : s2;
#pragma unroll
            for (int K = 0; K < NEN; ++K)
#pragma unroll
                for (int dd = 0; dd < D; ++dd) Xs[u][K][dd] = coords[(size_t)cns[u][K] * D + dd];
        }
#pragma unroll
        for (int u = 0; u < kUA; ++u) {
            const int s = sb + u;
            if (s >= s1e) break;
            const int a = pks[u] & 7;
            const double (*X)[D] = Xs[u];
            const double k1 = k1s[u], k2 = k2s[u];
            double row[NEN];
#pragma unroll
            for (int b = 0; b < NEN; ++b) row[b] = 0.0;
            for (int q = 0; q < nq; ++q) {
                double xi[3], w, phi[NEN], g[D][NEN];
                quad_point<KIND>(q, nq1d, xi, w);
                const double detJ = gderiv<KIND>(X, xi, phi, g);
                if (!(detJ > 0.0)) atomicOr(err, 4);
                if (PHYS == SFH_POISSON) {
                    // del2/fem/ElementPotential.ipp:71-72
                    const double sc = k1 * detJ * w;
#pragma unroll
                    for (int b = 0; b < NEN; ++b) {
                        double m = 0.0;
#pragma unroll
                        for (int dd = 0; dd < D; ++dd) m += g[dd][a] * g[dd][b];
                        row[b] += sc * m;
                    }
                } else {
                    // del2/fem/ElementMatern.ipp:68-79 (scalar diffusivity k2)
#pragma unroll
                    for (int b = 0; b < NEN; ++b) {
                        double m = 0.0;
#pragma unroll
                        for (int dd = 0; dd < D; ++dd) m += (k2 * g[dd][a]) * g[dd][b];
                        row[b] += detJ * w * m;
                    }
                    row[a] += detJ * w * (k1 * k1 * phi[a]);
                }
            }
            const unsigned char* em = emap + (size_t)s * NEN;
#pragma unroll
            for (int b = 0; b < NEN; ++b) buf[em[b] * kBlock + tid] += row[b] * factor;
        }
    }
    for (int p = 0; p < len; ++p) {
        const double v = buf[p * kBlock + tid];
        vals[r0 + p] = v;
        vals32[r0 + p] = (float)v;
    }
}

// hex8 linear elasticity at F = I: one thread per DOF row (node i, component c).
// K[(a,c),(n,k)] += [lambda g_c(a) g_k(n) + mu g_k(a) g_c(n) + mu delta_ck g(a).g(n)] detJ w,
// i.e. sum_JL C_iJkL dphi_J dphi_L of solid/fem/ElementStatic.ipp:300-316 with the
// SVenant tensor of solid/material/svenant/SVenant.ipp:107-140 at F = I.
constexpr int kBlockEl = 128;
__global__ __launch_bounds__(kBlockEl) void k_assemble_elastic(
    int ndof, int nq1d, const double* __restrict__ coords, const int* __restrict__ conn,
    const int* __restrict__ n2e_ptr, const int* __restrict__ n2e, const unsigned char* __restrict__ emap,
    const int* __restrict__ rowptr, const double* __restrict__ p1, const double* __restrict__ p2,
    double s1, double s2, double factor, double* __restrict__ vals, float* __restrict__ vals32, int* err) {
    extern __shared__ double buf[];
    const int tid = threadIdx.x;
    const int r = blockIdx.x * blockDim.x + tid;
    if (r >= ndof) return;
    const int i = r / 3, c = r % 3;
    const int r0 = rowptr[r], len = rowptr[r + 1] - r0;
    for (int p = 0; p < len; ++p) buf[p * kBlockEl + tid] = 0.0;
    const int nq = nq1d * nq1d * nq1d;
    for (int s = n2e_ptr[i]; s < n2e_ptr[i + 1]; ++s) {
        const int pk = n2e[s];
        const int e = pk >> 3, a = pk & 7;
        double X[8][3];
        load_X<SFH_HEX8>(coords, conn + (size_t)e * 8, X);
        const double E = p1 ? p1[e] : s1;
        const double nu = p2 ? p2[e] : s2;
        const double lambda = E * nu / (1.0 + nu) / (1.0 - 2.0 * nu);   // SVenant.ipp:32-37
        const double mu = E / 2.0 / (1.0 + nu);
        double row[8][3];
#pragma unroll
        for (int n = 0; n < 8; ++n) row[n][0] = row[n][1] = row[n][2] = 0.0;
        for (int q = 0; q < nq; ++q) {
            double xi[3], w, phi[8], g[3][8];
            quad_point<SFH_HEX8>(q, nq1d, xi, w);
            const double detJ = gderiv<SFH_HEX8>(X, xi, phi, g);
            if (!(detJ > 0.0)) atomicOr(err, 4);
            const double dw = detJ * w;
            const double gca = g[c][a];
#pragma unroll
            for (int n = 0; n < 8; ++n) {
                const double dot = g[0][a] * g[0][n] + g[1][a] * g[1][n] + g[2][a] * g[2][n];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    double kentry = lambda * gca * g[k][n] + mu * g[k][a] * g[c][n];
                    if (k == c) kentry += mu * dot;
                    row[n][k] += kentry * dw;
                }
            }
        }
        const unsigned char* em = emap + (size_t)s * 8;
#pragma unroll
        for (int n = 0; n < 8; ++n) {
            const int pbase = 3 * em[n];
#pragma unroll
            for (int k = 0; k < 3; ++k) buf[(pbase + k) * kBlockEl + tid] += row[n][k] * factor;
        }
    }
    for (int p = 0; p < len; ++p) {
        const double v = buf[p * kBlockEl + tid];
        vals[r0 + p] = v;
        vals32[r0 + p] = (float)v;
    }
}

// ------------------------------------- hex8 elasticity in two atomic-free phases
// Phase 1 (k_elastic_ke): every element's 24 x 24 stiffness once, the same
// integrand as k_assemble_elastic. A work-group holds 32 elements, 8 threads
// each: per batch of 8 Gauss points every thread forms one point's gradients
// g = J^-T dphi and detJ w in LDS, then every thread accumulates its share of
// the 36 node pairs (a <= n; the matrix is symmetric) over the batch. The
// element matrix is stored as 36 blocks of 3 x 3, block (a, n) holding
// K[(a, c), (n, k)] at c * 3 + k.
// Phase 2 (k_elastic_gather): one thread per (node i, neighbour slot q) of the
// node-level pattern sums the blocks of the elements that couple i and its
// neighbour, in a fixed order (the contribution list of the pair, built once
// from the element adjacency: elements in mesh order, then local nodes), and
// writes the 3 x 3 block into the three DOF rows of node i, FP64 and FP32.
constexpr int kElPerBlk = 32;                  // elements per work-group of phase 1
constexpr int kKeStride = 36 * 9;              // doubles per element matrix (upper block triangle)
constexpr int kGStride = 25;                   // g[3][8] and detJ w per Gauss point in LDS

__host__ __device__ inline int hex_pair(int a, int n) { return a * 8 - a * (a - 1) / 2 + (n - a); }   // a <= n

__global__ __launch_bounds__(256) void k_elastic_ke(int nel, int nq1d, const double* __restrict__ coords,
                                                    const int* __restrict__ conn, const double* __restrict__ p1,
                                                    const double* __restrict__ p2, double s1, double s2,
                                                    double factor, double* __restrict__ ke, int* err) {
    __shared__ double sX[kElPerBlk][8][3];
    __shared__ double sg[kElPerBlk][8 * kGStride];
    const int el = threadIdx.x >> 3, t = threadIdx.x & 7;
    const int e = blockIdx.x * kElPerBlk + el;
    const bool live = e < nel;
    if (live) {
        const int nd = conn[(size_t)e * 8 + t];
#pragma unroll
        for (int d = 0; d < 3; ++d) sX[el][t][d] = coords[(size_t)nd * 3 + d];
    }
    double E = 0.0, nu = 0.0;
    if (live) {
        E = p1 ? p1[e] : s1;
        nu = p2 ? p2[e] : s2;
    }
    const double lambda = E * nu / (1.0 + nu) / (1.0 - 2.0 * nu);   // SVenant.ipp:32-37
    const double mu = E / 2.0 / (1.0 + nu);
    // this thread's node pairs p = t + 8 j (j < 5, p < 36)
    int pa[5], pn[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        int r = t + 8 * j, a = 0;
        if (r >= 36) r = 35;
        while (r >= 8 - a) { r -= 8 - a; ++a; }
        pa[j] = a;
        pn[j] = a + r;
    }
    double B[5][9];
#pragma unroll
    for (int j = 0; j < 5; ++j)
#pragma unroll
        for (int x = 0; x < 9; ++x) B[j][x] = 0.0;
    __syncthreads();
    const int nq = nq1d * nq1d * nq1d;
    for (int qb = 0; qb < nq; qb += 8) {
        const int q = qb + t;
        double* gq = &sg[el][t * kGStride];
        if (live && q < nq) {
            double X[8][3];
#pragma unroll
            for (int K = 0; K < 8; ++K)
#pragma unroll
                for (int d = 0; d < 3; ++d) X[K][d] = sX[el][K][d];
            double xi[3], w, phi[8], g[3][8];
            quad_point<SFH_HEX8>(q, nq1d, xi, w);
            const double detJ = gderiv<SFH_HEX8>(X, xi, phi, g);
            if (!(detJ > 0.0)) atomicOr(err, 4);
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int K = 0; K < 8; ++K) gq[d * 8 + K] = g[d][K];
            gq[24] = detJ * w;
        } else {
            gq[24] = 0.0;
        }
        __syncthreads();
        const int nb = min(8, nq - qb);
        for (int qq = 0; qq < nb; ++qq) {
            const double* G = &sg[el][qq * kGStride];
            const double dw = G[24];
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                const double ga[3] = {G[pa[j]], G[8 + pa[j]], G[16 + pa[j]]};
                const double gn[3] = {G[pn[j]], G[8 + pn[j]], G[16 + pn[j]]};
                const double dot = ga[0] * gn[0] + ga[1] * gn[1] + ga[2] * gn[2];
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        double v = lambda * ga[c] * gn[k] + mu * ga[k] * gn[c];
                        if (k == c) v += mu * dot;
                        B[j][c * 3 + k] += v * dw;
                    }
            }
        }
        __syncthreads();
    }
    if (!live) return;
    double* out = ke + (size_t)e * kKeStride;
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        const int p = t + 8 * j;
        if (p < 36)
#pragma unroll
            for (int x = 0; x < 9; ++x) out[p * 9 + x] = B[j][x] * factor;
    }
}

__global__ __launch_bounds__(256) void k_elastic_gather(int nnodes, const int* __restrict__ rpn,
                                                        const int* __restrict__ cptr, const int* __restrict__ cent,
                                                        const double* __restrict__ ke, const int* __restrict__ rowptr,
                                                        double* __restrict__ vals, float* __restrict__ vals32) {
    // a half-wave per node, a lane per neighbour slot; consecutive nodes
    // (which share elements) on one XCD: hardware block b runs on XCD b % 8,
    // so each XCD takes a contiguous range of nodes
    const int nb = gridDim.x;
    const int lb = (nb % 8) ? (int)blockIdx.x : ((int)blockIdx.x % 8) * (nb / 8) + (int)blockIdx.x / 8;
    const int i = lb * 8 + (threadIdx.x >> 5), lane = threadIdx.x & 31;
    if (i >= nnodes) return;
    const int b = rpn[i], len = rpn[i + 1] - b;
    const int r0 = rowptr[3 * i], rstride = 3 * len;   // rows 3i + c start at r0 + c * 3 len
    for (int q = lane; q < len; q += 32) {
        const int t = b + q;
        double acc[9];
#pragma unroll
        for (int x = 0; x < 9; ++x) acc[x] = 0.0;
        // kUC contributions at a time: their codes, then all their blocks, are
        // requested before any is added (a block (n, a) with n < a is read and
        // used transposed: K[(a,c),(n,k)] = K[(n,k),(a,c)])
        constexpr int kUC = 4;
        const int c0 = cptr[t], c1 = cptr[t + 1];
        for (int cb = c0; cb < c1; cb += kUC) {
            int code[kUC];
#pragma unroll
            for (int u = 0; u < kUC; ++u) code[u] = cent[cb + u < c1 ? cb + u : c0];
            double v[kUC][9];
            bool tr[kUC];
#pragma unroll
            for (int u = 0; u < kUC; ++u) {
                const int e = code[u] >> 6, a = (code[u] >> 3) & 7, n = code[u] & 7;
                tr[u] = a > n;
                const double* Bk = ke + (size_t)e * kKeStride + (tr[u] ? hex_pair(n, a) : hex_pair(a, n)) * 9;
#pragma unroll
                for (int x = 0; x < 9; ++x) v[u][x] = Bk[x];
            }
#pragma unroll
            for (int u = 0; u < kUC; ++u) {
                if (cb + u >= c1) break;
#pragma unroll
                for (int x = 0; x < 9; ++x) acc[x] += tr[u] ? v[u][(x % 3) * 3 + x / 3] : v[u][x];
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int pos = r0 + c * rstride + 3 * q;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                vals[pos + k] = acc[c * 3 + k];
                vals32[pos + k] = (float)acc[c * 3 + k];
            }
        }
    }
}

// The same two phases for the scalar integrands on hex8 meshes (Poisson,
// Matern; the integrands of k_assemble_scalar): the element matrix is the
// 36 entries K[a][n], a <= n, and the gather writes one CSR entry per thread.
constexpr int kGStrideS = 33;                  // g[3][8], phi[8] and detJ w per Gauss point in LDS
template <int PHYS>
__global__ __launch_bounds__(256) void k_hex_scalar_ke(int nel, int nq1d, const double* __restrict__ coords,
                                                       const int* __restrict__ conn, const double* __restrict__ p1,
                                                       const double* __restrict__ p2, double s1, double s2,
                                                       double factor, double* __restrict__ ke, int* err) {
    __shared__ double sX[kElPerBlk][8][3];
    __shared__ double sg[kElPerBlk][8 * kGStrideS];
    const int el = threadIdx.x >> 3, t = threadIdx.x & 7;
    const int e = blockIdx.x * kElPerBlk + el;
    const bool live = e < nel;
    if (live) {
        const int nd = conn[(size_t)e * 8 + t];
#pragma unroll
        for (int d = 0; d < 3; ++d) sX[el][t][d] = coords[(size_t)nd * 3 + d];
    }
    const double k1 = live ? (p1 ? p1[e] : s1) : 0.0;
    const double k2 = live ? (p2 ? p2[e] : s2) : 0.0;
    int pa[5], pn[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        int r = t + 8 * j, a = 0;
        if (r >= 36) r = 35;
        while (r >= 8 - a) { r -= 8 - a; ++a; }
        pa[j] = a;
        pn[j] = a + r;
    }
    double B[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    __syncthreads();
    const int nq = nq1d * nq1d * nq1d;
    for (int qb = 0; qb < nq; qb += 8) {
        const int q = qb + t;
        double* gq = &sg[el][t * kGStrideS];
        if (live && q < nq) {
            double X[8][3];
#pragma unroll
            for (int K = 0; K < 8; ++K)
#pragma unroll
                for (int d = 0; d < 3; ++d) X[K][d] = sX[el][K][d];
            double xi[3], w, phi[8], g[3][8];
            quad_point<SFH_HEX8>(q, nq1d, xi, w);
            const double detJ = gderiv<SFH_HEX8>(X, xi, phi, g);
            if (!(detJ > 0.0)) atomicOr(err, 4);
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int K = 0; K < 8; ++K) gq[d * 8 + K] = g[d][K];
#pragma unroll
            for (int K = 0; K < 8; ++K) gq[24 + K] = phi[K];
            gq[32] = detJ * w;
        } else {
            gq[32] = 0.0;
        }
        __syncthreads();
        const int nb = min(8, nq - qb);
        for (int qq = 0; qq < nb; ++qq) {
            const double* G = &sg[el][qq * kGStrideS];
            const double dw = G[32];
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                const double dot = G[pa[j]] * G[pn[j]] + G[8 + pa[j]] * G[8 + pn[j]] + G[16 + pa[j]] * G[16 + pn[j]];
                if (PHYS == SFH_POISSON) {
                    B[j] += k1 * dw * dot;                                   // ElementPotential.ipp:71-72
                } else {
                    B[j] += dw * (k2 * dot);                                 // ElementMatern.ipp:68-79
                    if (pa[j] == pn[j]) B[j] += dw * (k1 * k1 * G[24 + pa[j]]);
                }
            }
        }
        __syncthreads();
    }
    if (!live) return;
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        const int p = t + 8 * j;
        if (p < 36) ke[(size_t)e * 36 + p] = B[j] * factor;
    }
}

__global__ __launch_bounds__(256) void k_hex_scalar_gather(int nnz, const int* __restrict__ cptr,
                                                           const int* __restrict__ cent, const double* __restrict__ ke,
                                                           double* __restrict__ vals, float* __restrict__ vals32) {
    const int nb = gridDim.x;
    const int lb = (nb % 8) ? (int)blockIdx.x : ((int)blockIdx.x % 8) * (nb / 8) + (int)blockIdx.x / 8;
    const int t = lb * blockDim.x + threadIdx.x;
    if (t >= nnz) return;
    double acc = 0.0;
    constexpr int kUC = 4;   // contributions requested together before any is added
    const int c0 = cptr[t], c1 = cptr[t + 1];
    for (int cb = c0; cb < c1; cb += kUC) {
        double v[kUC];
#pragma unroll
        for (int u = 0; u < kUC; ++u) {
            const int code = cent[cb + u < c1 ? cb + u : c0];
            const int e = code >> 6, a = (code >> 3) & 7, n = code & 7;
            v[u] = ke[(size_t)e * 36 + (a <= n ? hex_pair(a, n) : hex_pair(n, a))];
        }
#pragma unroll
        for (int u = 0; u < kUC; ++u)
            if (cb + u < c1) acc += v[u];
    }
    vals[t] = acc;
    vals32[t] = (float)acc;
}

// contribution lists of the node-level pattern: for node i's slot q, the
// (element, local index of i, local index of the neighbour) triples, counted
// (fill = 0) or written at cptr (fill = 1, cur = zeroed cursors)
__global__ void k_elastic_contrib(int nnodes, int fill, const int* __restrict__ n2e_ptr, const int* __restrict__ n2e,
                                  const unsigned char* __restrict__ emap, const int* __restrict__ rpn,
                                  const int* __restrict__ cptr, int* cnt, int* cent) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnodes) return;
    const int b = rpn[i];
    for (int s = n2e_ptr[i]; s < n2e_ptr[i + 1]; ++s) {
        const int pk = n2e[s];
        const int e = pk >> 3, a = pk & 7;
        for (int n = 0; n < 8; ++n) {
            const int slot = b + emap[(size_t)s * 8 + n];
            if (fill) cent[cptr[slot] + cnt[slot]] = (e << 6) | (a << 3) | n;
            ++cnt[slot];
        }
    }
}

// Body force (del2/fem/ElementPotential.ipp:146-175) gathered per node:
// F[i*d+c] += w detJ factor f_c(x_q) phi_a(xi_q), f interpolated from nodal values.
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_body_force(
    int nnodes, int nq1d, int dofpn, const double* __restrict__ coords, const int* __restrict__ conn,
    const int* __restrict__ n2e_ptr, const int* __restrict__ n2e, double f0, double f1, double f2,
    const double* __restrict__ fnodal, double factor, int accumulate, double* __restrict__ F, int* err) {
    constexpr int NEN = ET<KIND>::nen, D = ET<KIND>::dim;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnodes) return;
    const int nq = (KIND == SFH_TRI3) ? 1 : (D == 1 ? nq1d : (D == 2 ? nq1d * nq1d : nq1d * nq1d * nq1d));
    double acc[3] = {0.0, 0.0, 0.0};
    for (int s = n2e_ptr[i]; s < n2e_ptr[i + 1]; ++s) {
        const int pk = n2e[s];
        const int e = pk >> 3, a = pk & 7;
        const int* cn = conn + (size_t)e * NEN;
        double X[NEN][D];
        load_X<KIND>(coords, cn, X);
        for (int q = 0; q < nq; ++q) {
            double xi[3], w, phi[NEN], g[D][NEN];
            quad_point<KIND>(q, nq1d, xi, w);
            const double detJ = gderiv<KIND>(X, xi, phi, g);
            if (!(detJ > 0.0)) atomicOr(err, 4);
            for (int c = 0; c < dofpn; ++c) {
                double f = (c == 0) ? f0 : (c == 1 ? f1 : f2);
                if (fnodal) {
#pragma unroll
                    for (int K = 0; K < NEN; ++K) f += fnodal[(size_t)cn[K] * dofpn + c] * phi[K];
                }
                acc[c] += w * detJ * factor * (f * phi[a]);
            }
        }
    }
    for (int c = 0; c < dofpn; ++c) {
        const size_t o = (size_t)i * dofpn + c;
        F[o] = accumulate ? F[o] + acc[c] : acc[c];
    }
}

// corlib/Triplet.hpp:219-260 zeroRow + SystemSolveEigenSparse.hpp:208-210
__global__ void k_constraints(int nc, const int* __restrict__ dofs, const double* __restrict__ values,
                              double factor, double scalar, const int* __restrict__ rowptr,
                              const int* __restrict__ colind, double* vals, float* vals32, double* F,
                              unsigned char* cmask) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nc) return;
    const int r = dofs[t];
    for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) {
        const double v = (colind[p] == r) ? scalar : 0.0;
        vals[p] = v;
        vals32[p] = (float)v;
    }
    F[r] = factor * (values ? values[t] : 0.0) * scalar;
    cmask[r] = 1;
}

// Y = A X for k right-hand sides (FP64, row-zero operator as assembled)
__global__ void k_spmv64(int n, int k, const int* __restrict__ rp, const int* __restrict__ ci,
                         const double* __restrict__ va, const double* __restrict__ X, double* __restrict__ Y) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * k) return;
    const int i = (int)(t / k), c = (int)(t % k);
    double s = 0.0;
    for (int e = rp[i]; e < rp[i + 1]; ++e) s += va[e] * X[(size_t)ci[e] * k + c];
    Y[t] = s;
}

// Y = s X ./ m (row scaling)
__global__ void k_rowdiv(long long total, int k, double s, const double* __restrict__ m,
                         const double* __restrict__ X, double* __restrict__ Y) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    Y[t] = m ? s * X[t] / m[t / k] : s * X[t];
}

inline int grid_for(long long n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

// ------------------------------------------------------------------ host
int build_pattern(sfh_system* s) {
    hipStream_t st = s->ctx->stream;
    const int N = s->nnodes, nen = s->nen;
    const long long nadj = (long long)s->nel * nen;
    DevBuf cnt, fill, err;
    SFH_TRY(cnt.ensure(sizeof(int) * (N + 1)));
    SFH_TRY(fill.ensure(sizeof(int) * (N + 1)));
    SFH_TRY(err.ensure(sizeof(int)));
    SFH_TRY(s->n2e_ptr.ensure(sizeof(int) * (N + 1)));
    SFH_TRY(s->n2e.ensure(sizeof(int) * nadj));
    SFH_TRY(s->emap.ensure(nadj * nen));
    SFH_CHECK_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int) * (N + 1), st));
    SFH_CHECK_HIP(hipMemsetAsync(fill.p, 0, sizeof(int) * (N + 1), st));
    SFH_CHECK_HIP(hipMemsetAsync(err.p, 0, sizeof(int), st));
    const int* conn = s->conn.as<int>();
    k_adj_count<<<grid_for(nadj, kBlock), kBlock, 0, st>>>(s->nel, nen, conn, cnt.as<int>());
    SFH_LAUNCH_CHECK();
    SFH_TRY(scan_exclusive(N, cnt.as<int>(), s->n2e_ptr.as<int>(), st));
    SFH_LAUNCH_CHECK();
    k_adj_fill<<<grid_for(nadj, kBlock), kBlock, 0, st>>>(s->nel, nen, conn, s->n2e_ptr.as<int>(),
                                                          fill.as<int>(), s->n2e.as<int>());
    SFH_LAUNCH_CHECK();
    k_adj_sort<<<grid_for(N, kBlock), kBlock, 0, st>>>(N, s->n2e_ptr.as<int>(), s->n2e.as<int>(),
                                                       err.as<int>());
    SFH_LAUNCH_CHECK();
    k_row_count<<<grid_for(N, 128), 128, 0, st>>>(N, nen, conn, s->n2e_ptr.as<int>(),
                                                  s->n2e.as<int>(), cnt.as<int>(), err.as<int>());
    SFH_LAUNCH_CHECK();
    SFH_TRY(s->rowptr_node.ensure(sizeof(int) * (N + 1)));
    SFH_TRY(scan_exclusive(N, cnt.as<int>(), s->rowptr_node.as<int>(), st));
    SFH_LAUNCH_CHECK();
    int nnz_node = 0, herr = 0;
    SFH_CHECK_HIP(hipMemcpyAsync(&nnz_node, s->rowptr_node.as<int>() + N, sizeof(int),
                                 hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    if (herr) {
        set_error("mesh exceeds adjacency limits (%d elements per node, %d coupled nodes)", kMaxAdj,
                  kMaxCand);
        return SFH_EARG;
    }
    s->nnz_node = nnz_node;
    SFH_TRY(s->colind_node.ensure(sizeof(int) * (size_t)nnz_node));
    k_row_fill<<<grid_for(N, 128), 128, 0, st>>>(N, nen, conn, s->n2e_ptr.as<int>(), s->n2e.as<int>(),
                                                 s->rowptr_node.as<int>(), s->colind_node.as<int>(),
                                                 s->emap.as<unsigned char>(), err.as<int>());
    SFH_LAUNCH_CHECK();
    const int d = s->dofpn;
    s->ndof = N * d;
    s->nnz = (long long)nnz_node * d * d;
    if (s->nnz > 0x7fffffffLL) {
        set_error("nnz %lld exceeds 32-bit indexing", s->nnz);
        return SFH_EARG;
    }
    SFH_TRY(s->diagpos.ensure(sizeof(int) * s->ndof));
    if (d == 1) {
        SFH_TRY(s->rowptr.ensure(sizeof(int) * (N + 1)));
        SFH_TRY(s->colind.ensure(sizeof(int) * (size_t)nnz_node));
        SFH_CHECK_HIP(hipMemcpyAsync(s->rowptr.p, s->rowptr_node.p, sizeof(int) * (N + 1),
                                     hipMemcpyDeviceToDevice, st));
        SFH_CHECK_HIP(hipMemcpyAsync(s->colind.p, s->colind_node.p, sizeof(int) * (size_t)nnz_node,
                                     hipMemcpyDeviceToDevice, st));
        k_diagpos<<<grid_for(N, kBlock), kBlock, 0, st>>>(N, s->rowptr.as<int>(), s->colind.as<int>(),
                                                          s->diagpos.as<int>());
        SFH_LAUNCH_CHECK();
    } else {
        SFH_TRY(s->rowptr.ensure(sizeof(int) * (s->ndof + 1)));
        SFH_TRY(s->colind.ensure(sizeof(int) * (size_t)s->nnz));
        k_expand_pattern<<<grid_for(s->ndof, kBlock), kBlock, 0, st>>>(
            N, d, s->rowptr_node.as<int>(), s->colind_node.as<int>(), s->rowptr.as<int>(),
            s->colind.as<int>(), s->diagpos.as<int>());
        SFH_LAUNCH_CHECK();
    }
    // longest row bounds the LDS row buffer of the assembly kernels
    std::vector<int> rp(s->ndof + 1);
    SFH_CHECK_HIP(hipMemcpyAsync(rp.data(), s->rowptr.p, sizeof(int) * (s->ndof + 1),
                                 hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    if (herr) {
        set_error("pattern build failed (code %d)", herr);
        return SFH_EARG;
    }
    int mx = 0;
    for (int r = 0; r < s->ndof; ++r) {
        if (rp[r + 1] - rp[r] <= 0) {
            set_error("DOF %d belongs to no element", r);
            return SFH_EARG;
        }
        mx = std::max(mx, rp[r + 1] - rp[r]);
    }
    s->maxrow = mx;
    SFH_TRY(s->vals.ensure(sizeof(double) * (size_t)s->nnz));
    SFH_TRY(s->vals32.ensure(sizeof(float) * (size_t)s->nnz));
    SFH_TRY(s->F.ensure(sizeof(double) * s->ndof));
    SFH_TRY(s->cmask.ensure(s->ndof));
    SFH_TRY(s->dinv32.ensure(sizeof(float) * s->ndof));
    SFH_TRY(s->dinv64.ensure(sizeof(double) * s->ndof));
    SFH_TRY(s->diag32.ensure(sizeof(float) * s->ndof));
    SFH_TRY(s->diag64.ensure(sizeof(double) * s->ndof));
    SFH_TRY(s->cv32.ensure(sizeof(int2) * (size_t)s->nnz));
    SFH_CHECK_HIP(hipMemsetAsync(s->vals.p, 0, sizeof(double) * (size_t)s->nnz, st));
    SFH_CHECK_HIP(hipMemsetAsync(s->vals32.p, 0, sizeof(float) * (size_t)s->nnz, st));
    SFH_CHECK_HIP(hipMemsetAsync(s->F.p, 0, sizeof(double) * s->ndof, st));
    SFH_CHECK_HIP(hipMemsetAsync(s->cmask.p, 0, s->ndof, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    return SFH_OK;
}

// -------------------- quad4 on a declared 2D grid: closed-form cell integrals
// The 2D case of the hex8 kernel below: rectangles, N_a = l_a(x) l_a(y),
// grad N_i . grad N_j integrated as D_x M_y + M_x D_y (D = int l_i' l_j' =
// +-1/h, M = int l_i l_j = h/3 or h/6), exact for Gauss rules of two or more
// points per axis. One thread per node forms its row (the 9-point
// neighbourhood, each entry summed over the shared cells in the order y, x of
// the cell) in LDS; the work-group's nodes (kQgB consecutive nodes of the
// grid, across line ends) own one contiguous range of the CSR values, written
// out coalesced, FP64 and FP32.
// Poisson: p1 grad . grad; Matern: p2 grad . grad plus p1^2 int N_i on the
// diagonal (the lumped mass, a quarter of the cell's area).
constexpr int kQgB = 256;
template <int PHYS>
__global__ __launch_bounds__(kQgB) void k_quad_grid(int nx, int ny, const double* __restrict__ gx,
                                                    const double* __restrict__ gy, const int* __restrict__ cell,
                                                    const double* __restrict__ p1, const double* __restrict__ p2,
                                                    double s1, double s2, double factor,
                                                    const int* __restrict__ rowptr, double* __restrict__ vals,
                                                    float* __restrict__ vals32, int* err) {
    __shared__ double rows[kQgB * 9];
    const int i0 = blockIdx.x * kQgB, nl = min(kQgB, nx * ny - i0);
    const int base = rowptr[i0], end = rowptr[i0 + nl];
    const int ix = (i0 + (int)threadIdx.x) % nx, iy = (i0 + (int)threadIdx.x) / nx;
    if (threadIdx.x < nl) {
        const int ii[2] = {ix, iy}, n[2] = {nx, ny};
        const double* g[2] = {gx, gy};
        // per axis: the cell below (0) and above (1) the node, whether it
        // exists, its width and reciprocal width
        bool has[2][2];
        double h[2][2], r[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int c = ii[a] - 1 + u;
                has[a][u] = c >= 0 && c <= n[a] - 2;
                const int cc = has[a][u] ? c : 0;
                h[a][u] = g[a][cc + 1] - g[a][cc];
                if (has[a][u] && !(h[a][u] > 0.0)) atomicOr(err, 4);
                r[a][u] = 1.0 / h[a][u];
            }
        double k1[2][2], k2[2][2];   // material of cell (y slot, x slot)
#pragma unroll
        for (int uy = 0; uy < 2; ++uy)
#pragma unroll
            for (int ux = 0; ux < 2; ++ux) {
                const bool in = has[1][uy] && has[0][ux];
                const int c = in ? (iy - 1 + uy) * (nx - 1) + ix - 1 + ux : 0;
                const int e = cell ? cell[c] : c;
                k1[uy][ux] = p1 ? p1[e] : s1;
                k2[uy][ux] = p2 ? p2[e] : s2;
            }
        double* out = rows + (rowptr[ix + nx * iy] - base);
        int q = 0;
#pragma unroll
        for (int oy = -1; oy <= 1; ++oy)
#pragma unroll
            for (int ox = -1; ox <= 1; ++ox) {
                const int off[2] = {ox, oy};
                bool inr = true;
#pragma unroll
                for (int a = 0; a < 2; ++a) inr = inr && ii[a] + off[a] >= 0 && ii[a] + off[a] < n[a];
                if (!inr) continue;
                double acc = 0.0;
#pragma unroll
                for (int uy = 0; uy < 2; ++uy)
#pragma unroll
                    for (int ux = 0; ux < 2; ++ux) {
                        // cell slot u shares the pair along an axis when the
                        // neighbour lies in it: offset -1 needs the cell below,
                        // +1 the cell above, 0 either
                        const int u[2] = {ux, uy};
                        bool sh = true;
                        double D[2], Mm[2];
#pragma unroll
                        for (int a = 0; a < 2; ++a) {
                            sh = sh && has[a][u[a]] && (off[a] == 0 || off[a] == 2 * u[a] - 1);
                            const bool same = off[a] == 0;   // node and neighbour at the same end of the cell
                            D[a] = same ? r[a][u[a]] : -r[a][u[a]];
                            Mm[a] = same ? h[a][u[a]] * (1.0 / 3.0) : h[a][u[a]] * (1.0 / 6.0);
                        }
                        if (!sh) continue;
                        const double tr = D[0] * Mm[1] + Mm[0] * D[1];
                        if constexpr (PHYS == SFH_POISSON) {
                            acc += k1[uy][ux] * tr;                                  // ElementPotential.ipp:71-72
                        } else {
                            acc += k2[uy][ux] * tr;                                  // ElementMatern.ipp:68-79
                            if (ox == 0 && oy == 0)
                                acc += k1[uy][ux] * k1[uy][ux] * (h[0][ux] * h[1][uy] * 0.25);
                        }
                    }
                out[q++] = acc;
            }
    }
    __syncthreads();
    for (int p = threadIdx.x; p < end - base; p += kQgB) {
        const double v = rows[p] * factor;
        vals[(size_t)base + p] = v;
        vals32[(size_t)base + p] = (float)v;
    }
}

// -------------------- hex8 on a declared 3D grid: closed-form cell integrals
// When the elements are the cells of a declared tensor grid, every element is
// an axis-aligned box and its trilinear shape functions factor per axis,
// N_a = l_a(x) l_a(y) l_a(z), so every integral of the element forms is a
// product of three 1D integrals over the cell's widths h:
//   int l_i l_j = h/3 (same end of the cell) or h/6,   int l_i' l_j' = +-1/h,
//   int l_i' l_j = -1/2 (i at the low end) or +1/2,
// the values that Gauss quadrature with two or more points per axis gives
// (the integrands are at most quadratic per axis). One thread per (node,
// neighbour offset) sums the 3 x 3 block (elasticity) or the entry (scalar
// forms) over the cells the two nodes share (1, 2, 4 or 8, in the order z, y,
// x of the cell) and writes it into the node's rows, FP64 and FP32: no element
// matrix goes through memory. Integrands as in k_elastic_ke and
// k_hex_scalar_ke (elasticity: lambda d_c N_i d_k N_j + mu d_k N_i d_c N_j +
// delta_ck mu grad N_i . grad N_j; Poisson: p1 grad . grad; Matern: p2 grad .
// grad plus p1^2 int N_i on the diagonal).
// A work-group takes kHgN consecutive nodes of one grid line. Phase 0: the
// material and the widths (and their reciprocals) of the cells around them.
// Phase 1: one thread per (node, cell around it, corner of that cell) forms
// the 3 x 3 block (or entry) of the node and the corner over the cell, in LDS;
// every lane of a wave does the same work. Phase 2: one thread per (node,
// neighbour offset) sums the blocks of the cells the two nodes share, in a
// fixed order (cell z, y, x), and writes them into the node's rows (staging
// the rows in LDS for contiguous stores was measured slower: the extra LDS
// halves the work-groups per CU).
// nodes per work-group: 8 for elasticity (64 blocks of 9 doubles per node in
// LDS), 32 for the scalar forms
template <int PHYS>
constexpr int hex_grid_nodes() { return PHYS == SFH_ELASTICITY ? 8 : 32; }
template <int PHYS>
__global__ __launch_bounds__(256) void k_hex_grid(int nx, int ny, int nz, const double* __restrict__ gx,
                                                  const double* __restrict__ gy, const double* __restrict__ gz,
                                                  const int* __restrict__ cell, const double* __restrict__ p1,
                                                  const double* __restrict__ p2, double s1, double s2, double factor,
                                                  const int* __restrict__ rowptr, double* __restrict__ vals,
                                                  float* __restrict__ vals32, int* err) {
    constexpr int NV = PHYS == SFH_ELASTICITY ? 9 : 1;
    constexpr int kHgN = hex_grid_nodes<PHYS>();
    constexpr int kCx = kHgN + 1;                         // cells along x around the work-group's nodes
    __shared__ double blk[kHgN * 64 * NV];               // (node, cell slot, corner) -> block
    __shared__ double mat[2][2][2][kCx];                 // (lambda, mu) or (p1, p2) of cell (z, y, x slot)
    __shared__ double hw[kCx + 4], rw[kCx + 4];          // widths, reciprocals: x cells, then y, z pairs
    const int n[3] = {nx, ny, nz};
    const int lx0 = blockIdx.x * kHgN, nl = min(kHgN, nx - lx0);
    const int iy = blockIdx.y, iz = blockIdx.z;
    const int tid = threadIdx.x;
    // phase 0: cell (lx0 - 1 + x slot, iy - 1 + y slot, iz - 1 + z slot)
    if (tid < 4 * kCx) {
        const int xs = tid % kCx, ys = (tid / kCx) & 1, zs = tid / (2 * kCx);
        const int cx = lx0 - 1 + xs, cy = iy - 1 + ys, cz = iz - 1 + zs;
        double m0 = 0.0, m1 = 0.0;
        if (cx >= 0 && cx <= nx - 2 && cy >= 0 && cy <= ny - 2 && cz >= 0 && cz <= nz - 2) {
            const int c = (int)(((size_t)cz * (ny - 1) + cy) * (nx - 1) + cx);
            const int e = (p1 || p2) && cell ? cell[c] : c;
            const double a = p1 ? p1[e] : s1, b = p2 ? p2[e] : s2;
            if constexpr (PHYS == SFH_ELASTICITY) {
                m0 = a * b / (1.0 + b) / (1.0 - 2.0 * b);   // SVenant.ipp:32-37
                m1 = a / 2.0 / (1.0 + b);
            } else {
                m0 = a;
                m1 = b;
            }
        }
        mat[0][zs][ys][xs] = m0;
        mat[1][zs][ys][xs] = m1;
    } else if (tid < 4 * kCx + kCx + 4) {
        const int w = tid - 4 * kCx;
        const int ax = w < kCx ? 0 : (w < kCx + 2 ? 1 : 2);
        const int c = ax == 0 ? lx0 - 1 + w : (ax == 1 ? iy - 1 + (w - kCx) : iz - 1 + (w - kCx - 2));
        const double* g = ax == 0 ? gx : (ax == 1 ? gy : gz);
        double h = 0.0, r = 0.0;
        if (c >= 0 && c <= n[ax] - 2) {
            h = g[c + 1] - g[c];
            if (!(h > 0.0)) atomicOr(err, 4);
            r = 1.0 / h;
        }
        hw[w] = h;
        rw[w] = r;
    }
    // the row start of this thread's first phase-2 node, loaded before the
    // phases that only wait on LDS
    const int rp_first = tid < nl * 27
                             ? rowptr[(NV == 9 ? 3 : 1) * (lx0 + tid / 27 + nx * (iy + ny * iz))]
                             : 0;
    __syncthreads();
    // phase 1: node il, cell slot cs (bit a: the cell above the node along
    // axis a), corner cn (bit a: the cell's upper node along a)
    for (int b = tid; b < nl * 64; b += blockDim.x) {
        const int il = b >> 6, cs = (b >> 3) & 7, cn = b & 7;
        const int ix = lx0 + il;
        const int sl[3] = {cs & 1, (cs >> 1) & 1, cs >> 2};
        const int ii[3] = {ix, iy, iz};
        bool ok = true;
        double M[3], D[3], Gi[3], Gj[3], vol = 1.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int c = ii[a] - 1 + sl[a];
            ok = ok && c >= 0 && c <= n[a] - 2;
            const int w = a == 0 ? il + sl[0] : kCx + 2 * (a - 1) + sl[a];
            const double h = hw[w], r = rw[w];
            const int si = 1 - sl[a], tj = (cn >> a) & 1;   // the node's and the corner's end of the cell
            M[a] = si == tj ? h * (1.0 / 3.0) : h * (1.0 / 6.0);
            D[a] = si == tj ? r : -r;
            Gi[a] = si ? 0.5 : -0.5;
            Gj[a] = tj ? 0.5 : -0.5;
            vol *= h;
        }
        if (!ok) continue;
        // I[c][k] = int d_c N_i d_k N_j: the derivative axes' stiffness or
        // mixed factors times the other axes' mass
        double I[3][3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int k = 0; k < 3; ++k)
                I[c][k] = c == k ? D[c] * M[(c + 1) % 3] * M[(c + 2) % 3] : Gi[c] * Gj[k] * M[3 - c - k];
        const double tr = I[0][0] + I[1][1] + I[2][2];
        const double m0 = mat[0][sl[2]][sl[1]][il + sl[0]], m1 = mat[1][sl[2]][sl[1]][il + sl[0]];
        double* out = blk + (size_t)b * NV;
        if constexpr (PHYS == SFH_ELASTICITY) {
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    double v = m0 * I[c][k] + m1 * I[k][c];
                    if (k == c) v += m1 * tr;
                    out[c * 3 + k] = v;
                }
        } else if constexpr (PHYS == SFH_POISSON) {
            out[0] = m0 * tr;                                                   // ElementPotential.ipp:71-72
        } else {
            double v = m1 * tr;                                                 // ElementMatern.ipp:68-79
            if (cn == 7 - cs) v += m0 * m0 * (vol * 0.125);                     // the node itself: lumped mass
            out[0] = v;
        }
    }
    __syncthreads();
    // phase 2: node il, neighbour offset o (x fastest)
    for (int b = tid; b < nl * 27; b += blockDim.x) {
        const int il = b / 27, o = b - 27 * il;
        const int ix = lx0 + il;
        const int ii[3] = {ix, iy, iz};
        const int off[3] = {o % 3 - 1, (o / 3) % 3 - 1, o / 9 - 1};
        int cnt[3], bef[3], s0[3], s1v[3];
        bool in = true;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int j = ii[a] + off[a];
            in = in && j >= 0 && j < n[a];
            cnt[a] = (ii[a] > 0) + 1 + (ii[a] < n[a] - 1);      // neighbour offsets in range on this axis
            bef[a] = (off[a] >= 0 && ii[a] > 0) + (off[a] >= 1);  // of them, those below off[a]
            // cell slots sharing the pair along a: {off} for off = -1 / +1 (slot 0 / 1), both for 0,
            // limited to cells inside the grid
            s0[a] = off[a] > 0 ? 1 : 0;
            s1v[a] = off[a] < 0 ? 0 : 1;
            if (ii[a] == 0) s0[a] = max(s0[a], 1);
            if (ii[a] == n[a] - 1) s1v[a] = min(s1v[a], 0);
        }
        if (!in) continue;
        double acc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = 0.0;
        for (int sz = s0[2]; sz <= s1v[2]; ++sz)
            for (int sy = s0[1]; sy <= s1v[1]; ++sy)
                for (int sx = s0[0]; sx <= s1v[0]; ++sx) {
                    const int cs = sx | (sy << 1) | (sz << 2);
                    const int cn = (off[0] + 1 - sx) | ((off[1] + 1 - sy) << 1) | ((off[2] + 1 - sz) << 2);
                    const double* src = blk + (size_t)((il * 8 + cs) * 8 + cn) * NV;
#pragma unroll
                    for (int v = 0; v < NV; ++v) acc[v] += src[v];
                }
        const int i = ix + nx * (iy + ny * iz);
        const int len = cnt[0] * cnt[1] * cnt[2];
        const int q = (bef[2] * cnt[1] + bef[1]) * cnt[0] + bef[0];   // the neighbour's slot in the sorted row
        const int rs = b == tid ? rp_first : rowptr[(NV == 9 ? 3 : 1) * i];
        if constexpr (PHYS == SFH_ELASTICITY) {
            const size_t r0 = (size_t)rs;   // rows 3i + c start at r0 + c * 3 len
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const size_t pos = r0 + (size_t)c * 3 * len + 3 * q + k;
                    const double v = acc[c * 3 + k] * factor;
                    vals[pos] = v;
                    vals32[pos] = (float)v;
                }
        } else {
            const size_t pos = (size_t)rs + q;
            const double v = acc[0] * factor;
            vals[pos] = v;
            vals32[pos] = (float)v;
        }
    }
}

// Whether the elements are the cells of the declared grid (quad4 on a 2D
// grid, hex8 on a 3D grid; every cell exactly once, in any element order,
// local nodes in the reference order, i.e. the orientation the quadrature
// integrates with a positive Jacobian) and every node's row holds exactly its
// in-range 9- or 27-point neighbours in increasing order: then k_quad_grid /
// k_hex_grid apply, with the cell -> element table on the device (none when
// the elements are in cell order). Checked once per grid declaration, on the
// device. Per element: its corners must be one cell's in the reference order
// (counter-clockwise, then the top face of a hex8: dx + 2 dy + 4 dz of local
// node k), and no two elements may share a cell. flags: 1 not a grid mesh,
// 2 element order differs from the cell order. cell[] starts at -1.
__global__ void k_grid_cells(int nel, int nen, int nnodes, int nx, int ny, int nz, const int* __restrict__ conn,
                             int* __restrict__ cell, int* flags) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nel) return;
    constexpr int kCorner[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    int lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX};
    int xyz[8][3];
    for (int k = 0; k < nen; ++k) {
        const int v = conn[(size_t)e * nen + k];
        if (v < 0 || v >= nnodes) { atomicOr(flags, 1); return; }
        xyz[k][0] = v % nx;
        xyz[k][1] = (v / nx) % ny;
        xyz[k][2] = v / (nx * ny);
        for (int a = 0; a < 3; ++a) lo[a] = min(lo[a], xyz[k][a]);
    }
    for (int k = 0; k < nen; ++k) {
        const int dx = xyz[k][0] - lo[0], dy = xyz[k][1] - lo[1], dz = xyz[k][2] - lo[2];
        if (dx > 1 || dy > 1 || dz > 1 || dx + 2 * dy + 4 * dz != kCorner[k]) { atomicOr(flags, 1); return; }
    }
    if (lo[0] > nx - 2 || lo[1] > ny - 2 || (nz > 1 && lo[2] > nz - 2)) { atomicOr(flags, 1); return; }
    const size_t c = ((size_t)lo[2] * (ny - 1) + lo[1]) * (nx - 1) + lo[0];
    if (atomicCAS(cell + c, -1, e) != -1) { atomicOr(flags, 1); return; }
    if ((size_t)e != c) atomicOr(flags, 2);
}

// the node pattern must list each node's in-range 9- or 27-point neighbours
// in increasing order (flags |= 1 otherwise)
__global__ void k_grid_pattern(int nx, int ny, int nz, const int* __restrict__ rpn, const int* __restrict__ cin,
                               int* flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx * ny * nz) return;
    const int ix = i % nx, iy = (i / nx) % ny, iz = i / (nx * ny);
    const int oz0 = nz > 1 ? -1 : 0, oz1 = nz > 1 ? 1 : 0;
    int p = rpn[i];
    const int pe = rpn[i + 1];
    for (int oz = oz0; oz <= oz1; ++oz)
        for (int oy = -1; oy <= 1; ++oy)
            for (int ox = -1; ox <= 1; ++ox) {
                const int jx = ix + ox, jy = iy + oy, jz = iz + oz;
                if (jx < 0 || jx >= nx || jy < 0 || jy >= ny || jz < 0 || jz >= nz) continue;
                if (p >= pe || cin[p] != jx + nx * (jy + ny * jz)) { atomicOr(flags, 1); return; }
                ++p;
            }
    if (p != pe) atomicOr(flags, 1);
}

int grid_cells_setup(sfh_system* s, hipStream_t st) {
    if (s->gcell_state != 0) return SFH_OK;
    s->gcell_state = -1;
    const int d = s->dim;
    const int nx = s->gn[0], ny = s->gn[1], nz = d == 3 ? s->gn[2] : 1;
    const int nen = d == 3 ? 8 : 4;
    if (!((d == 3 && s->kind == SFH_HEX8) || (d == 2 && s->kind == SFH_QUAD4)) || nx < 2 || ny < 2 ||
        (d == 3 && nz < 2) || (long long)nx * ny * nz != s->nnodes ||
        (long long)(nx - 1) * (ny - 1) * (d == 3 ? nz - 1 : 1) != s->nel ||
        (long long)s->nnodes * 27 >= INT32_MAX || ny > 65535 || nz > 65535)
        return SFH_OK;
    SFH_TRY(s->gcell.ensure(sizeof(int) * (size_t)s->nel));
    SFH_TRY(s->asm_err.ensure(sizeof(int)));
    SFH_CHECK_HIP(hipMemsetAsync(s->gcell.p, 0xff, sizeof(int) * (size_t)s->nel, st));   // -1
    SFH_CHECK_HIP(hipMemsetAsync(s->asm_err.p, 0, sizeof(int), st));
    k_grid_cells<<<(unsigned)((s->nel + 255) / 256), 256, 0, st>>>(s->nel, nen, s->nnodes, nx, ny, nz,
                                                                   s->conn.as<int>(), s->gcell.as<int>(),
                                                                   s->asm_err.as<int>());
    SFH_LAUNCH_CHECK();
    k_grid_pattern<<<(unsigned)((s->nnodes + 255) / 256), 256, 0, st>>>(nx, ny, nz, s->rowptr_node.as<int>(),
                                                                        s->colind_node.as<int>(),
                                                                        s->asm_err.as<int>());
    SFH_LAUNCH_CHECK();
    int flags = 0;
    SFH_CHECK_HIP(hipMemcpyAsync(&flags, s->asm_err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipMemsetAsync(s->asm_err.p, 0, sizeof(int), st));   // the assembly's error word again
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    if (flags & 1) return SFH_OK;
    s->gcell_identity = !(flags & 2);
    s->gcell_state = 1;
    return SFH_OK;
}

// contribution lists of the two-phase hex8 assemblies, built once per system
int elastic_contrib_setup(sfh_system* s, hipStream_t st) {
    if (s->el_ready) return SFH_OK;
    const long long nzn = s->nnz_node;
    const int nn = s->nnodes;
    DevBuf cnt;
    SFH_TRY(cnt.ensure(sizeof(int) * (size_t)nzn));
    SFH_TRY(s->el_cptr.ensure(sizeof(int) * (size_t)(nzn + 1)));
    SFH_TRY(s->el_cent.ensure(sizeof(int) * (size_t)s->nel * 64));   // 8 nodes x 8 neighbours per element
    SFH_CHECK_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int) * (size_t)nzn, st));
    const unsigned g = (unsigned)((nn + 255) / 256);
    k_elastic_contrib<<<g, 256, 0, st>>>(nn, 0, s->n2e_ptr.as<int>(), s->n2e.as<int>(), s->emap.as<unsigned char>(),
                                         s->rowptr_node.as<int>(), nullptr, cnt.as<int>(), nullptr);
    SFH_LAUNCH_CHECK();
    SFH_TRY(scan_exclusive((int)nzn, cnt.as<int>(), s->el_cptr.as<int>(), st));
    SFH_CHECK_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int) * (size_t)nzn, st));
    k_elastic_contrib<<<g, 256, 0, st>>>(nn, 1, s->n2e_ptr.as<int>(), s->n2e.as<int>(), s->emap.as<unsigned char>(),
                                         s->rowptr_node.as<int>(), s->el_cptr.as<int>(), cnt.as<int>(),
                                         s->el_cent.as<int>());
    SFH_LAUNCH_CHECK();
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    s->el_ready = true;
    return SFH_OK;
}

int assemble_matrix(sfh_system* s, int physics, int nq1d, const double* p1, const double* p2,
                    double s1, double s2, double factor) {
    hipStream_t st = s->ctx->stream;
    if (nq1d < 1 || nq1d > 3) { set_error("nq1d must be 1..3"); return SFH_EARG; }
    if (s->kind == SFH_TRI3 && nq1d != 1) { set_error("TRI3 uses the 1-point rule"); return SFH_EARG; }
    SFH_TRY(s->asm_err.ensure(sizeof(int)));
    SFH_TRY(s->asm_host.ensure(sizeof(int)));
    int* errp = s->asm_err.as<int>();
    SFH_CHECK_HIP(hipMemsetAsync(errp, 0, sizeof(int), st));
    const int* c = s->conn.as<int>();
    const double* X = s->coords.as<double>();
    // hex8 cells of a declared 3D grid, integrated exactly by the closed form
    // when the rule has two or more points per axis
    bool grid_cells = false;
    if (nq1d >= 2 && ((s->kind == SFH_HEX8 && s->dim == 3) || (s->kind == SFH_QUAD4 && s->dim == 2))) {
        SFH_TRY(grid_cells_setup(s, st));
        grid_cells = s->gcell_state == 1;
    }
    const int* gcell = s->gcell_identity ? nullptr : s->gcell.as<int>();
    sfh_ctx* cx = s->ctx;
    const bool timed = cx->ksample > 0;
    if (timed) {
        if (!cx->aev0) SFH_CHECK_HIP(hipEventCreate(&cx->aev0));
        if (!cx->aev1) SFH_CHECK_HIP(hipEventCreate(&cx->aev1));
        SFH_CHECK_HIP(hipEventRecord(cx->aev0, st));
    }
    if (physics == SFH_ELASTICITY) {
        if (s->kind != SFH_HEX8 || s->dofpn != 3) {
            set_error("elasticity needs HEX8 elements with 3 DOFs per node");
            return SFH_EARG;
        }
        if (grid_cells) {
            constexpr int nw = hex_grid_nodes<SFH_ELASTICITY>();
            const dim3 gb((unsigned)((s->gn[0] + nw - 1) / nw), (unsigned)s->gn[1], (unsigned)s->gn[2]);
            k_hex_grid<SFH_ELASTICITY><<<gb, 256, 0, st>>>(
                s->gn[0], s->gn[1], s->gn[2], s->gaxis[0].as<double>(), s->gaxis[1].as<double>(),
                s->gaxis[2].as<double>(), gcell, p1, p2, s1, s2, factor, s->rowptr.as<int>(),
                s->vals.as<double>(), s->vals32.as<float>(), errp);
            SFH_LAUNCH_CHECK();
        } else if (s->nel < (1 << 25) && 64LL * s->nel < INT32_MAX && s->nnz_node < INT32_MAX) {
            SFH_TRY(elastic_contrib_setup(s, st));
            SFH_TRY(s->el_ke.ensure(sizeof(double) * kKeStride * (size_t)s->nel));
            k_elastic_ke<<<(unsigned)((s->nel + kElPerBlk - 1) / kElPerBlk), 256, 0, st>>>(
                s->nel, nq1d, X, c, p1, p2, s1, s2, factor, s->el_ke.as<double>(), errp);
            SFH_LAUNCH_CHECK();
            k_elastic_gather<<<(unsigned)(((s->nnodes + 7) / 8 + 7) / 8 * 8), 256, 0, st>>>(
                s->nnodes, s->rowptr_node.as<int>(), s->el_cptr.as<int>(), s->el_cent.as<int>(),
                s->el_ke.as<double>(), s->rowptr.as<int>(), s->vals.as<double>(), s->vals32.as<float>());
            SFH_LAUNCH_CHECK();
        } else {   // meshes beyond the contribution lists' 32-bit codes: one thread per DOF row
            const size_t lds = sizeof(double) * kBlockEl * s->maxrow;
            if (lds > s->ctx->lds_per_block) { set_error("row too long for LDS (%d)", s->maxrow); return SFH_EARG; }
            k_assemble_elastic<<<grid_for(s->ndof, kBlockEl), kBlockEl, lds, st>>>(
                s->ndof, nq1d, X, c, s->n2e_ptr.as<int>(), s->n2e.as<int>(), s->emap.as<unsigned char>(),
                s->rowptr.as<int>(), p1, p2, s1, s2, factor, s->vals.as<double>(), s->vals32.as<float>(),
                errp);
            SFH_LAUNCH_CHECK();
        }
    } else if ((physics == SFH_POISSON || physics == SFH_MATERN) && s->kind == SFH_HEX8 && s->dofpn == 1 &&
               grid_cells) {
        constexpr int nw = hex_grid_nodes<SFH_POISSON>();   // (the same for SFH_MATERN)
        const dim3 gb((unsigned)((s->gn[0] + nw - 1) / nw), (unsigned)s->gn[1], (unsigned)s->gn[2]);
        if (physics == SFH_POISSON)
            k_hex_grid<SFH_POISSON><<<gb, 256, 0, st>>>(s->gn[0], s->gn[1], s->gn[2], s->gaxis[0].as<double>(),
                                                         s->gaxis[1].as<double>(), s->gaxis[2].as<double>(),
                                                         gcell, p1, p2, s1, s2, factor,
                                                         s->rowptr.as<int>(), s->vals.as<double>(),
                                                         s->vals32.as<float>(), errp);
        else
            k_hex_grid<SFH_MATERN><<<gb, 256, 0, st>>>(s->gn[0], s->gn[1], s->gn[2], s->gaxis[0].as<double>(),
                                                        s->gaxis[1].as<double>(), s->gaxis[2].as<double>(),
                                                        gcell, p1, p2, s1, s2, factor,
                                                        s->rowptr.as<int>(), s->vals.as<double>(),
                                                        s->vals32.as<float>(), errp);
        SFH_LAUNCH_CHECK();
    } else if ((physics == SFH_POISSON || physics == SFH_MATERN) && s->kind == SFH_HEX8 && s->dofpn == 1 &&
               s->nel < (1 << 25) && 64LL * s->nel < INT32_MAX) {
        SFH_TRY(elastic_contrib_setup(s, st));
        SFH_TRY(s->el_ke.ensure(sizeof(double) * 36 * (size_t)s->nel));
        const unsigned ge = (unsigned)((s->nel + kElPerBlk - 1) / kElPerBlk);
        if (physics == SFH_POISSON)
            k_hex_scalar_ke<SFH_POISSON><<<ge, 256, 0, st>>>(s->nel, nq1d, X, c, p1, p2, s1, s2, factor,
                                                              s->el_ke.as<double>(), errp);
        else
            k_hex_scalar_ke<SFH_MATERN><<<ge, 256, 0, st>>>(s->nel, nq1d, X, c, p1, p2, s1, s2, factor,
                                                             s->el_ke.as<double>(), errp);
        SFH_LAUNCH_CHECK();
        k_hex_scalar_gather<<<(unsigned)(((s->nnz + 255) / 256 + 7) / 8 * 8), 256, 0, st>>>(
            (int)s->nnz, s->el_cptr.as<int>(), s->el_cent.as<int>(), s->el_ke.as<double>(), s->vals.as<double>(),
            s->vals32.as<float>());
        SFH_LAUNCH_CHECK();
    } else if ((physics == SFH_POISSON || physics == SFH_MATERN) && s->kind == SFH_QUAD4 && s->dofpn == 1 &&
               grid_cells) {
        const unsigned gb = (unsigned)(((long long)s->gn[0] * s->gn[1] + kQgB - 1) / kQgB);
        if (physics == SFH_POISSON)
            k_quad_grid<SFH_POISSON><<<gb, kQgB, 0, st>>>(s->gn[0], s->gn[1], s->gaxis[0].as<double>(),
                                                          s->gaxis[1].as<double>(), gcell, p1, p2, s1, s2, factor,
                                                          s->rowptr.as<int>(), s->vals.as<double>(),
                                                          s->vals32.as<float>(), errp);
        else
            k_quad_grid<SFH_MATERN><<<gb, kQgB, 0, st>>>(s->gn[0], s->gn[1], s->gaxis[0].as<double>(),
                                                         s->gaxis[1].as<double>(), gcell, p1, p2, s1, s2, factor,
                                                         s->rowptr.as<int>(), s->vals.as<double>(),
                                                         s->vals32.as<float>(), errp);
        SFH_LAUNCH_CHECK();
    } else if (physics == SFH_POISSON || physics == SFH_MATERN) {
        if (s->dofpn != 1) { set_error("scalar integrand needs 1 DOF per node"); return SFH_EARG; }
        const size_t lds = sizeof(double) * kBlock * s->maxrow;
        if (lds > s->ctx->lds_per_block) { set_error("row too long for LDS (%d)", s->maxrow); return SFH_EARG; }
        const int g = grid_for(s->nnodes, kBlock);
#define SFH_ASM(K, P)                                                                           \
    k_assemble_scalar<K, P><<<g, kBlock, lds, st>>>(                                            \
        s->nnodes, nq1d, X, c, s->n2e_ptr.as<int>(), s->n2e.as<int>(), s->emap.as<unsigned char>(), \
        s->rowptr.as<int>(), p1, p2, s1, s2, factor, s->maxrow, s->vals.as<double>(),          \
        s->vals32.as<float>(), errp)
        if (physics == SFH_POISSON) {
            switch (s->kind) {
                case SFH_LINE2: SFH_ASM(SFH_LINE2, SFH_POISSON); break;
                case SFH_TRI3: SFH_ASM(SFH_TRI3, SFH_POISSON); break;
                case SFH_QUAD4: SFH_ASM(SFH_QUAD4, SFH_POISSON); break;
                case SFH_HEX8: SFH_ASM(SFH_HEX8, SFH_POISSON); break;
            }
        } else {
            switch (s->kind) {
                case SFH_LINE2: SFH_ASM(SFH_LINE2, SFH_MATERN); break;
                case SFH_TRI3: SFH_ASM(SFH_TRI3, SFH_MATERN); break;
                case SFH_QUAD4: SFH_ASM(SFH_QUAD4, SFH_MATERN); break;
                case SFH_HEX8: SFH_ASM(SFH_HEX8, SFH_MATERN); break;
            }
        }
#undef SFH_ASM
        SFH_LAUNCH_CHECK();
    } else {
        set_error("unknown physics %d", physics);
        return SFH_EARG;
    }
    if (timed) SFH_CHECK_HIP(hipEventRecord(cx->aev1, st));
    SFH_CHECK_HIP(hipMemcpyAsync(s->asm_host.p, errp, sizeof(int), hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    if (timed) {
        float ams = 0.f;
        SFH_CHECK_HIP(hipEventElapsedTime(&ams, cx->aev0, cx->aev1));
        cx->asum += ams;
        ++cx->acount;
    }
    const int herr = *s->asm_host.as<int>();
    if (herr & 4) { set_error("Negative Jacobian."); return SFH_ENEGJAC; }
    SFH_CHECK_HIP(hipMemsetAsync(s->cmask.p, 0, s->ndof, st));
    s->nc = 0;
    s->assembled = true;
    s->dirty = true;
    if (s->dia_state > 0) s->dia_state = 0;
    return SFH_OK;
}

int assemble_rhs(sfh_system* s, int nq1d, const double* fconst, const double* fnodal, double factor,
                 int accumulate, double* out) {
    hipStream_t st = s->ctx->stream;
    if (nq1d < 1 || nq1d > 3) { set_error("nq1d must be 1..3"); return SFH_EARG; }
    SFH_TRY(s->asm_err.ensure(sizeof(int)));
    SFH_TRY(s->asm_host.ensure(sizeof(int)));
    int* errp = s->asm_err.as<int>();
    SFH_CHECK_HIP(hipMemsetAsync(errp, 0, sizeof(int), st));
    const double f0 = fconst ? fconst[0] : 0.0;
    const double f1 = (fconst && s->dofpn > 1) ? fconst[1] : 0.0;
    const double f2 = (fconst && s->dofpn > 2) ? fconst[2] : 0.0;
    const int g = grid_for(s->nnodes, kBlock);
#define SFH_BF(K)                                                                                 \
    k_body_force<K><<<g, kBlock, 0, st>>>(s->nnodes, nq1d, s->dofpn, s->coords.as<double>(),      \
                                          s->conn.as<int>(), s->n2e_ptr.as<int>(), s->n2e.as<int>(), \
                                          f0, f1, f2, fnodal, factor, accumulate, out, errp)
    switch (s->kind) {
        case SFH_LINE2: SFH_BF(SFH_LINE2); break;
        case SFH_TRI3: SFH_BF(SFH_TRI3); break;
        case SFH_QUAD4: SFH_BF(SFH_QUAD4); break;
        case SFH_HEX8: SFH_BF(SFH_HEX8); break;
    }
#undef SFH_BF
    SFH_LAUNCH_CHECK();
    SFH_CHECK_HIP(hipMemcpyAsync(s->asm_host.p, errp, sizeof(int), hipMemcpyDeviceToHost, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    const int herr = *s->asm_host.as<int>();
    if (herr & 4) { set_error("Negative Jacobian."); return SFH_ENEGJAC; }
    return SFH_OK;
}

int apply_constraints(sfh_system* s, int nc, const int* dofs, const double* values, double factor,
                      double scalar) {
    hipStream_t st = s->ctx->stream;
    if (nc <= 0) return SFH_OK;
    for (int i = 0; i < nc; ++i)
        if (dofs[i] < 0 || dofs[i] >= s->ndof) { set_error("constraint dof %d out of range", dofs[i]); return SFH_EARG; }
    DevBuf d, v;
    SFH_TRY(d.ensure(sizeof(int) * nc));
    SFH_CHECK_HIP(hipMemcpyAsync(d.p, dofs, sizeof(int) * nc, hipMemcpyHostToDevice, st));
    if (values) {
        SFH_TRY(v.ensure(sizeof(double) * nc));
        SFH_CHECK_HIP(hipMemcpyAsync(v.p, values, sizeof(double) * nc, hipMemcpyHostToDevice, st));
    }
    k_constraints<<<grid_for(nc, kBlock), kBlock, 0, st>>>(nc, d.as<int>(), values ? v.as<double>() : nullptr,
                                                           factor, scalar, s->rowptr.as<int>(),
                                                           s->colind.as<int>(), s->vals.as<double>(),
                                                           s->vals32.as<float>(), s->F.as<double>(),
                                                           s->cmask.as<unsigned char>());
    SFH_LAUNCH_CHECK();
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    s->dirty = true;
    if (s->dia_state > 0) s->dia_state = 0;
    return SFH_OK;
}

// the same constraints from a DOF list already on the device (values 0), with
// no allocation and no wait: for loops that re-assemble and re-constrain
int apply_constraints_dev(sfh_system* s, int nc, const int* dofs_dev, double factor, double scalar) {
    if (nc <= 0) return SFH_OK;
    k_constraints<<<grid_for(nc, kBlock), kBlock, 0, s->ctx->stream>>>(
        nc, dofs_dev, nullptr, factor, scalar, s->rowptr.as<int>(), s->colind.as<int>(), s->vals.as<double>(),
        s->vals32.as<float>(), s->F.as<double>(), s->cmask.as<unsigned char>());
    SFH_LAUNCH_CHECK();
    s->dirty = true;
    if (s->dia_state > 0) s->dia_state = 0;
    return SFH_OK;
}

int spmv(sfh_system* s, int k, const double* X, double* Y) {
    hipStream_t st = s->ctx->stream;
    k_spmv64<<<grid_for((long long)s->ndof * k, 256), 256, 0, st>>>(s->ndof, k, s->rowptr.as<int>(), s->colind.as<int>(),
                                                                    s->vals.as<double>(), X, Y);
    SFH_LAUNCH_CHECK();
    return SFH_OK;
}

// V -= c M^-1 T (m: the lumped mass) or V -= c T (m == nullptr), N x k
__global__ void k_frac_step(long long total, int k, double c, const double* __restrict__ m,
                            const double* __restrict__ T, double* __restrict__ V) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    V[t] -= m ? c * T[t] / m[t / k] : c * T[t];
}

// Y = lead prod_j (I - r_j M^-1 A) X (transpose = 0: P_l or P_r of
// statX/inferQ/PrecisionMatrixRational.ipp:41-161, the factors applied right to
// left as that matrix product applies them to X) or the transpose
// lead prod_j (I - r_j A M^-1) X (transpose = 1, the factors in the reverse
// order); roots = (lead, r_0 .. r_{degree-1}) on the host.
int frac_apply(sfh_system* s, const double* mass, int degree, const double* roots, int transpose, int k,
               const double* X, double* Y) {
    hipStream_t st = s->ctx->stream;
    const long long total = (long long)s->ndof * k;
    DevBuf v, t, u;
    SFH_TRY(v.ensure(sizeof(double) * total));
    SFH_TRY(t.ensure(sizeof(double) * total));
    if (transpose) SFH_TRY(u.ensure(sizeof(double) * total));
    double* V = v.as<double>();
    SFH_CHECK_HIP(hipMemcpyAsync(V, X, sizeof(double) * total, hipMemcpyDeviceToDevice, st));
    for (int q = 0; q < degree; ++q) {
        const double r = roots[1 + (transpose ? q : degree - 1 - q)];
        if (!transpose) {
            SFH_TRY(spmv(s, k, V, t.as<double>()));
            k_frac_step<<<grid_for(total, 256), 256, 0, st>>>(total, k, r, mass, t.as<double>(), V);
        } else {
            k_rowdiv<<<grid_for(total, 256), 256, 0, st>>>(total, k, 1.0, mass, V, t.as<double>());
            SFH_LAUNCH_CHECK();
            SFH_TRY(spmv(s, k, t.as<double>(), u.as<double>()));
            k_frac_step<<<grid_for(total, 256), 256, 0, st>>>(total, k, r, nullptr, u.as<double>(), V);
        }
        SFH_LAUNCH_CHECK();
    }
    k_rowdiv<<<grid_for(total, 256), 256, 0, st>>>(total, 1, roots[0], nullptr, V, Y);
    SFH_LAUNCH_CHECK();
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    return SFH_OK;
}

// Q X = tau^2 kappa^{4 beta} (A M^-1)^{2 alpha - 1} A X  (integer beta,
// statX/inferQ/PrecisionMatrixRational.ipp:177-209,278-298)
int precision_apply(sfh_system* s, const double* mass, double tau2k4b, int alpha, int k, const double* X, double* Y) {
    hipStream_t st = s->ctx->stream;
    const long long total = (long long)s->ndof * k;
    DevBuf t;
    SFH_TRY(t.ensure(sizeof(double) * total));
    double* a = Y;
    double* b = t.as<double>();
    SFH_TRY(spmv(s, k, X, a));
    for (int rep = 0; rep < 2 * alpha - 1; ++rep) {
        k_rowdiv<<<grid_for(total, 256), 256, 0, st>>>(total, k, 1.0, mass, a, b);
        SFH_LAUNCH_CHECK();
        SFH_TRY(spmv(s, k, b, a));
    }
    k_rowdiv<<<grid_for(total, 256), 256, 0, st>>>(total, 1, tau2k4b, nullptr, a, a);
    SFH_LAUNCH_CHECK();
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    return SFH_OK;
}

}  // namespace sfh
