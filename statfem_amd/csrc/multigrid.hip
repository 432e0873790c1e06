// Geometric multigrid hierarchy for the preconditioner of a single 2D tensor
// grid. No reference counterpart: the reference factorises the assembled
// operator directly (corlib/SystemSolveEigenSparse.hpp:273-311); on a declared
// grid this library solves iteratively instead, and this hierarchy is what
// makes the iteration count independent of the grid size.
//
// Levels: level 0 is the W x H node grid of the operator (its 9-coefficient
// stencil); a batch of grids gets one hierarchy per grid, levels side by side. Level l+1 keeps every other node of level l in each direction,
// ((W+1)/2) x ((H+1)/2) nodes. Prolongation P is bilinear: a fine node at an
// even index copies its coarse parent, one at an odd index averages the two
// neighbouring parents, and an odd last node (even W) copies its single parent
// (P reproduces constants up to the grid edge). On level 0, P is restricted to
// the free (unconstrained) nodes, so the coarse spaces see the operator's free
// block. Coarse operators are Galerkin products A_{l+1} = P^T A_l P, formed in
// FP64; for a 9-point stencil and bilinear P they are 9-point again.
//
// Smoother: l1-Jacobi (Baker, Falgout, Kolev, Yang, SIAM J. Sci. Comput. 33,
// 2011) with D_l1 = sum_j |a_ij| and weight omega = 4/3. A <= D_l1 for any
// symmetric A, so the sweep contracts on every level whatever the
// coefficients; on a constant-coefficient interior row it is 2/3-damped
// Jacobi. The coarsest level (at most kMGCoarse nodes) is solved with a dense
// inverse computed in FP64 on the host.
#include <cmath>

#include "common.hpp"

namespace sfh {
namespace {

constexpr double kOmega = 4.0 / 3.0;

// weight of coarse index X in fine index x (grid of n fine nodes)
__host__ __device__ inline double pweight(int x, int X, int n) {
    if ((x & 1) == 0) return x == 2 * X ? 1.0 : 0.0;
    if (x == n - 1) return (x - 1) / 2 == X ? 1.0 : 0.0;
    return ((x - 1) / 2 == X || (x + 1) / 2 == X) ? 0.5 : 0.0;
}

// coarse parents of fine index x: count, indices and weights
__device__ inline int parents(int x, int n, int* P, double* w) {
    if ((x & 1) == 0) { P[0] = x / 2; w[0] = 1.0; return 1; }
    if (x == n - 1) { P[0] = (x - 1) / 2; w[0] = 1.0; return 1; }
    P[0] = (x - 1) / 2; w[0] = 0.5;
    P[1] = (x + 1) / 2; w[1] = 0.5;
    return 2;
}

// A_c = P^T A P on coarse node I: sum over the fine nodes i of I's support and
// their stencil neighbours j of P[i][I] a_ij P[j][J]. cmask (level 0 only)
// drops constrained fine nodes from P. A coarse node without free support
// gets a decoupled unit row.
__global__ void k_mg_galerkin(int W, int H, int Wc, int Hc, const double* __restrict__ A,
                              const unsigned char* __restrict__ cmask, double* __restrict__ Ac) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= Wc * Hc) return;
    const int Ix = I % Wc, Iy = I / Wc;
    double a9[9];
    for (int k = 0; k < 9; ++k) a9[k] = 0.0;
    for (int dy = -1; dy <= 1; ++dy) {
        const int iy = 2 * Iy + dy;
        if (iy < 0 || iy >= H) continue;
        const double wy = pweight(iy, Iy, H);
        if (wy == 0.0) continue;
        for (int dx = -1; dx <= 1; ++dx) {
            const int ix = 2 * Ix + dx;
            if (ix < 0 || ix >= W) continue;
            const double wi = wy * pweight(ix, Ix, W);
            const int i = iy * W + ix;
            if (wi == 0.0 || (cmask && cmask[i])) continue;
            for (int k = 0; k < 9; ++k) {
                const int jx = ix + k % 3 - 1, jy = iy + k / 3 - 1;
                if (jx < 0 || jx >= W || jy < 0 || jy >= H) continue;
                const double a = A[(size_t)i * 9 + k];
                if (a == 0.0 || (cmask && cmask[jy * W + jx])) continue;
                int PX[2], PY[2];
                double WX[2], WY[2];
                const int nx = parents(jx, W, PX, WX), ny = parents(jy, H, PY, WY);
                for (int u = 0; u < ny; ++u)
                    for (int v = 0; v < nx; ++v)
                        a9[(PY[u] - Iy + 1) * 3 + (PX[v] - Ix + 1)] += wi * a * WY[u] * WX[v];
            }
        }
    }
    if (!(a9[4] > 0.0)) {
        for (int k = 0; k < 9; ++k) a9[k] = 0.0;
        a9[4] = 1.0;
    }
    for (int k = 0; k < 9; ++k) Ac[(size_t)I * 9 + k] = a9[k];
}

// smoother weights omega / sum_j |a_ij| and (levels >= 1) the FP32 stencil
__global__ void k_mg_level(int n, const double* __restrict__ A, float* __restrict__ dsm, float* __restrict__ A32) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int k = 0; k < 9; ++k) {
        const double a = A[(size_t)i * 9 + k];
        s += fabs(a);
        if (A32) A32[(size_t)i * 9 + k] = (float)a;
    }
    dsm[i] = s > 0.0 ? (float)(kOmega / s) : 0.f;
}

// dense inverse of the coarsest stencil operator (Cholesky, FP64); false when
// the operator is not positive definite
bool coarse_inverse(int W, int H, const std::vector<double>& a9, std::vector<float>& inv) {
    const int n = W * H;
    std::vector<double> A((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        const int x = i % W, y = i / W;
        for (int k = 0; k < 9; ++k) {
            const int jx = x + k % 3 - 1, jy = y + k / 3 - 1;
            if (jx < 0 || jx >= W || jy < 0 || jy >= H) continue;
            A[(size_t)i * n + jy * W + jx] = a9[(size_t)i * 9 + k];
        }
    }
    // symmetrise the FP64 rounding of the Galerkin sums
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) A[(size_t)i * n + j] = A[(size_t)j * n + i] = 0.5 * (A[(size_t)i * n + j] + A[(size_t)j * n + i]);
    for (int j = 0; j < n; ++j) {   // lower Cholesky factor in place
        double d = A[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
        if (!(d > 0.0)) return false;
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
        for (int i = 0; i < n; ++i) {   // L y = e_c
            double s = i == c ? 1.0 : 0.0;
            for (int k = 0; k < i; ++k) s -= A[(size_t)i * n + k] * col[k];
            col[i] = s / A[(size_t)i * n + i];
        }
        for (int i = n - 1; i >= 0; --i) {   // L^T x = y
            double s = col[i];
            for (int k = i + 1; k < n; ++k) s -= A[(size_t)k * n + i] * col[k];
            col[i] = s / A[(size_t)i * n + i];
        }
        for (int i = 0; i < n; ++i) inv[(size_t)i * n + c] = (float)col[i];
    }
    return true;
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
            w = (w + 1) / 2;
            h = (h + 1) / 2;
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
            for (int u = 0; u < l; ++u) { w = (w + 1) / 2; h = (h + 1) / 2; }
            const size_t o = (size_t)l * ng + q;
            m.gW[o] = w;
            m.gH[o] = h;
            m.gbase[o] = l == 0 ? grids[3 * q] : (int)base;
            base += (long long)w * h;
        }
        m.N[l] = l == 0 ? s->ndof : base;
    }
    SFH_TRY(m.dsm[0].ensure(sizeof(float) * (size_t)s->ndof));
    k_mg_level<<<(s->ndof + 255) / 256, 256, 0, st>>>(s->ndof, s->dia64.as<double>(), m.dsm[0].as<float>(), nullptr);
    SFH_LAUNCH_CHECK();
    for (int l = 0; l < m.L; ++l) {
        const long long nc = m.N[l + 1];
        SFH_TRY(m.coef64[l + 1].ensure(sizeof(double) * 9 * (size_t)nc));
        SFH_TRY(m.coef32[l + 1].ensure(sizeof(float) * 9 * (size_t)nc));
        SFH_TRY(m.dsm[l + 1].ensure(sizeof(float) * (size_t)nc));
        const double* Af = l == 0 ? s->dia64.as<double>() : m.coef64[l].as<double>();
        for (int q = 0; q < ng; ++q) {
            if (m.segL[q] <= l) continue;
            const size_t o = (size_t)l * ng + q, oc = (size_t)(l + 1) * ng + q;
            const int ncq = m.gW[oc] * m.gH[oc];
            k_mg_galerkin<<<(ncq + 255) / 256, 256, 0, st>>>(
                m.gW[o], m.gH[o], m.gW[oc], m.gH[oc], Af + (size_t)m.gbase[o] * 9,
                l == 0 ? s->cmask.as<unsigned char>() + m.gbase[o] : nullptr,
                m.coef64[l + 1].as<double>() + (size_t)m.gbase[oc] * 9);
            SFH_LAUNCH_CHECK();
        }
        k_mg_level<<<(unsigned)((nc + 255) / 256), 256, 0, st>>>((int)nc, m.coef64[l + 1].as<double>(),
                                                                 m.dsm[l + 1].as<float>(),
                                                                 m.coef32[l + 1].as<float>());
        SFH_LAUNCH_CHECK();
    }
    // dense inverses of the coarsest levels
    std::vector<float> all;
    m.cinv_off.assign(ng, 0);
    m.cmax = 0;
    for (int q = 0; q < ng; ++q) {
        const size_t o = (size_t)m.segL[q] * ng + q;
        const int w = m.gW[o], h = m.gH[o], n = w * h;
        std::vector<double> a9((size_t)n * 9);
        SFH_CHECK_HIP(hipMemcpyAsync(a9.data(), m.coef64[m.segL[q]].as<double>() + (size_t)m.gbase[o] * 9,
                                     sizeof(double) * 9 * (size_t)n, hipMemcpyDeviceToHost, st));
        SFH_CHECK_HIP(hipStreamSynchronize(st));
        std::vector<float> inv;
        if (!coarse_inverse(w, h, a9, inv)) { m.L = 0; return SFH_OK; }   // not SPD: Jacobi only
        m.cinv_off[q] = (int)all.size();
        all.insert(all.end(), inv.begin(), inv.end());
        m.cmax = std::max(m.cmax, n);
    }
    SFH_TRY(m.cinv.ensure(sizeof(float) * all.size()));
    SFH_CHECK_HIP(hipMemcpyAsync(m.cinv.p, all.data(), sizeof(float) * all.size(), hipMemcpyHostToDevice, st));
    SFH_CHECK_HIP(hipStreamSynchronize(st));
    return SFH_OK;
}

}  // namespace sfh
