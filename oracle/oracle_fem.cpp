// CPU FP64 oracle — finite-element half (test infrastructure only).
//
// This file restates, on the CPU and in plain double precision, the reference
// statFEM algorithms that the HIP library implements on the GPU. It is a
// checker: only the test suite, the smoke check and the CPU-baseline leg of the
// benchmark load it. No product code path calls into it.
//
// Parity status of each entry point is listed in oracle/README.md.
//
// Reference sources restated here (paths relative to the statFEM tree):
//   corlib/GaussLegendre.hpp:58-80           1D Gauss-Legendre rules on (0,1), 15-digit constants
//   corlib/TensorQuadrature.hpp:100-140      tensor-product ordering (last axis outermost)
//   corlib/SimplexQuadrature.hpp:162-171     1-point triangle rule
//   corlib/Shapefun.ipp:39-55,181-197,460-476,926-975   LINE2/TRI3/QUAD4/HEX8 on [0,1]^d
//   corlib/ElementBasic.ipp:177-205          globalDerivatives (J^-T dphi/dxi, detJ > 0 check)
//   corlib/linalg.hpp:200-223                square-case contravariant basis
//   del2/fem/ElementPotential.ipp:56-74      Poisson stiffness integrand
//   del2/fem/ElementPotential.ipp:146-175    body-force integrand
//   del2/fem/ElementMatern.ipp:57-82         Matern integrand (H grad.grad + lumped kappa^2 phi)
//   solid/fem/ElementStatic.ipp:284-321      hex8 stiffness  sum_JL C_iJkL dphi_J dphi_L
//   solid/material/svenant/SVenant.ipp:32-37,107-140    Lame constants, C_iJkL at F
//   corlib/MatrixAssembler.hpp:198-232       per-element compute, *factor, insertToMatrix
//   corlib/SystemSolveEigenSparse.hpp:171-215   insertToMatrix / applyConstraints
//   corlib/Triplet.hpp:104-170,219-304       COO insert, sort+merge (auto-merge past 1e7), zeroRow, CSR
//   corlib/SystemSolveEigenSparse.hpp:273-404   solveSystem / solveSystemTranspose (direct)
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <utility>
#include <thread>
#include <vector>

#include "oracle.h"

namespace {

// ---------------------------------------------------------------- quadrature
struct QP {
    double w;
    double xi[3];
};

// corlib/GaussLegendre.hpp:58-80 (weights and points exactly as stored there)
void gauss_legendre(int n, double* w, double* x) {
    if (n == 1) {
        w[0] = 1.0; x[0] = 0.5;
    } else if (n == 2) {
        w[0] = 0.5; x[0] = 0.788675134594813;
        w[1] = 0.5; x[1] = 0.211324865405187;
    } else {
        w[0] = 0.277777777777777; x[0] = 0.887298334620741;
        w[1] = 0.444444444444444; x[1] = 0.5;
        w[2] = 0.277777777777777; x[2] = 0.112701665379259;
    }
}

int kind_dim(int kind) {
    switch (kind) {
        case OR_LINE2: return 1;
        case OR_TRI3: return 2;
        case OR_QUAD4: return 2;
        case OR_HEX8: return 3;
    }
    return 0;
}

int kind_nen(int kind) {
    switch (kind) {
        case OR_LINE2: return 2;
        case OR_TRI3: return 3;
        case OR_QUAD4: return 4;
        case OR_HEX8: return 8;
    }
    return 0;
}

// Tensor rule: corlib/TensorQuadrature.hpp:100-140 -- the lower-dimensional rule is
// the inner loop, the new (last) axis the outer one; weight = w_lower * w_outer.
std::vector<QP> quad_rule(int kind, int nq1d) {
    std::vector<QP> q;
    if (kind == OR_TRI3) {
        QP p;
        p.w = 0.5;
        p.xi[0] = 0.333333333333333; p.xi[1] = 0.333333333333333; p.xi[2] = 0.0;
        q.push_back(p);
        return q;
    }
    double w[3], x[3];
    gauss_legendre(nq1d, w, x);
    const int d = kind_dim(kind);
    if (d == 1) {
        for (int i = 0; i < nq1d; ++i) {
            QP p; p.w = w[i]; p.xi[0] = x[i]; p.xi[1] = p.xi[2] = 0.0; q.push_back(p);
        }
    } else if (d == 2) {
        for (int o = 0; o < nq1d; ++o)
            for (int i = 0; i < nq1d; ++i) {
                QP p; p.w = w[i]; p.w *= w[o];
                p.xi[0] = x[i]; p.xi[1] = x[o]; p.xi[2] = 0.0; q.push_back(p);
            }
    } else {
        for (int o = 0; o < nq1d; ++o)
            for (int m = 0; m < nq1d; ++m)
                for (int i = 0; i < nq1d; ++i) {
                    QP p; p.w = w[i]; p.w *= w[m]; p.w *= w[o];
                    p.xi[0] = x[i]; p.xi[1] = x[m]; p.xi[2] = x[o]; q.push_back(p);
                }
    }
    return q;
}

// ------------------------------------------------------------ shape functions
// phi[K], dphi[alpha*nen + K]
void shape(int kind, const double* xi, double* phi, double* dphi) {
    const double a = xi[0], b = xi[1], c = xi[2];
    switch (kind) {
        case OR_LINE2:   // corlib/Shapefun.ipp:39-55
            phi[0] = 1. - a; phi[1] = a;
            dphi[0] = -1.; dphi[1] = 1.;
            break;
        case OR_TRI3:    // corlib/Shapefun.ipp:181-197
            phi[0] = 1. - (a + b); phi[1] = a; phi[2] = b;
            dphi[0] = -1.; dphi[1] = 1.; dphi[2] = 0.;
            dphi[3] = -1.; dphi[4] = 0.; dphi[5] = 1.;
            break;
        case OR_QUAD4:   // corlib/Shapefun.ipp:460-476
            phi[0] = (1. - a) * (1. - b); phi[1] = a * (1. - b);
            phi[2] = a * b;               phi[3] = (1. - a) * b;
            dphi[0] = -(1. - b); dphi[1] = (1. - b); dphi[2] = b;  dphi[3] = -b;
            dphi[4] = -(1. - a); dphi[5] = -a;       dphi[6] = a;  dphi[7] = (1. - a);
            break;
        case OR_HEX8:    // corlib/Shapefun.ipp:926-975
            phi[0] = (1. - a) * (1. - b) * (1. - c); phi[1] = a * (1. - b) * (1. - c);
            phi[2] = a * b * (1. - c);               phi[3] = (1. - a) * b * (1. - c);
            phi[4] = (1. - a) * (1. - b) * c;        phi[5] = a * (1. - b) * c;
            phi[6] = a * b * c;                      phi[7] = (1. - a) * b * c;
            dphi[0] = -(1. - b) * (1. - c); dphi[1] = (1. - b) * (1. - c);
            dphi[2] = b * (1. - c);         dphi[3] = -b * (1. - c);
            dphi[4] = -(1. - b) * c;        dphi[5] = (1. - b) * c;
            dphi[6] = b * c;                dphi[7] = -b * c;
            dphi[8] = -(1. - a) * (1. - c); dphi[9] = -a * (1. - c);
            dphi[10] = a * (1. - c);        dphi[11] = (1. - a) * (1. - c);
            dphi[12] = -(1. - a) * c;       dphi[13] = -a * c;
            dphi[14] = a * c;               dphi[15] = (1. - a) * c;
            dphi[16] = -(1. - a) * (1. - b); dphi[17] = -a * (1. - b);
            dphi[18] = -a * b;               dphi[19] = -(1. - a) * b;
            dphi[20] = (1. - a) * (1. - b);  dphi[21] = a * (1. - b);
            dphi[22] = a * b;                dphi[23] = (1. - a) * b;
            break;
    }
}

// Square inverse with determinant, corlib/linalg.hpp (inverse of J^T, returns det).
double invert(int d, const double* m, double* inv) {
    if (d == 1) {
        inv[0] = 1.0 / m[0];
        return m[0];
    }
    if (d == 2) {
        const double det = m[0] * m[3] - m[1] * m[2];
        inv[0] = m[3] / det; inv[1] = -m[1] / det;
        inv[2] = -m[2] / det; inv[3] = m[0] / det;
        return det;
    }
    const double det = m[0] * (m[4] * m[8] - m[5] * m[7])
                     - m[1] * (m[3] * m[8] - m[5] * m[6])
                     + m[2] * (m[3] * m[7] - m[4] * m[6]);
    inv[0] = (m[4] * m[8] - m[5] * m[7]) / det;
    inv[1] = (m[2] * m[7] - m[1] * m[8]) / det;
    inv[2] = (m[1] * m[5] - m[2] * m[4]) / det;
    inv[3] = (m[5] * m[6] - m[3] * m[8]) / det;
    inv[4] = (m[0] * m[8] - m[2] * m[6]) / det;
    inv[5] = (m[2] * m[3] - m[0] * m[5]) / det;
    inv[6] = (m[3] * m[7] - m[4] * m[6]) / det;
    inv[7] = (m[1] * m[6] - m[0] * m[7]) / det;
    inv[8] = (m[0] * m[4] - m[1] * m[3]) / det;
    return det;
}

// corlib/ElementBasic.ipp:177-205: coBase = X dphi^T, contra = (coBase^T)^-1,
// dphiDx = contra * dphiDxi. X[K*dim + i]. Returns detJ (<= 0 flags an error).
double global_derivatives(int kind, const double* X, const double* xi, double* phi,
                          double* dphiDx) {
    const int d = kind_dim(kind), nen = kind_nen(kind);
    double dphi[24];
    shape(kind, xi, phi, dphi);
    double coT[9] = {0};   // coT[alpha*d + i] = (coBase^T)[alpha][i] = sum_K dphi[alpha][K] X[K][i]
    for (int al = 0; al < d; ++al)
        for (int i = 0; i < d; ++i) {
            double s = 0.0;
            for (int K = 0; K < nen; ++K) s += X[K * d + i] * dphi[al * nen + K];
            coT[al * d + i] = s;
        }
    double contra[9];
    const double detJ = invert(d, coT, contra);
    for (int i = 0; i < d; ++i)
        for (int K = 0; K < nen; ++K) {
            double s = 0.0;
            for (int al = 0; al < d; ++al) s += contra[i * d + al] * dphi[al * nen + K];
            dphiDx[i * nen + K] = s;
        }
    return detJ;
}

void gather_coords(int d, int nen, const double* coords, const int* conn, double* X) {
    for (int K = 0; K < nen; ++K)
        for (int i = 0; i < d; ++i) X[K * d + i] = coords[(size_t)conn[K] * d + i];
}

// ------------------------------------------------------------------ Triplet
struct Trip {
    std::pair<unsigned, unsigned> ij;
    double v;
};

struct TripLess {
    bool operator()(const Trip& a, const Trip& b) const { return a.ij < b.ij; }
};

const size_t kAllowedOverhead = 10000000;   // corlib/Triplet.hpp:16

struct Triplet {
    std::vector<Trip> t;
    size_t sortedSize = 0;
    bool sorted = false;
    unsigned n = 0;

    // corlib/Triplet.hpp:126-170: std::sort by (row,col), merge equal keys by
    // summation in sorted order, truncate.
    void finish() {
        std::sort(t.begin(), t.end(), TripLess());
        size_t store = 0, b = 0;
        const size_t e = t.size();
        while (b != e) {
            size_t en = b + 1;
            while (en < e && t[en].ij == t[b].ij) ++en;
            double val = 0.0;
            for (size_t q = b; q < en; ++q) val += t[q].v;
            Trip m; m.ij = t[b].ij; m.v = val;
            t[store++] = m;
            b = en;
        }
        t.resize(store);
        sortedSize = store;
        sorted = true;
    }

    // corlib/Triplet.hpp:104-124
    void insert(unsigned i, unsigned j, double v) {
        Trip tr; tr.ij = std::make_pair(i, j); tr.v = v;
        t.push_back(tr);
        sorted = false;
        if ((t.size() - sortedSize) > kAllowedOverhead) finish();
    }

    // corlib/Triplet.hpp:219-260
    void zeroRow(unsigned row, double factor) {
        if (!sorted) finish();
        auto lo = std::lower_bound(t.begin(), t.end(), row,
                                   [](const Trip& a, unsigned r) { return a.ij.first < r; });
        auto hi = lo;
        while (hi != t.end() && hi->ij.first == row) ++hi;
        auto diag = lo;
        while (diag != hi && diag->ij.first != diag->ij.second) ++diag;
        for (auto it = lo; it != hi; ++it) it->v = 0.0;
        Trip tmp; tmp.ij = std::make_pair(row, row); tmp.v = factor;
        if (diag != hi) {
            *diag = tmp;
        } else {
            t.push_back(tmp);
            sorted = false;
        }
    }
};

}  // namespace

extern "C" {

int or_quad_rule(int kind, int nq1d, double* w, double* xi) {
    if (kind_dim(kind) == 0 || nq1d < 1 || nq1d > 3) return OR_EARG;
    auto q = quad_rule(kind, nq1d);
    for (size_t p = 0; p < q.size(); ++p) {
        w[p] = q[p].w;
        for (int i = 0; i < 3; ++i) xi[p * 3 + i] = q[p].xi[i];
    }
    return (int)q.size();
}

// Element matrices for all elements, in element order (what MatrixAssembler
// hands to insertToMatrix, corlib/MatrixAssembler.hpp:198-232).
//  physics OR_POISSON:    p1[e] = conductivity          (ElementPotential.ipp:56-74)
//  physics OR_ELASTICITY: p1[e] = E, p2[e] = nu, hex8   (ElementStatic.ipp:284-321)
//  physics OR_MATERN:     p1[e] = kappa, p2[e] = scalar diffusivity H
//                                                       (ElementMatern.ipp:57-82)
// Ke is nel x ned x ned, ned = nen*dofpn; interleaved DOFs (node-major).
int or_element_matrices(int physics, int kind, int nq1d, int nnodes, const double* coords,
                        int nel, const int* conn, const double* p1, const double* p2,
                        double factor, double* Ke) {
    const int d = kind_dim(kind), nen = kind_nen(kind);
    if (d == 0) return OR_EARG;
    if (physics == OR_ELASTICITY && kind != OR_HEX8) return OR_EARG;
    const int dofpn = physics == OR_ELASTICITY ? 3 : 1;
    const int ned = nen * dofpn;
    auto q = quad_rule(kind, nq1d);
    (void)nnodes;
    std::vector<double> K(ned * ned);
    for (int e = 0; e < nel; ++e) {
        double X[24], phi[8], g[24];
        gather_coords(d, nen, coords, conn + (size_t)e * nen, X);
        std::fill(K.begin(), K.end(), 0.0);
        for (const QP& qp : q) {
            const double detJ = global_derivatives(kind, X, qp.xi, phi, g);
            if (!(detJ > 0.0)) return OR_ENEGJAC;
            if (physics == OR_POISSON) {
                const double s = p1[e] * detJ * qp.w;
                for (int a = 0; a < nen; ++a)
                    for (int b = 0; b < nen; ++b) {
                        double m = 0.0;
                        for (int i = 0; i < d; ++i) m += g[i * nen + a] * g[i * nen + b];
                        K[a * ned + b] += s * m;
                    }
            } else if (physics == OR_MATERN) {
                const double kap = p1[e], H = p2[e];
                for (int a = 0; a < nen; ++a)
                    for (int b = 0; b < nen; ++b) {
                        double m = 0.0;
                        for (int i = 0; i < d; ++i) m += (H * g[i * nen + a]) * g[i * nen + b];
                        K[a * ned + b] += detJ * qp.w * m;
                    }
                for (int a = 0; a < nen; ++a) K[a * ned + a] += detJ * qp.w * (kap * kap * phi[a]);
            } else {
                // SVenant::elasticityTensor at F = I (SVenant.ipp:107-140):
                // E = 0.5(F^T F - I) = 0, S = 0, FF^T = I.
                const double E = p1[e], nu = p2[e];
                const double lambda = E * nu / (1.0 + nu) / (1.0 - 2.0 * nu);
                const double mu = E / 2.0 / (1.0 + nu);
                double C[81];
                for (int i = 0; i < 3; ++i)
                    for (int J = 0; J < 3; ++J)
                        for (int k = 0; k < 3; ++k)
                            for (int L = 0; L < 3; ++L) {
                                const double Fi_J = (i == J), Fk_L = (k == L), Fi_L = (i == L),
                                             Fk_J = (k == J);
                                double c = lambda * Fi_J * Fk_L;
                                c += mu * Fi_L * Fk_J;
                                if (J == L) c += mu * (i == k ? 1.0 : 0.0);
                                if (i == k) c += 0.0;   // S(J,L) = 0 at F = I
                                C[((i * 3 + J) * 3 + k) * 3 + L] = c;
                            }
                for (int m = 0; m < nen; ++m)
                    for (int i = 0; i < 3; ++i)
                        for (int n = 0; n < nen; ++n)
                            for (int k = 0; k < 3; ++k) {
                                double ke = 0.0;
                                for (int J = 0; J < 3; ++J)
                                    for (int L = 0; L < 3; ++L)
                                        ke += C[((i * 3 + J) * 3 + k) * 3 + L] * g[J * nen + m] *
                                              g[L * nen + n];
                                K[(m * 3 + i) * ned + n * 3 + k] += ke * detJ * qp.w;
                            }
            }
        }
        double* out = Ke + (size_t)e * ned * ned;
        for (int r = 0; r < ned * ned; ++r) out[r] = K[r] * factor;
    }
    return OR_OK;
}

// Body force / unit lumped mass: F[dof] = factor * sum_e sum_q w detJ f(x_q) phi_a(xi_q)
// (del2/fem/ElementPotential.ipp:146-175 via corlib/Integrator.hpp:62-74 and
// corlib/VectorAssembler.hpp:101-114). f(x) = fconst[c] + sum_K fnodal[K*dofpn+c] phi_K.
int or_body_force(int kind, int nq1d, int nnodes, const double* coords, int nel,
                  const int* conn, int dofpn, const double* fconst, const double* fnodal,
                  double factor, double* F) {
    const int d = kind_dim(kind), nen = kind_nen(kind);
    if (d == 0) return OR_EARG;
    auto q = quad_rule(kind, nq1d);
    std::fill(F, F + (size_t)nnodes * dofpn, 0.0);
    for (int e = 0; e < nel; ++e) {
        const int* cn = conn + (size_t)e * nen;
        double X[24], phi[8], g[24];
        gather_coords(d, nen, coords, cn, X);
        for (const QP& qp : q) {
            const double detJ = global_derivatives(kind, X, qp.xi, phi, g);
            if (!(detJ > 0.0)) return OR_ENEGJAC;
            for (int c = 0; c < dofpn; ++c) {
                double f = fconst ? fconst[c] : 0.0;
                if (fnodal)
                    for (int K = 0; K < nen; ++K) f += fnodal[(size_t)cn[K] * dofpn + c] * phi[K];
                for (int a = 0; a < nen; ++a)
                    F[(size_t)cn[a] * dofpn + c] += qp.w * detJ * factor * (f * phi[a]);
            }
        }
    }
    return OR_OK;
}

// --------------------------------------------------------------- triplets
void* or_triplet_new(int n) {
    Triplet* t = new Triplet;
    t->n = (unsigned)n;
    return t;
}

void or_triplet_free(void* h) { delete static_cast<Triplet*>(h); }

// insertToMatrix for every element in element order
// (corlib/SystemSolveEigenSparse.hpp:171-182).
int or_triplet_insert_elements(void* h, int nel, int ned, const int* edofs, const double* Ke) {
    Triplet* t = static_cast<Triplet*>(h);
    for (int e = 0; e < nel; ++e) {
        const int* r = edofs + (size_t)e * ned;
        const double* k = Ke + (size_t)e * ned * ned;
        for (int i = 0; i < ned; ++i)
            for (int j = 0; j < ned; ++j) t->insert((unsigned)r[i], (unsigned)r[j], k[i * ned + j]);
    }
    return OR_OK;
}

// applyConstraints (corlib/SystemSolveEigenSparse.hpp:197-215): zeroRow + F[c] = value*scalar.
int or_triplet_apply_constraints(void* h, double* F, int nc, const int* dofs,
                                 const double* values, double factor, double scalar) {
    Triplet* t = static_cast<Triplet*>(h);
    for (int c = 0; c < nc; ++c) {
        t->zeroRow((unsigned)dofs[c], scalar);
        if (F) F[dofs[c]] = factor * (values ? values[c] : 0.0) * scalar;
    }
    return OR_OK;
}

long long or_triplet_nnz(void* h) {
    Triplet* t = static_cast<Triplet*>(h);
    if (!t->sorted) t->finish();
    return (long long)t->t.size();
}

// setCompressedRowArrays (corlib/Triplet.hpp:264-304)
int or_triplet_csr(void* h, int* rowptr, int* colind, double* vals) {
    Triplet* t = static_cast<Triplet*>(h);
    if (!t->sorted) t->finish();
    int prev = -1;
    int idx = 0, r = 0;
    for (size_t q = 0; q < t->t.size(); ++q) {
        const int row = (int)t->t[q].ij.first;
        if (colind) colind[q] = (int)t->t[q].ij.second;
        if (vals) vals[q] = t->t[q].v;
        if (row != prev) {
            if (rowptr) rowptr[r] = idx;
            ++r;
        }
        ++idx;
        prev = row;
    }
    if (rowptr) rowptr[r] = (int)t->t.size();
    return r == (int)t->n ? OR_OK : OR_EARG;
}

// ------------------------------------------------------ banded direct solver
// Stand-in for Eigen::SparseLU on the row-zeroed, non-symmetric operator
// (corlib/SystemSolveEigenSparse.hpp:273-404): LU without pivoting in band
// storage; transposed solves reuse the same factors.
struct BandLU {
    int n = 0, kl = 0, ku = 0, ld = 0;
    std::vector<double> a;   // row-major band: a[i*ld + (j - i + kl)]
};

void* or_lu_factor(int n, const int* rowptr, const int* colind, const double* vals, int* status) {
    BandLU* f = new BandLU;
    f->n = n;
    for (int i = 0; i < n; ++i)
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            const int j = colind[p];
            f->kl = std::max(f->kl, i - j);
            f->ku = std::max(f->ku, j - i);
        }
    f->ld = f->kl + f->ku + 1;
    f->a.assign((size_t)n * f->ld, 0.0);
    for (int i = 0; i < n; ++i)
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p)
            f->a[(size_t)i * f->ld + (colind[p] - i + f->kl)] += vals[p];
    const int kl = f->kl, ku = f->ku, ld = f->ld;
    double* a = f->a.data();
    for (int k = 0; k < n; ++k) {
        const double piv = a[(size_t)k * ld + kl];
        if (piv == 0.0 || !std::isfinite(piv)) {
            delete f;
            if (status) *status = OR_ESINGULAR;
            return nullptr;
        }
        const int iend = std::min(n - 1, k + kl);
        const int jend = std::min(n - 1, k + ku);
        const double* rk = a + (size_t)k * ld + kl - k;   // rk[j] = A[k][j]
        for (int i = k + 1; i <= iend; ++i) {
            double* ri = a + (size_t)i * ld + kl - i;       // ri[j] = A[i][j]
            if (ri[k] == 0.0) continue;
            const double l = ri[k] / piv;
            ri[k] = l;
            for (int j = k + 1; j <= jend; ++j) ri[j] -= l * rk[j];
        }
    }
    if (status) *status = OR_OK;
    return f;
}

void or_lu_free(void* h) { delete static_cast<BandLU*>(h); }

// X = A^-1 B (trans = 0) or A^-T B (trans = 1); B, X are n x k row-major.
int or_lu_solve(void* h, int trans, int k, const double* B, double* X) {
    BandLU* f = static_cast<BandLU*>(h);
    const int n = f->n, kl = f->kl, ku = f->ku, ld = f->ld;
    const double* a = f->a.data();
    std::vector<double> x((size_t)n);
    for (int c = 0; c < k; ++c) {
        for (int i = 0; i < n; ++i) x[i] = B[(size_t)i * k + c];
        if (!trans) {
            for (int i = 0; i < n; ++i) {       // L y = b (unit lower)
                const double* ri = a + (size_t)i * ld + kl - i;
                double s = x[i];
                for (int j = std::max(0, i - kl); j < i; ++j) s -= ri[j] * x[j];
                x[i] = s;
            }
            for (int i = n - 1; i >= 0; --i) {  // U x = y
                const double* ri = a + (size_t)i * ld + kl - i;
                double s = x[i];
                const int jend = std::min(n - 1, i + ku);
                for (int j = i + 1; j <= jend; ++j) s -= ri[j] * x[j];
                x[i] = s / ri[i];
            }
        } else {
            for (int i = 0; i < n; ++i) {       // U^T y = b
                const double* ri = a + (size_t)i * ld + kl - i;
                x[i] /= ri[i];
                const double xi = x[i];
                const int jend = std::min(n - 1, i + ku);
                for (int j = i + 1; j <= jend; ++j) x[j] -= ri[j] * xi;
            }
            for (int i = n - 1; i >= 0; --i) {  // L^T x = y
                const double* ri = a + (size_t)i * ld + kl - i;
                const double xi = x[i];
                for (int j = std::max(0, i - kl); j < i; ++j) x[j] -= ri[j] * xi;
            }
        }
        for (int i = 0; i < n; ++i) X[(size_t)i * k + c] = x[i];
    }
    return OR_OK;
}

// ------------------------------------- banded Cholesky (symmetric elimination)
// CPU-baseline direct solver: homogeneous-Dirichlet rows *and* columns removed
// (identity on constrained DOFs), then L L^T in band storage. For homogeneous
// constraints this gives the same free-DOF solution as the row-zero operator.
struct BandChol {
    int n = 0, b = 0, ld = 0;
    std::vector<double> l;   // l[i*ld + (j - i + b)], j in [i-b, i]
};

void* or_chol_factor(int n, const int* rowptr, const int* colind, const double* vals,
                     const unsigned char* constrained, int* status) {
    BandChol* f = new BandChol;
    f->n = n;
    for (int i = 0; i < n; ++i)
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) f->b = std::max(f->b, i - colind[p]);
    f->ld = f->b + 1;
    const int b = f->b, ld = f->ld;
    f->l.assign((size_t)n * ld, 0.0);
    double* L = f->l.data();
    for (int i = 0; i < n; ++i) {
        const bool ci = constrained && constrained[i];
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            const int j = colind[p];
            if (j > i) continue;
            const bool cj = constrained && constrained[j];
            double v = vals[p];
            if (ci || cj) v = (i == j) ? 1.0 : 0.0;
            L[(size_t)i * ld + (j - i + b)] += v;
        }
    }
    // right-looking band Cholesky, row-oriented inner products
    for (int i = 0; i < n; ++i) {
        double* ri = L + (size_t)i * ld + b - i;   // ri[j] = L[i][j]
        const int j0 = std::max(0, i - b);
        for (int j = j0; j <= i; ++j) {
            const double* rj = L + (size_t)j * ld + b - j;
            double s = ri[j];
            const int k0 = std::max(j0, std::max(0, j - b));
            for (int k = k0; k < j; ++k) s -= ri[k] * rj[k];
            if (j < i) {
                ri[j] = s / rj[j];
            } else {
                if (!(s > 0.0)) {
                    delete f;
                    if (status) *status = OR_ESINGULAR;
                    return nullptr;
                }
                ri[i] = std::sqrt(s);
            }
        }
    }
    if (status) *status = OR_OK;
    return f;
}

void or_chol_free(void* h) { delete static_cast<BandChol*>(h); }

// Solves L L^T X = B for columns c0 .. c1-1 of k right-hand sides, row-major n x k.
static void chol_solve_cols(const BandChol* f, int k, int c0, int c1, const double* B, double* X) {
    const int n = f->n, b = f->b, ld = f->ld;
    const double* L = f->l.data();
    std::vector<double> x((size_t)n);
    for (int c = c0; c < c1; ++c) {
        for (int i = 0; i < n; ++i) x[i] = B[(size_t)i * k + c];
        for (int i = 0; i < n; ++i) {
            const double* ri = L + (size_t)i * ld + b - i;
            double s = x[i];
            for (int j = std::max(0, i - b); j < i; ++j) s -= ri[j] * x[j];
            x[i] = s / ri[i];
        }
        for (int i = n - 1; i >= 0; --i) {
            const double* ri = L + (size_t)i * ld + b - i;
            x[i] /= ri[i];
            const double xi = x[i];
            for (int j = std::max(0, i - b); j < i; ++j) x[j] -= ri[j] * xi;
        }
        for (int i = 0; i < n; ++i) X[(size_t)i * k + c] = x[i];
    }
}

int or_chol_solve(void* h, int k, const double* B, double* X) {
    chol_solve_cols(static_cast<BandChol*>(h), k, 0, k, B, X);
    return OR_OK;
}

// The same solves with the columns split over nthreads std::threads (the
// right-hand sides are independent; each thread keeps its own work vector).
int or_chol_solve_mt(void* h, int k, const double* B, double* X, int nthreads) {
    const BandChol* f = static_cast<BandChol*>(h);
    nthreads = std::max(1, std::min(nthreads, k));
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t) {
        const int c0 = (int)((long long)k * t / nthreads), c1 = (int)((long long)k * (t + 1) / nthreads);
        pool.emplace_back(chol_solve_cols, f, k, c0, c1, B, X);
    }
    for (auto& th : pool) th.join();
    return OR_OK;
}

// FP64 sparse y = A x for k right-hand sides (row-major), used by residual checks.
int or_spmv(int n, const int* rowptr, const int* colind, const double* vals, int k,
            const double* X, double* Y) {
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < k; ++c) {
            double s = 0.0;
            for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) s += vals[p] * X[(size_t)colind[p] * k + c];
            Y[(size_t)i * k + c] = s;
        }
    return OR_OK;
}

}  // extern "C"
