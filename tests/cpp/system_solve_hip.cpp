// Drives include/statfem/SystemSolveHip.hpp the way a statFEM driver drives
// SystemSolveEigenSparse (del2/apps/reference/laplaceRHSinfer/fem.hpp:69-96):
// per-element insertToMatrix / insertToRhs of a Q1 Poisson problem, boundary
// constraints, solveSystem, giveSolution, solveSystemTranspose and the batched
// transpose. Checks against the library's whole-mesh path through the C ABI
// (same CSR pattern, same solution) and prints OK.
#include <cmath>
#include <cstdio>
#include <vector>

#include "statfem/SystemSolveHip.hpp"

namespace {

struct Mat {   // minimal dense matrix with the interface the concept uses
    int r, c;
    std::vector<double> a;
    Mat(int r_, int c_) : r(r_), c(c_), a((size_t)r_ * c_, 0.0) {}
    double& operator()(size_t i, size_t j) { return a[i * c + j]; }
    double operator()(size_t i, size_t j) const { return a[i * c + j]; }
    int rows() const { return r; }
    int cols() const { return c; }
};

int fail(const char* what, double v) {
    std::printf("FAIL %s %.3e\n", what, v);
    return 1;
}

}  // namespace

int main() {
    const int n = 24, N = (n + 1) * (n + 1);
    const double h = 1.0 / n;
    // exact Q1 Laplacian element stiffness on a square cell (h-independent in 2D),
    // local nodes counter-clockwise from the lower-left corner
    Mat Ke(4, 4);
    const double k0[4][4] = {{4, -1, -2, -1}, {-1, 4, -1, -2}, {-2, -1, 4, -1}, {-1, -2, -1, 4}};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) Ke(i, j) = k0[i][j] / 6.0;
    std::vector<double> Fe(4, h * h / 4.0);   // f = 1

    std::vector<double> coords(2 * N);
    for (int j = 0; j <= n; ++j)
        for (int i = 0; i <= n; ++i) {
            coords[2 * (j * (n + 1) + i)] = i * h;
            coords[2 * (j * (n + 1) + i) + 1] = j * h;
        }
    std::vector<int> conn;
    statfem_hip::SystemSolveHip solver(N);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const unsigned a = j * (n + 1) + i;
            std::vector<unsigned> dofs = {a, a + 1, a + 1 + (n + 1), a + (n + 1)};
            for (unsigned d : dofs) conn.push_back((int)d);
            solver.insertToMatrix(Ke, dofs, dofs);
            solver.insertToRhs(Fe, dofs);
        }
    statfem_hip::SystemSolveHip::Constraints bc;
    std::vector<int> bcd;
    for (int j = 0; j <= n; ++j)
        for (int i = 0; i <= n; ++i)
            if (i == 0 || j == 0 || i == n || j == n) {
                bc.emplace_back(j * (n + 1) + i, 0.0);
                bcd.push_back(j * (n + 1) + i);
            }
    solver.applyConstraints(bc, 1.0, 1.0);
    solver.options().tol = 1e-12;
    solver.solveSystem();
    std::vector<unsigned> all(N);
    for (int i = 0; i < N; ++i) all[i] = i;
    std::vector<double> u;
    solver.giveSolution(all, u);

    // the library's whole-mesh path
    sfh_ctx* ctx = solver.context();
    sfh_system* sys = nullptr;
    statfem_hip::detail::check(sfh_system_create(ctx, SFH_QUAD4, 1, N, coords.data(), n * n, conn.data(), &sys),
                               "mesh system");
    statfem_hip::detail::check(sfh_assemble_matrix(sys, SFH_POISSON, 2, nullptr, nullptr, 1.0, 0.0, 1.0), "assemble");
    const double one = 1.0;
    statfem_hip::detail::check(sfh_assemble_rhs(sys, 2, &one, nullptr, 1.0, 0), "rhs");
    statfem_hip::detail::check(sfh_apply_constraints(sys, (int)bcd.size(), bcd.data(), nullptr, 1.0, 1.0), "bc");
    int nd = 0;
    long long nnz = 0;
    statfem_hip::detail::check(sfh_system_info(sys, &nd, &nnz), "info");
    if ((long long)solver.numNonZeros() != nnz) return fail("nnz", (double)nnz);
    std::vector<int> rp1(N + 1), ci1(nnz), rp2(N + 1), ci2(nnz);
    statfem_hip::detail::check(sfh_system_pattern(sys, rp1.data(), ci1.data()), "pattern");
    statfem_hip::detail::check(sfh_system_pattern(solver.handle(), rp2.data(), ci2.data()), "pattern");
    if (rp1 != rp2 || ci1 != ci2) return fail("pattern", 0.0);
    std::vector<double> v1(nnz), v2(nnz);
    statfem_hip::detail::check(sfh_matrix_values(sys, v1.data()), "values");
    statfem_hip::detail::check(sfh_matrix_values(solver.handle(), v2.data()), "values");
    double dv = 0.0;
    for (long long e = 0; e < nnz; ++e) dv = std::fmax(dv, std::fabs(v1[e] - v2[e]));
    if (dv > 1e-14) return fail("values", dv);
    statfem_hip::DeviceVector Xd(ctx, N);
    sfh_solve_opts o;
    sfh_solve_opts_default(&o);
    o.tol = 1e-12;
    statfem_hip::detail::check(sfh_solve_rhs(sys, Xd.data(), &o, nullptr), "solve");
    std::vector<double> ref(N);
    Xd.download(ref.data());
    double du = 0.0, um = 0.0;
    for (int i = 0; i < N; ++i) {
        du = std::fmax(du, std::fabs(u[i] - ref[i]));
        um = std::fmax(um, std::fabs(ref[i]));
    }
    if (du > 1e-9 * um) return fail("solution", du / um);

    // transposed solve with a unit vector = a row of the inverse of the row-zero operator
    std::vector<double> e(N, 0.0), r;
    e[N / 2] = 1.0;
    solver.solveSystemTranspose(e, r);
    Mat Pt(N, 3), Rt(N, 3);
    for (int i = 0; i < N; ++i) { Pt(i, 0) = e[i]; Pt(i, 1) = 0.5 * e[i]; Pt(i, 2) = (i % 7 == 0) ? 1.0 : 0.0; }
    solver.solveSystemTransposeMulti(Pt, Rt);
    double dt = 0.0, rm = 0.0;
    for (int i = 0; i < N; ++i) {
        dt = std::fmax(dt, std::fabs(Rt(i, 0) - r[i]) + std::fabs(Rt(i, 1) - 0.5 * r[i]));
        rm = std::fmax(rm, std::fabs(r[i]));
    }
    if (dt > 1e-9 * rm) return fail("transpose", dt / rm);
    // u_c = (A^-T e_c)^T F for the centre DOF c
    double uc = 0.0;
    std::vector<double> F;
    solver.giveForce(all, F);
    for (int i = 0; i < N; ++i) uc += r[i] * F[i];
    if (std::fabs(uc - u[N / 2]) > 1e-9 * um) return fail("adjoint", std::fabs(uc - u[N / 2]));

    // clearMatrix + the same insertion sequence with doubled stiffness: u halves
    solver.clearMatrix();
    Mat K2 = Ke;
    for (auto& a : K2.a) a *= 2.0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const unsigned a = j * (n + 1) + i;
            std::vector<unsigned> dofs = {a, a + 1, a + 1 + (n + 1), a + (n + 1)};
            solver.insertToMatrix(K2, dofs, dofs);
        }
    solver.applyConstraints(bc, 1.0, 1.0);
    solver.solveSystem(true);
    std::vector<double> u2;
    solver.giveSolution(all, u2);
    double dh = 0.0;
    for (int i = 0; i < N; ++i) dh = std::fmax(dh, std::fabs(2.0 * u2[i] - u[i]));
    if (dh > 1e-9 * um) return fail("clearMatrix", dh / um);
    double f2 = 0.0;
    for (double f : F) f2 += f * f;
    if (std::fabs(solver.normRhs() - std::sqrt(f2)) > 1e-14 * std::sqrt(f2)) return fail("normRhs", solver.normRhs());
    sfh_system_destroy(sys);
    std::printf("OK nnz=%lld max|u|=%.6e\n", nnz, um);
    return 0;
}
