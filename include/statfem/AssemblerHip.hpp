// AssemblerHip — whole-mesh element assembly on an MI355X, the GPU form of
// statFEM's per-element assembly loop
//
//     MatrixComputeAndAssembleFun maf(...);   mesh.iterateOverElements(maf);
//     (corlib/MatrixAssembler.hpp:42-100,166-242 over ComputeElementMatrix,
//      corlib/ComputeElementMatrix.hpp:62-77)
//
// and of the right-hand side loop Integrator + BodyForce + VectorAssembler
// (corlib/Integrator.hpp:62-74, corlib/BodyForce.hpp:66-71,
// corlib/VectorAssembler.hpp:101-114). A GPU cannot be driven one element at a
// time through the assembler's callback, so one call assembles every element:
// the mesh (node coordinates in file order, element connectivity in the
// reference's local node order) is uploaded once, the CSR pattern
// (Triplet::finishAssembly, corlib/Triplet.hpp:126-170,264-339) and the
// element-to-CSR map are built on the device, and each assembly overwrites
// the operator's values.
//
// The target solver is a SystemSolveHip: its first assembly hands it the
// device system, and from then on its constraints, solves and accessors act
// on that system. Right-hand sides can go to any target with the SystemSolve
// concept's insertToRhs (for instance CovarianceForcingHip, which receives the
// unit-load integrals exactly as CovarianceForcing does in the reference's
// drivers, del2/apps/reference/laplaceRHSinfer/inference.hpp:103-120).
//
// Header-only C++17 over statfem_hip.h; errors end the program like FTL_VERIFY.
#pragma once

#include <string>
#include <vector>

#include "statfem/SystemSolveHip.hpp"

namespace statfem_hip {

class AssemblerHip {
public:
    // Element kinds and integrands of statfem_hip.h (SFH_LINE2 .. SFH_HEX8;
    // SFH_POISSON, SFH_ELASTICITY, SFH_MATERN). coords: numNodes x dim
    // (dim 1, 2, 3 by element kind), conn: numElements x nodes per element.
    AssemblerHip(int elementKind, int dofsPerNode, int numNodes, const std::vector<double>& coords,
                 int numElements, const std::vector<int>& conn, sfh_ctx* ctx = nullptr)
        : ctx_(ctx ? ctx : default_context()), kind_(elementKind), dofpn_(dofsPerNode), nnodes_(numNodes),
          nel_(numElements), coords_(coords), conn_(conn) {
        STATFEM_HIP_VERIFY(dofpn_ == 1 || dofpn_ == 3, "dofs per node must be 1 or 3");
    }
    ~AssemblerHip() {
        if (sys_ && !bound_) sfh_system_destroy(sys_);
    }
    AssemblerHip(const AssemblerHip&) = delete;
    AssemblerHip& operator=(const AssemblerHip&) = delete;

    unsigned numDofs() const { return static_cast<unsigned>(nnodes_ * dofpn_); }
    int numElements() const { return nel_; }

    // Declares the mesh a tensor grid (nodes x fastest) with these axis
    // coordinates: separable forcing covariance and the multigrid solver.
    void declareGrid(const std::vector<double>& gx, const std::vector<double>& gy = {},
                     const std::vector<double>& gz = {}) {
        gx_ = gx;
        gy_ = gy;
        gz_ = gz;
        if (sys_) applyGrid();
    }

    // MatrixAssembler over all elements: K_e from the integrand `physics` with
    // nq1d Gauss points per direction (per-element parameters p1/p2 of size
    // numElements, or a single value each), times `factor`, into the solver.
    void assemble(SystemSolveHip* solver, int physics, int nq1d, const std::vector<double>& p1,
                  const std::vector<double>& p2 = {}, double factor = 1.0) {
        bind(solver);
        const double s1 = p1.size() == 1 ? p1[0] : 0.0;
        const double s2 = p2.size() == 1 ? p2[0] : 0.0;
        DeviceVector d1, d2;
        if (p1.size() > 1) d1 = upload(p1);
        if (p2.size() > 1) d2 = upload(p2);
        STATFEM_HIP_VERIFY(p1.size() == 1 || static_cast<int>(p1.size()) == nel_, "p1: one value or one per element");
        STATFEM_HIP_VERIFY(p2.size() <= 1 || static_cast<int>(p2.size()) == nel_, "p2: one value or one per element");
        detail::check(sfh_assemble_matrix(sys_, physics, nq1d, p1.size() > 1 ? d1.data() : nullptr,
                                          p2.size() > 1 ? d2.data() : nullptr, s1, s2, factor),
                      "sfh_assemble_matrix");
    }

    // Integrator + BodyForce + VectorAssembler over all elements for the
    // load f(x) = fconst + sum_K fnodal[K] phi_K(x) (fnodal: numDofs or
    // empty), times factor; the element vectors are added to `target` with
    // insertToRhs (all DOFs at once).
    template <class TARGET>
    void assembleRhs(TARGET* target, int nq1d, const std::vector<double>& fconst,
                     const std::vector<double>& fnodal = {}, double factor = 1.0) {
        std::vector<double> F = bodyForce(nq1d, fconst, fnodal, factor);
        std::vector<unsigned> dofs(F.size());
        for (unsigned i = 0; i < dofs.size(); ++i) dofs[i] = i;
        target->insertToRhs(F, dofs);
    }
    void assembleRhs(SystemSolveHip* solver, int nq1d, const std::vector<double>& fconst,
                     const std::vector<double>& fnodal = {}, double factor = 1.0) {
        bind(solver);
        solver->addForce(bodyForce(nq1d, fconst, fnodal, factor));
    }

    // The unit-load integrals int phi_i (ElementMatern::lumpedMassIntegrand,
    // del2/fem/ElementMatern.ipp:96-119), host copy.
    std::vector<double> lumpedMass(int nq1d) {
        ensureSystem();
        DeviceVector d(ctx_, numDofs());
        detail::check(sfh_lumped_mass(sys_, nq1d, d.data()), "sfh_lumped_mass");
        std::vector<double> h(numDofs());
        d.download(h.data());
        return h;
    }

private:
    DeviceVector upload(const std::vector<double>& h) {
        DeviceVector d(ctx_, h.size());
        d.upload(h.data());
        return d;
    }
    void ensureSystem() {
        if (sys_) return;
        STATFEM_HIP_VERIFY(coords_.size() % static_cast<std::size_t>(nnodes_) == 0, "coordinates: numNodes x dim");
        detail::check(sfh_system_create(ctx_, kind_, dofpn_, nnodes_, coords_.data(), nel_, conn_.data(), &sys_),
                      "sfh_system_create");
        if (!gx_.empty()) applyGrid();
    }
    void applyGrid() {
        detail::check(sfh_system_set_grid(sys_, static_cast<int>(gx_.size()), gy_.empty() ? 1 : static_cast<int>(gy_.size()),
                                          gz_.empty() ? 1 : static_cast<int>(gz_.size()), gx_.data(),
                                          gy_.empty() ? nullptr : gy_.data(), gz_.empty() ? nullptr : gz_.data()),
                      "sfh_system_set_grid");
    }
    // the first assembly into a solver hands it the device system
    void bind(SystemSolveHip* solver) {
        STATFEM_HIP_VERIFY(solver != nullptr, "no solver");
        STATFEM_HIP_VERIFY(solver->numDofs() == numDofs(), "solver size differs from the mesh's DOFs");
        if (bound_ == solver) return;
        STATFEM_HIP_VERIFY(bound_ == nullptr, "an AssemblerHip feeds one solver");
        ensureSystem();
        solver->bindMesh(sys_);
        bound_ = solver;
    }
    std::vector<double> bodyForce(int nq1d, const std::vector<double>& fconst, const std::vector<double>& fnodal,
                                  double factor) {
        ensureSystem();
        STATFEM_HIP_VERIFY(static_cast<int>(fconst.size()) == dofpn_, "fconst: one value per DOF of a node");
        DeviceVector dn;
        if (!fnodal.empty()) {
            STATFEM_HIP_VERIFY(fnodal.size() == numDofs(), "fnodal: one value per DOF");
            dn = upload(fnodal);
        }
        // the library assembles into the system's F; the solver's own F is the
        // host copy uploaded before each solve, so nothing is lost here
        detail::check(sfh_assemble_rhs(sys_, nq1d, fconst.data(), fnodal.empty() ? nullptr : dn.data(), factor, 0),
                      "sfh_assemble_rhs");
        std::vector<double> F(numDofs());
        detail::check(sfh_rhs(sys_, F.data()), "sfh_rhs");
        return F;
    }

    sfh_ctx* ctx_;
    int kind_, dofpn_, nnodes_, nel_;
    std::vector<double> coords_;
    std::vector<int> conn_;
    std::vector<double> gx_, gy_, gz_;
    sfh_system* sys_ = nullptr;          // owned by bound_ once bound
    SystemSolveHip* bound_ = nullptr;
};

}  // namespace statfem_hip
