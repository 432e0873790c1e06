// ForwardGaussRHSHip and CovarianceForcingHip — drop-ins of statX's
// ForwardGaussRHS<SOLVER, COVARFORCING> (statX/infer/ForwardGaussRHS.hpp:61-139)
// and CovarianceForcing<DOF, ExpKernel> (statX/infer/CovarianceForcing.hpp:70-114,
// CovarianceForcing.ipp:14-109) whose solves and covariance products run on an
// MI355X.
//
// The reference loops over the n_s rows of the projection P with one
// transposed SparseLU solve each (ForwardGaussRHS.hpp:105-112), then forms
// C_u = R C_f R^T from a dense O(N^2) C_f (CovarianceForcing.ipp:37-109).
// Here updateSolution is one call into the library (sfh_forward_gauss): all
// n_s transposed solves at once, C_f R^T matrix-free (separable on a declared
// tensor grid), and C_u = R (C_f R^T) + (sigma_f regEps)^2 P P^T with FP64
// MFMA products. The keys are the reference's: <prefix>sigma and
// <prefix>length of the forcing kernel (prefix "forcing."), and the
// regularisation's "forcing.sigma" (ForwardGaussRHS.hpp:130-136).
//
// The solver must hold a whole-mesh system (SystemSolveHip after
// AssemblerHip::assemble); its right-hand side is the assembled force
// (giveForce, ForwardGaussRHS.hpp:115-124). CovarianceForcingHip receives the
// unit-load integrals through insertToRhs and the constraints through
// applyConstraints, as CovarianceForcing does in the reference's drivers; its
// covVec is what the GPU's C_f uses.
//
// Beyond the reference: updateSolution can also leave the field-by-sensor
// prior covariance C_u P^T (N x n_s) and ubar = A^-1 F on the device, the
// inputs of the Woodbury posterior (LikelihoodKOHHip); priorFieldCovariance
// forms the unprojected C_u (the reference's updateSolution with P = I,
// LikelihoodKOH.ipp:476-480) for moderate N.
#pragma once

#include <algorithm>
#include <map>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "statfem/SystemSolveHip.hpp"

namespace statfem_hip {

namespace detail {
template <class M>
auto resize2(M& m, std::size_t r, std::size_t c, int) -> decltype(m.resize(r, c), void()) { m.resize(r, c); }
template <class M>
void resize2(M&, std::size_t, std::size_t, long) {}
template <class V>
auto resize1(V& v, std::size_t n, int) -> decltype(v.resize(n), void()) { v.resize(n); }
template <class V>
void resize1(V&, std::size_t, long) {}
template <class V>
auto set1(V& v, std::size_t i, double x, int) -> decltype(v(i) = x, void()) { v(i) = x; }
template <class V>
auto set1(V& v, std::size_t i, double x, long) -> decltype(v[i] = x, void()) { v[i] = x; }

inline double key(const std::map<std::string, double>& hp, const std::string& k) {
    const auto it = hp.find(k);
    STATFEM_HIP_VERIFY(it != hp.end(), ("hyperparameter missing: " + k).c_str());
    return it->second;
}
}  // namespace detail

// ------------------------------------------------------------------------
template <unsigned DOF = 1>
class CovarianceForcingHip {
public:
    typedef std::map<std::string, double> HpsValueMap;
    typedef std::vector<std::pair<unsigned, double>> VectorConstraint;
    static constexpr unsigned dof = DOF;

    // cutEps > 0 drops small C_f entries in the reference (CovarianceForcing.ipp:72,91);
    // the GPU product keeps every entry, so only 0 is accepted
    explicit CovarianceForcingHip(const unsigned numNodes, const double cutEps = 0.0)
        : numNodes_(numNodes), covVec_(numNodes * DOF, 0.0), kernelPre_("forcing.") {
        STATFEM_HIP_VERIFY(cutEps == 0.0, "CovarianceForcingHip keeps every C_f entry (cutEps = 0)");
        static_assert(DOF == 1, "the GPU forcing covariance acts on scalar DOFs");
    }

    void setHypIdPre(const std::string& preamble) { hypIdPre_ = preamble; }

    template <class VEC, class IDX>
    void insertToRhs(const VEC& eCovVec, const IDX& dofIndices) {
        for (std::size_t i = 0; i < dofIndices.size(); ++i) covVec_[dofIndices[i]] += detail::get(eCovVec, i);
    }
    // coordinates come from the mesh of the solver's whole-mesh system
    template <class X>
    void insertToCoordinates(const X& /*nCoords*/) { ++numCoords_; }

    void applyConstraints(const VectorConstraint& constraints, const double /*factor*/, const double /*scalar*/) {
        for (const auto& c : constraints) {
            bcIndices_.push_back(c.first);
            covVec_[c.first] = 0.;
        }
    }
    void getConstrainedIndices(std::vector<unsigned>& bcIndices) { bcIndices = bcIndices_; }

    // forcing kernel hyperparameters (ExpKernel keys with the "forcing." preamble)
    double sigma(const HpsValueMap& hp) const { return detail::key(hp, kernelPre_ + "sigma"); }
    double length(const HpsValueMap& hp) const { return detail::key(hp, kernelPre_ + "length"); }
    const std::vector<double>& covVec() const { return covVec_; }
    unsigned numNodes() const { return numNodes_; }

private:
    unsigned numNodes_;
    std::size_t numCoords_ = 0;
    std::vector<double> covVec_;
    std::vector<unsigned> bcIndices_;
    std::string hypIdPre_, kernelPre_;
};

// ------------------------------------------------------------------------
template <class SOLVER = SystemSolveHip, class COVARFORCING = CovarianceForcingHip<1>>
class ForwardGaussRHSHip {
public:
    typedef SOLVER Solver;
    typedef COVARFORCING CovarForcing;
    typedef std::map<std::string, double> HpsValueMap;

    ForwardGaussRHSHip(Solver* solver, CovarForcing* covarForcing, const double regEps = 1.e-4)
        : solver_(solver), covarForcing_(covarForcing), regEps_(regEps) {
        sfh_solve_opts_default(&opts_);
    }

    void setHypIdPre(const std::string& preamble) { hypIdPre_ = preamble; }

    // P: n_s x N dense (rows are sensors); stored as its nonzeros per row
    template <class MAT>
    void setProjectionMatrix(MAT& pMat) {
        ns_ = static_cast<int>(pMat.rows());
        N_ = static_cast<int>(pMat.cols());
        std::vector<std::vector<std::pair<int, double>>> rows(ns_);
        pn_ = 1;
        for (int s = 0; s < ns_; ++s) {
            for (int j = 0; j < N_; ++j) {
                const double v = pMat(s, j);
                if (v != 0.0) rows[s].emplace_back(j, v);
            }
            pn_ = std::max<int>(pn_, static_cast<int>(rows[s].size()));
        }
        pcols_.assign(static_cast<std::size_t>(ns_) * pn_, 0);
        pvals_.assign(static_cast<std::size_t>(ns_) * pn_, 0.0);
        for (int s = 0; s < ns_; ++s)
            for (std::size_t q = 0; q < rows[s].size(); ++q) {
                pcols_[static_cast<std::size_t>(s) * pn_ + q] = rows[s][q].first;
                pvals_[static_cast<std::size_t>(s) * pn_ + q] = rows[s][q].second;
            }
    }

    template <class VEC>
    void giveProjectedMeanU(VEC& projMean) const {
        STATFEM_HIP_VERIFY(!feMean_.empty(), "giveProjectedMeanU before updateSolution");
        detail::resize1(projMean, feMean_.size(), 0);
        for (std::size_t i = 0; i < feMean_.size(); ++i) detail::set1(projMean, i, feMean_[i], 0);
    }
    template <class MAT>
    void giveProjectedCovarU(MAT& projCovmat) const {
        STATFEM_HIP_VERIFY(!feCovMat_.empty(), "giveProjectedCovarU before updateSolution");
        detail::resize2(projCovmat, ns_, ns_, 0);
        for (int a = 0; a < ns_; ++a)
            for (int b = 0; b < ns_; ++b) projCovmat(a, b) = feCovMat_[static_cast<std::size_t>(a) * ns_ + b];
    }

    // mean P A^-1 f and C_u = R C_f R^T + (sigma_f regEps)^2 P P^T
    void updateSolution(const HpsValueMap& params) { update(params, false); }

    // GPU extension: the same, keeping C_u P^T (N x n_s) and ubar on the device
    void updateSolutionField(const HpsValueMap& params) { update(params, true); }
    const double* fieldDevice() const { return field_.data(); }
    const double* ubarDevice() const { return ubar_.data(); }
    int numSensors() const { return ns_; }
    int numDofs() const { return N_; }

    // The unprojected prior covariance C_u (N x N, row-major host), as the
    // reference forms it with P = I (LikelihoodKOH.ipp:476-480); the
    // projection and the projected moments are left as they were.
    std::vector<double> priorFieldCovariance(const HpsValueMap& params) {
        const auto keep = std::make_tuple(ns_, pn_, pcols_, pvals_, feMean_, feCovMat_);
        ns_ = N_;
        pn_ = 1;
        pcols_.resize(N_);
        pvals_.assign(N_, 1.0);
        for (int i = 0; i < N_; ++i) pcols_[i] = i;
        update(params, false);
        std::vector<double> full;
        full.swap(feCovMat_);
        std::tie(ns_, pn_, pcols_, pvals_, feMean_, feCovMat_) = keep;
        return full;
    }

    sfh_solve_opts& options() { return opts_; }
    const sfh_solve_stats& lastStats() const { return stats_; }
    Solver* solver() const { return solver_; }

private:
    void update(const HpsValueMap& params, bool wantField) {
        STATFEM_HIP_VERIFY(ns_ > 0, "updateSolution before setProjectionMatrix");
        sfh_system* sys = solver_->handle();   // uploads the right-hand side
        STATFEM_HIP_VERIFY(static_cast<int>(solver_->numDofs()) == N_, "projection width differs from the DOFs");
        // the caller's covVec (unit-load integrals with constraints), when given
        const auto& cv = covarForcing_->covVec();
        bool given = false;
        for (double v : cv) given = given || v != 0.0;
        if (given) STATFEM_HIP_VERIFY(cv.size() == static_cast<std::size_t>(N_), "covVec size differs from the DOFs");
        detail::check(sfh_system_set_covvec(sys, given ? cv.data() : nullptr), "sfh_system_set_covvec");
        const double sigma = covarForcing_->sigma(params), length = covarForcing_->length(params);
        // the regularisation reads forcing.sigma itself (ForwardGaussRHS.hpp:130-133)
        const double regSigma = detail::key(params, "forcing.sigma");
        STATFEM_HIP_VERIFY(regSigma == sigma, "forcing.sigma of the kernel and of the regularisation differ");
        feMean_.assign(ns_, 0.0);
        feCovMat_.assign(static_cast<std::size_t>(ns_) * ns_, 0.0);
        if (wantField) {   // (the field buffers of an earlier call stay valid otherwise)
            field_ = DeviceVector(solver_->context(), static_cast<std::size_t>(N_) * ns_);
            ubar_ = DeviceVector(solver_->context(), N_);
        }
        detail::check(sfh_forward_gauss(sys, ns_, pn_, pcols_.data(), pvals_.data(), sigma, length, regEps_, &opts_,
                                        feMean_.data(), feCovMat_.data(), wantField ? field_.data() : nullptr,
                                        wantField ? ubar_.data() : nullptr, &stats_),
                      "sfh_forward_gauss");
    }

    Solver* solver_;
    CovarForcing* covarForcing_;
    double regEps_;
    std::string hypIdPre_;
    int ns_ = 0, N_ = 0, pn_ = 0;
    std::vector<int> pcols_;
    std::vector<double> pvals_;
    std::vector<double> feMean_, feCovMat_;
    DeviceVector field_, ubar_;
    sfh_solve_opts opts_;
    sfh_solve_stats stats_{};
};

}  // namespace statfem_hip
