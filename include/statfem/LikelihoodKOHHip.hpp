// LikelihoodKOHHip — a drop-in of statX's LikelihoodKOH<FORWARDGAUSS, MISMATCHCOV>
// (statX/infer/LikelihoodKOH.hpp:55-160, LikelihoodKOH.ipp) whose dense
// sensor-space algebra runs on an MI355X through the library's likelihood
// state (sfh_lik_*).
//
// Same interface and keys: setSensorValues (n_s x n_o, one column per
// reading), setProjectionMatrix, giveLogLikelihood / giveTypicalReadings /
// giveProjFEconditionedData / giveFEconditionedDataWoodbury with the
// hyperparameters <prefix>rho, mismatch.sigma, mismatch.length and
// mismatch.sensorEps. The forward model provides the projected mean and C_u
// (giveProjectedMeanU / giveProjectedCovarU, LikelihoodKOH.ipp:24,84), and for
// the Woodbury posterior also the device field block C_u P^T, ubar and the
// unprojected prior (ForwardGaussRHSHip's updateSolutionField and
// priorFieldCovariance).
//
// Numerics: the total covariance C = K_s + rho^2 C_u is factored once per
// evaluation (Cholesky, FP64) and the reference's FullPivLU inverses of the
// posterior (LikelihoodKOH.ipp:310-361,441-506) are replaced by the
// algebraically equal Cholesky forms S = C_u + K_s/(rho^2 n_o):
// m = ubar + C_u S^-1 (ybar/rho - ubar), C_{u|y} = C_u - C_u S^-1 C_u.
// Failures end the program like the reference's FTL_VERIFY: a non-SPD
// covariance (LikelihoodKOH.ipp:93) and a non-negative quadratic form (:113).
//
// GPU extension: giveLogLikelihoods evaluates many hyperparameter sets in one
// batched launch (the MCMC chains of statX/infer/MCMC.hpp run them in lockstep).
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "statfem/ForwardGaussRHSHip.hpp"
#include "statfem/MismatchCovarianceHip.hpp"

namespace statfem_hip {

namespace detail {
// row-major dense matrix with the (i, j) / rows() / cols() / resize(r, c) interface
struct Dense {
    std::size_t r = 0, c = 0;
    std::vector<double> a;
    void resize(std::size_t r_, std::size_t c_) { r = r_; c = c_; a.assign(r * c, 0.0); }
    double& operator()(std::size_t i, std::size_t j) { return a[i * c + j]; }
    double operator()(std::size_t i, std::size_t j) const { return a[i * c + j]; }
    std::size_t rows() const { return r; }
    std::size_t cols() const { return c; }
};
}  // namespace detail

template <class FORWARDGAUSS, class MISMATCHCOV = MismatchCovarianceHip>
class LikelihoodKOHHip {
public:
    typedef FORWARDGAUSS ForwardGauss;
    typedef MISMATCHCOV MismatchCov;
    typedef std::map<std::string, double> HpsValueMap;
    static constexpr unsigned numHps = 1;

    LikelihoodKOHHip(ForwardGauss* forwardGauss, MismatchCov* mismatchCov)
        : hypId_("rho"), mismatchCov_(mismatchCov), forwardGauss_(forwardGauss) {
        mismatchCov_->setHypIdPre("mismatch.");
    }
    ~LikelihoodKOHHip() { if (lik_) sfh_lik_destroy(lik_); }
    LikelihoodKOHHip(const LikelihoodKOHHip&) = delete;
    LikelihoodKOHHip& operator=(const LikelihoodKOHHip&) = delete;

    void setHypIdPre(const std::string& preamble) { hypIdPre_ = preamble; }

    template <class MAT>
    void setSensorValues(const MAT& sensorValues) {
        ns_ = static_cast<int>(sensorValues.rows());
        no_ = static_cast<int>(sensorValues.cols());
        Y_.assign(static_cast<std::size_t>(ns_) * no_, 0.0);   // reading-major for the library
        for (int i = 0; i < ns_; ++i)
            for (int k = 0; k < no_; ++k) Y_[static_cast<std::size_t>(k) * ns_ + i] = sensorValues(i, k);
        dirty_ = true;
    }

    template <class MAT>
    void setProjectionMatrix(MAT& pMat) {
        pRows_ = static_cast<int>(pMat.rows());
        pCols_ = static_cast<int>(pMat.cols());
    }

    double giveLogLikelihood(const HpsValueMap& hp) {
        std::vector<double> out;
        std::vector<int> st;
        giveLogLikelihoods(std::vector<HpsValueMap>{hp}, out, st);
        STATFEM_HIP_VERIFY(st[0] == SFH_OK, st[0] == SFH_EQUADFORM ? "non-negative quadratic form in the likelihood"
                                                                    : "total covariance not SPD");
        return out[0];
    }

    // many hyperparameter sets (sharing mismatch.sensorEps) in one batched
    // evaluation; status per set (SFH_OK, SFH_ENOTSPD, SFH_EQUADFORM)
    void giveLogLikelihoods(const std::vector<HpsValueMap>& hps, std::vector<double>& out, std::vector<int>& status) {
        STATFEM_HIP_VERIFY(!hps.empty(), "no hyperparameter sets");
        sync(mismatchCov_->sensorEps(hps[0]));
        std::vector<double> theta;
        for (const auto& hp : hps) {
            STATFEM_HIP_VERIFY(mismatchCov_->sensorEps(hp) == eps_, "sets of one batch share mismatch.sensorEps");
            theta.push_back(rho(hp));
            theta.push_back(mismatchCov_->sigma(hp));
            theta.push_back(mismatchCov_->length(hp));
        }
        out.assign(hps.size(), 0.0);
        status.assign(hps.size(), 0);
        detail::check(sfh_lik_loglik_batch(lik_, static_cast<int>(hps.size()), theta.data(), out.data(), status.data()),
                      "sfh_lik_loglik_batch");
    }

    // y = rho P ubar + L z with L L^T = K_s + rho^2 C_u (LikelihoodKOH.ipp:42-71);
    // z from the library's counter-based normal stream under `seed`
    template <class MAT>
    void giveTypicalReadings(const HpsValueMap& hp, unsigned numReadings, MAT& yTypicalMat, unsigned seed = 0) {
        sync(mismatchCov_->sensorEps(hp));
        std::vector<double> Y(static_cast<std::size_t>(numReadings) * ns_);
        detail::check(sfh_lik_typical_readings(lik_, rho(hp), mismatchCov_->sigma(hp), mismatchCov_->length(hp),
                                               static_cast<int>(numReadings), static_cast<uint64_t>(seed), Y.data()),
                      "sfh_lik_typical_readings");
        detail::resize2(yTypicalMat, ns_, numReadings, 0);
        for (int i = 0; i < ns_; ++i)
            for (unsigned k = 0; k < numReadings; ++k) yTypicalMat(i, k) = Y[static_cast<std::size_t>(k) * ns_ + i];
    }

    // p(P u | y): mean and covariance at the sensors (LikelihoodKOH.ipp:310-361)
    template <class VEC, class MAT>
    void giveProjFEconditionedData(const HpsValueMap& hp, VEC& feMeanCondData, MAT& feCovarCondData) {
        sync(mismatchCov_->sensorEps(hp));
        std::vector<double> m(ns_), c(static_cast<std::size_t>(ns_) * ns_);
        detail::check(sfh_lik_post_proj(lik_, rho(hp), mismatchCov_->sigma(hp), mismatchCov_->length(hp), m.data(),
                                        c.data()),
                      "sfh_lik_post_proj");
        detail::resize1(feMeanCondData, ns_, 0);
        for (int i = 0; i < ns_; ++i) detail::set1(feMeanCondData, i, m[i], 0);
        detail::resize2(feCovarCondData, ns_, ns_, 0);
        for (int i = 0; i < ns_; ++i)
            for (int j = 0; j < ns_; ++j) feCovarCondData(i, j) = c[static_cast<std::size_t>(i) * ns_ + j];
    }

    // p(u | y) on the whole field by the Woodbury expansion (LikelihoodKOH.ipp:441-506):
    // mean (N) and covariance (N x N; the prior C_u is formed with P = I as in the
    // reference, so for moderate N)
    template <class VEC, class MAT>
    void giveFEconditionedDataWoodbury(const HpsValueMap& hp, VEC& feMeanCondData, MAT& feCovarCondData) {
        const int N = fieldMean(hp, feMeanCondData);
        std::vector<double> prior = forwardGauss_->priorFieldCovariance(hp);
        DeviceVector pd(context(), static_cast<std::size_t>(N) * N);
        pd.upload(prior.data());
        detail::check(sfh_lik_post_field_cov(lik_, rho(hp), mismatchCov_->sigma(hp), mismatchCov_->length(hp), N,
                                             forwardGauss_->fieldDevice(), pd.data(), 1, pd.data()),
                      "sfh_lik_post_field_cov");
        pd.download(prior.data());
        detail::resize2(feCovarCondData, N, N, 0);
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) feCovarCondData(i, j) = prior[static_cast<std::size_t>(i) * N + j];
    }

    // GPU extension: the Woodbury mean alone (N), for fields too large for the
    // N x N covariance
    template <class VEC>
    void giveFEconditionedMeanWoodbury(const HpsValueMap& hp, VEC& feMeanCondData) { fieldMean(hp, feMeanCondData); }

private:
    double rho(const HpsValueMap& hp) const { return detail::key(hp, hypIdPre_ + hypId_); }
    sfh_ctx* context() const { return forwardGauss_->solver()->context(); }

    template <class VEC>
    int fieldMean(const HpsValueMap& hp, VEC& feMeanCondData) {
        STATFEM_HIP_VERIFY(pRows_ == 0 || pRows_ == ns_, "projection rows differ from the sensors");
        forwardGauss_->updateSolutionField(hp);
        sync(mismatchCov_->sensorEps(hp));
        const int N = forwardGauss_->numDofs();
        DeviceVector md(context(), N);
        detail::check(sfh_lik_post_field(lik_, rho(hp), mismatchCov_->sigma(hp), mismatchCov_->length(hp), N,
                                         forwardGauss_->fieldDevice(), forwardGauss_->ubarDevice(), md.data()),
                      "sfh_lik_post_field");
        std::vector<double> m(N);
        md.download(m.data());
        detail::resize1(feMeanCondData, N, 0);
        for (int i = 0; i < N; ++i) detail::set1(feMeanCondData, i, m[i], 0);
        return N;
    }

    // the library state holds C_u, the projected mean, the readings, the sensor
    // coordinates and sensorEps: rebuilt when any of them changed
    void sync(double eps) {
        STATFEM_HIP_VERIFY(ns_ > 0 && no_ > 0, "setSensorValues first");
        STATFEM_HIP_VERIFY(mismatchCov_->numSensors() == ns_, "sensor coordinates differ from the readings");
        std::vector<double> mean;
        detail::Dense Cu;
        forwardGauss_->giveProjectedMeanU(mean);
        forwardGauss_->giveProjectedCovarU(Cu);
        STATFEM_HIP_VERIFY(static_cast<int>(mean.size()) == ns_, "forward model and readings differ in sensors");
        if (lik_ && !dirty_ && eps == eps_ && mean == mean_ && Cu.a == Cu_) return;
        if (lik_) sfh_lik_destroy(lik_);
        lik_ = nullptr;
        detail::check(sfh_lik_create(context(), ns_, no_, mismatchCov_->dim(), mismatchCov_->coordinates().data(),
                                     Cu.a.data(), mean.data(), Y_.data(), eps, &lik_),
                      "sfh_lik_create");
        mean_.swap(mean);
        Cu_.swap(Cu.a);
        eps_ = eps;
        dirty_ = false;
    }

    const std::string hypId_;
    std::string hypIdPre_;
    MismatchCov* const mismatchCov_;
    ForwardGauss* const forwardGauss_;
    int ns_ = 0, no_ = 0, pRows_ = 0, pCols_ = 0;
    std::vector<double> Y_, mean_, Cu_;
    double eps_ = -1.0;
    bool dirty_ = true;
    sfh_lik* lik_ = nullptr;
};

}  // namespace statfem_hip
