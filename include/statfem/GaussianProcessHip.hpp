// GaussianProcessHip — a drop-in of statX's GaussianProcess<ExpKernel, NUMPTS>
// (statX/infer/GaussianProcess.hpp:48-270) whose factorisation and
// predictions run on an MI355X (sfh_gp, sfh_gp_predict_cov).
//
// Keys as in the reference: the kernel's "gaussianProcess.sigma" and
// "gaussianProcess.length", and <prefix>eps and the observations
// <prefix>w0 .. <prefix>w{NUMPTS-1} (GaussianProcess.hpp:62-75,98-113,245-250).
// The reference factors K + eps^2 I once in createGaussianProcess and reuses
// it for every prediction; here every call is one library call that fits and
// predicts on the device, and predictMeans / predictVariances (an extension)
// evaluate many test locations in one launch.
#pragma once

#include <string>
#include <vector>

#include "statfem/ForwardGaussRHSHip.hpp"

namespace statfem_hip {

template <unsigned NUMPTS>
class GaussianProcessHip {
public:
    typedef std::map<std::string, double> HpsValueMap;
    const unsigned numPts = NUMPTS;

    explicit GaussianProcessHip(sfh_ctx* ctx = nullptr) : ctx_(ctx ? ctx : default_context()), kernelPre_("gaussianProcess.") {}

    void setHypIdPre(const std::string& preamble) { hypIdPre_ = preamble; }

    // X: NUMPTS x dim observation locations (rows)
    template <class MAT>
    void createGaussianProcess(HpsValueMap& hParams, const MAT& xMat) {
        STATFEM_HIP_VERIFY(static_cast<unsigned>(xMat.rows()) == NUMPTS, "one row per observation point");
        hParams_ = hParams;
        dim_ = static_cast<int>(xMat.cols());
        X_.resize(static_cast<std::size_t>(NUMPTS) * dim_);
        for (unsigned i = 0; i < NUMPTS; ++i)
            for (int d = 0; d < dim_; ++d) X_[static_cast<std::size_t>(i) * dim_ + d] = xMat(i, d);
        // the factorisation must exist (GaussianProcess.hpp:266-268): checked with zero observations
        std::vector<double> y0(NUMPTS, 0.0);
        double ll = 0.0;
        detail::check(sfh_gp(ctx_, NUMPTS, dim_, X_.data(), y0.data(), sigma(), length(), eps(), 0, nullptr, &ll,
                             nullptr, nullptr),
                      "GaussianProcess covariance not SPD");
    }

    void setObservationValues(const HpsValueMap& hParams) {
        y_.resize(NUMPTS);
        for (unsigned i = 0; i < NUMPTS; ++i) y_[i] = detail::key(hParams, hypIdPre_ + "w" + std::to_string(i));
    }

    template <class VEC>
    double predictMean(const VEC& xStar) const {
        std::vector<double> m, v;
        predict(point(xStar), m, v);
        return m[0];
    }
    template <class VEC>
    double predictVariance(const VEC& xStar) const {
        std::vector<double> m, v;
        predict(point(xStar), m, v);
        return v[0];
    }
    // xStars: np x dim test locations; covMat np x np
    template <class MAT, class OUT>
    void predictCovariance(const MAT& xStars, OUT& covMat) const {
        STATFEM_HIP_VERIFY(static_cast<int>(xStars.cols()) == dim_, "test locations of the observations' dimension");
        const int np = static_cast<int>(xStars.rows());
        std::vector<double> Xp(static_cast<std::size_t>(np) * dim_), cov(static_cast<std::size_t>(np) * np);
        for (int p = 0; p < np; ++p)
            for (int d = 0; d < dim_; ++d) Xp[static_cast<std::size_t>(p) * dim_ + d] = xStars(p, d);
        detail::check(sfh_gp_predict_cov(ctx_, NUMPTS, dim_, X_.data(), obs().data(), sigma(), length(), eps(), np,
                                         Xp.data(), cov.data()),
                      "sfh_gp_predict_cov");
        detail::resize2(covMat, np, np, 0);
        for (int i = 0; i < np; ++i)
            for (int j = 0; j < np; ++j) covMat(i, j) = cov[static_cast<std::size_t>(i) * np + j];
    }
    double giveLogLikelihood() const {
        double ll = 0.0;
        detail::check(sfh_gp(ctx_, NUMPTS, dim_, X_.data(), obs().data(), sigma(), length(), eps(), 0, nullptr, &ll,
                             nullptr, nullptr),
                      "sfh_gp");
        return ll;
    }

    // extension: means and variances at np locations (np x dim, row-major) in one call
    void predict(const std::vector<double>& Xp, std::vector<double>& mean, std::vector<double>& var) const {
        const int np = static_cast<int>(Xp.size() / (dim_ > 0 ? dim_ : 1));
        mean.assign(np, 0.0);
        var.assign(np, 0.0);
        double ll = 0.0;
        detail::check(sfh_gp(ctx_, NUMPTS, dim_, X_.data(), obs().data(), sigma(), length(), eps(), np, Xp.data(), &ll,
                             mean.data(), var.data()),
                      "sfh_gp");
    }

private:
    double sigma() const { return detail::key(hParams_, kernelPre_ + "sigma"); }
    double length() const { return detail::key(hParams_, kernelPre_ + "length"); }
    double eps() const { return detail::key(hParams_, hypIdPre_ + "eps"); }
    const std::vector<double>& obs() const {
        STATFEM_HIP_VERIFY(y_.size() == NUMPTS, "setObservationValues first");
        return y_;
    }
    template <class VEC>
    std::vector<double> point(const VEC& x) const {
        STATFEM_HIP_VERIFY(static_cast<int>(x.size()) == dim_, "test location of the observations' dimension");
        std::vector<double> p(dim_);
        for (int d = 0; d < dim_; ++d) p[d] = detail::get(x, d);
        return p;
    }

    sfh_ctx* ctx_;
    std::string hypIdPre_;
    const std::string kernelPre_;
    HpsValueMap hParams_;
    int dim_ = 0;
    std::vector<double> X_, y_;
};

}  // namespace statfem_hip
