// MismatchCovarianceHip — a drop-in of statX's MismatchCovariance<ExpKernel>
// (statX/infer/MismatchCovariance.hpp:62-121): the model-discrepancy plus
// sensor-noise covariance K_s = [k_d(x_k, x_l)] + eps^2 I at the sensors, with
// k_d(x, y) = sigma^2 exp(-|x - y|^2 / (2 length^2)) (statX/infer/ExpKernel.hpp:74-92).
//
// Keys as in the reference: the kernel's "mismatch.sigma" and "mismatch.length"
// and the noise <prefix>sensorEps, the prefix set by the likelihood
// (LikelihoodKOH.hpp:64-70 sets "mismatch."). The covariance itself is small
// (n_s x n_s) and is formed on the host, as the reference does; on the GPU
// path LikelihoodKOHHip builds it inside its batched likelihood kernels from
// the same sensor coordinates and keys.
//
// giveCovarianceDerv returns the derivatives of the squared-exponential
// kernel itself: d/dsigma = 2 k / sigma, d/dlength = k |x - y|^2 / length^3.
// (The reference's ExpKernel::giveDerivHParam evaluates k at (x2, x2), which
// drops the exponential and the sign of the length derivative,
// ExpKernel.hpp:110-114; that form is not reproduced.)
#pragma once

#include <cmath>
#include <map>
#include <string>
#include <vector>

#include "statfem/ForwardGaussRHSHip.hpp"

namespace statfem_hip {

class MismatchCovarianceHip {
public:
    typedef std::map<std::string, double> HpsValueMap;
    static constexpr unsigned numHps = 1;

    // sensorCoords: n_s points, each with size() and operator[] or operator()
    template <class COORDS>
    explicit MismatchCovarianceHip(const COORDS& sensorCoords) : hypId_("sensorEps"), kernelPre_("mismatch.") {
        ns_ = static_cast<int>(sensorCoords.size());
        dim_ = ns_ > 0 ? static_cast<int>(sensorCoords[0].size()) : 0;
        xs_.reserve(static_cast<std::size_t>(ns_) * dim_);
        for (int k = 0; k < ns_; ++k) {
            STATFEM_HIP_VERIFY(static_cast<int>(sensorCoords[k].size()) == dim_, "sensor coordinates of one dimension");
            for (int d = 0; d < dim_; ++d) xs_.push_back(detail::get(sensorCoords[k], d));
        }
    }

    void setHypIdPre(const std::string& preamble) { hypIdPre_ = preamble; }

    template <class MAT>
    void giveCovariance(const HpsValueMap& hyperParams, MAT& kMat) const {
        const double eps = sensorEps(hyperParams);
        const double sigma = detail::key(hyperParams, kernelPre_ + "sigma");
        const double length = detail::key(hyperParams, kernelPre_ + "length");
        detail::resize2(kMat, ns_, ns_, 0);
        for (int k = 0; k < ns_; ++k)
            for (int l = 0; l < ns_; ++l) kMat(k, l) = value(k, l, sigma, length) + (k == l ? eps * eps : 0.0);
    }

    // dK_s / dtheta for theta = mismatch.sigma, mismatch.length (MismatchCovariance.hpp:103-118)
    template <class MATMAP>
    void giveCovarianceDerv(const HpsValueMap& hyperParams, MATMAP& mapHpsDMat) const {
        const double sigma = detail::key(hyperParams, kernelPre_ + "sigma");
        const double length = detail::key(hyperParams, kernelPre_ + "length");
        auto& ds = mapHpsDMat[kernelPre_ + "sigma"];
        auto& dl = mapHpsDMat[kernelPre_ + "length"];
        detail::resize2(ds, ns_, ns_, 0);
        detail::resize2(dl, ns_, ns_, 0);
        for (int k = 0; k < ns_; ++k)
            for (int l = 0; l < ns_; ++l) {
                const double v = value(k, l, sigma, length);
                ds(k, l) = 2.0 * v / sigma;
                dl(k, l) = v * dist2(k, l) / (length * length * length);
            }
    }

    // ---- used by LikelihoodKOHHip
    double sensorEps(const HpsValueMap& hp) const { return detail::key(hp, hypIdPre_ + hypId_); }
    double sigma(const HpsValueMap& hp) const { return detail::key(hp, kernelPre_ + "sigma"); }
    double length(const HpsValueMap& hp) const { return detail::key(hp, kernelPre_ + "length"); }
    const std::vector<double>& coordinates() const { return xs_; }
    int numSensors() const { return ns_; }
    int dim() const { return dim_; }

private:
    double dist2(int k, int l) const {
        double s = 0.0;
        for (int d = 0; d < dim_; ++d) {
            const double t = xs_[static_cast<std::size_t>(k) * dim_ + d] - xs_[static_cast<std::size_t>(l) * dim_ + d];
            s += t * t;
        }
        return s;
    }
    double value(int k, int l, double sigma, double length) const {
        return sigma * sigma * std::exp(-dist2(k, l) / (2. * length * length));
    }

    int ns_ = 0, dim_ = 0;
    std::vector<double> xs_;
    const std::string hypId_;
    std::string hypIdPre_;
    const std::string kernelPre_;
};

}  // namespace statfem_hip
