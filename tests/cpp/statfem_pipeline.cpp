// Drives the statX drop-ins the way del2/apps/reference/laplaceRHSinfer drives
// the reference classes (inference.hpp:60-176, fem.hpp:19-96): AssemblerHip
// assembles the LINE2 bar into a SystemSolveHip and the unit-load integrals
// into a CovarianceForcingHip, both constrained at the bar's ends;
// ForwardGaussRHSHip takes the phiMat projection and the allParam.dat
// hyperparameters; MismatchCovarianceHip the sensor coordinates;
// LikelihoodKOHHip the readings. At the recorded posterior means of the MCMC
// (watchDog.0.mk:20-22) the Woodbury field mean must match the recorded
// inferredFEcondData to its print precision. The GaussianProcessHip drop-in
// reproduces the gaussianProcess app's golden (watchDog.0.mk:18-19).
//
// Usage: statfem_pipeline <golden dir> <output file>. Prints OK and writes the
// computed quantities for the oracle comparison of the test suite.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "statfem/AssemblerHip.hpp"
#include "statfem/ForwardGaussRHSHip.hpp"
#include "statfem/GaussianProcessHip.hpp"
#include "statfem/LikelihoodKOHHip.hpp"
#include "statfem/MismatchCovarianceHip.hpp"

namespace {

struct Mat {   // dense row-major matrix with the interface the classes use
    std::size_t r = 0, c = 0;
    std::vector<double> a;
    Mat() = default;
    Mat(std::size_t r_, std::size_t c_) { resize(r_, c_); }
    void resize(std::size_t r_, std::size_t c_) { r = r_; c = c_; a.assign(r * c, 0.0); }
    double& operator()(std::size_t i, std::size_t j) { return a[i * c + j]; }
    double operator()(std::size_t i, std::size_t j) const { return a[i * c + j]; }
    std::size_t rows() const { return r; }
    std::size_t cols() const { return c; }
};

std::vector<std::vector<double>> readRows(const std::string& path, int skip = 0) {
    std::ifstream in(path);
    std::vector<std::vector<double>> rows;
    std::string line;
    int n = 0;
    while (std::getline(in, line)) {
        if (n++ < skip) continue;
        std::istringstream ss(line);
        std::vector<double> row;
        double v;
        while (ss >> v) row.push_back(v);
        if (!row.empty()) rows.push_back(row);
    }
    return rows;
}

std::map<std::string, double> readParams(const std::string& path) {
    std::ifstream in(path);
    std::map<std::string, double> m;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        std::string k;
        double v;
        if (ss >> k >> v) m[k] = v;
    }
    return m;
}

int fail(const char* what, double v) {
    std::printf("FAIL %s %.6e\n", what, v);
    return 1;
}

void write(std::ofstream& out, const char* name, const std::vector<double>& v) {
    out << name << " " << v.size();
    char buf[32];
    for (double x : v) {
        std::snprintf(buf, sizeof buf, " %.17g", x);
        out << buf;
    }
    out << "\n";
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::printf("usage: statfem_pipeline <golden dir> <output file>\n");
        return 2;
    }
    using namespace statfem_hip;
    const std::string g = std::string(argv[1]) + "/laplaceRHSinfer/";
    std::ofstream out(argv[2]);

    // ---- mesh (SMF: '!' header lines, "nNodes nElements", nodes, connectivity)
    std::ifstream smf(g + "bar.smf");
    std::string line;
    std::vector<std::string> body;
    while (std::getline(smf, line))
        if (!line.empty() && line[0] != '!') body.push_back(line);
    int nn = 0, ne = 0;
    std::istringstream(body[0]) >> nn >> ne;
    std::vector<double> x(nn);
    for (int i = 0; i < nn; ++i) std::istringstream(body[1 + i]) >> x[i];
    std::vector<int> conn(2 * ne);
    for (int e = 0; e < ne; ++e) std::istringstream(body[1 + nn + e]) >> conn[2 * e] >> conn[2 * e + 1];

    // ---- FE system (fem.hpp:19-96): conductivity 1, f = pi^2/5, u = 0 at both ends
    AssemblerHip assembler(SFH_LINE2, 1, nn, x, ne, conn);
    SystemSolveHip solver(nn);
    solver.options().tol = 1e-13;
    assembler.assemble(&solver, SFH_POISSON, 3, {1.0});
    assembler.assembleRhs(&solver, 3, {M_PI * M_PI / 5.0});
    SystemSolveHip::Constraints bc;
    for (int i = 0; i < nn; ++i)
        if (std::fabs(x[i]) < 1e-8 || std::fabs(x[i] - 1.0) < 1e-8) bc.emplace_back(i, 0.0);
    solver.applyConstraints(bc, 1.0, 1.0);
    solver.solveSystem();

    // ---- forcing covariance weights (inference.hpp:89-120)
    CovarianceForcingHip<1> covarForcing(nn);
    for (int i = 0; i < nn; ++i) covarForcing.insertToCoordinates(std::vector<double>{x[i]});
    assembler.assembleRhs(&covarForcing, 3, {1.0});
    covarForcing.applyConstraints(bc, 1.0, 1.0);

    // ---- projection (readMatrixPhi, ioHelpers.hpp:88-123) and parameters
    const auto trip = readRows(g + "phiMat.dat");
    int ns = 0;
    for (const auto& t : trip) ns = std::max(ns, (int)t[0] + 1);
    Mat P(ns, nn);
    for (const auto& t : trip) P((int)t[0], (int)t[1]) = t[2];
    std::map<std::string, double> hp = readParams(g + "allParam.dat");

    ForwardGaussRHSHip<> forwardGauss(&solver, &covarForcing, 0.001);
    forwardGauss.options().tol = 1e-13;
    forwardGauss.setProjectionMatrix(P);
    forwardGauss.updateSolution(hp);

    // ---- mismatch covariance and likelihood (inference.hpp:141-160)
    const auto xsRows = readRows(g + "sensorCoordinates.dat");
    std::vector<std::vector<double>> sensorCoord(xsRows.begin(), xsRows.end());
    MismatchCovarianceHip misMatchCov(sensorCoord);
    LikelihoodKOHHip<ForwardGaussRHSHip<>> likelihood(&forwardGauss, &misMatchCov);
    const auto yRows = readRows(g + "sensorValues.dat", 1);
    Mat Y(yRows.size(), yRows[0].size());
    for (std::size_t i = 0; i < Y.rows(); ++i)
        for (std::size_t k = 0; k < Y.cols(); ++k) Y(i, k) = yRows[i][k];
    likelihood.setSensorValues(Y);

    // the MCMC posterior means the reference recorded (watchDog.0.mk:20-22)
    hp["rho"] = 0.804412;
    hp["mismatch.sigma"] = 0.0191724;
    hp["mismatch.length"] = 0.0740285;
    const double logL = likelihood.giveLogLikelihood(hp);
    std::vector<std::map<std::string, double>> sets(3, hp);
    sets[1]["rho"] = 1.1;
    sets[2]["mismatch.length"] = 0.2;
    std::vector<double> lls;
    std::vector<int> sts;
    likelihood.giveLogLikelihoods(sets, lls, sts);
    if (std::fabs(lls[0] - logL) > 1e-12 * std::fabs(logL) || sts[1] != SFH_OK || sts[2] != SFH_OK)
        return fail("batched likelihood", lls[0] - logL);
    Mat Ks;
    misMatchCov.giveCovariance(hp, Ks);
    std::map<std::string, Mat> dK;
    misMatchCov.giveCovarianceDerv(hp, dK);

    std::vector<double> pm, wm;
    Mat pc, wc;
    likelihood.giveProjFEconditionedData(hp, pm, pc);
    likelihood.setProjectionMatrix(P);
    likelihood.giveFEconditionedDataWoodbury(hp, wm, wc);
    const auto ref = readRows(g + "watchDog.0.mk.inferredFEcondData.dat");
    double dm = 0.0;
    for (int i = 0; i < nn; ++i) dm = std::fmax(dm, std::fabs(wm[i] - ref[i][0]));
    if (dm > 1e-6) return fail("Woodbury mean vs golden", dm);   // printed to 6 digits

    std::vector<double> mean, u;
    Mat Cu;
    forwardGauss.giveProjectedMeanU(mean);
    forwardGauss.giveProjectedCovarU(Cu);
    std::vector<unsigned> all(nn);
    for (int i = 0; i < nn; ++i) all[i] = i;
    solver.giveSolution(all, u);
    write(out, "u", u);
    write(out, "covvec", covarForcing.covVec());
    write(out, "mean", mean);
    write(out, "Cu", Cu.a);
    write(out, "logL", {logL});
    write(out, "logL_batch", lls);
    write(out, "Ks", Ks.a);
    write(out, "dKs_sigma", dK["mismatch.sigma"].a);
    write(out, "dKs_length", dK["mismatch.length"].a);
    write(out, "proj_mean", pm);
    write(out, "proj_cov", pc.a);
    write(out, "wood_mean", wm);
    write(out, "wood_cov", wc.a);

    // ---- GaussianProcess drop-in against the gaussianProcess golden
    const std::string gg = std::string(argv[1]) + "/gp/";
    std::map<std::string, double> ghp = readParams(gg + "gpHps.dat");
    const auto gx = readRows(gg + "gpCoordinates.dat", 1);
    Mat X(gx.size(), 1);
    for (std::size_t i = 0; i < gx.size(); ++i) X(i, 0) = gx[i][0];
    GaussianProcessHip<6> gp;
    gp.setHypIdPre("gaussianProcess.");
    gp.createGaussianProcess(ghp, X);
    gp.setObservationValues(ghp);
    const double gll = gp.giveLogLikelihood();
    if (std::fabs(gll - (-4.837171259)) > 1e-9) return fail("GP log-likelihood", gll);
    const auto xp = readRows(gg + "predictionCoord.dat");
    const auto gmean = readRows(gg + "watchDog.0.mk.predictionMean.dat");
    const auto gvar = readRows(gg + "watchDog.0.mk.predictionVar.dat");
    Mat Xs(xp.size(), 1);
    double de = 0.0;
    for (std::size_t p = 0; p < xp.size(); ++p) {
        Xs(p, 0) = xp[p][0];
        const std::vector<double> xs1{xp[p][0]};
        de = std::fmax(de, std::fabs(gp.predictMean(xs1) - gmean[p][0]));
        de = std::fmax(de, std::fabs(gp.predictVariance(xs1) - gvar[p][0]));
    }
    if (de > 1e-8) return fail("GP predictions vs golden", de);
    Mat gcov;
    gp.predictCovariance(Xs, gcov);
    double dd = 0.0;
    for (std::size_t p = 0; p < xp.size(); ++p) dd = std::fmax(dd, std::fabs(gcov(p, p) - gvar[p][0]));
    if (dd > 1e-8) return fail("GP predictCovariance diagonal vs golden", dd);
    write(out, "gp_cov", gcov.a);
    std::printf("OK logL %.12g Woodbury mean max|d| vs golden %.2e, N %d, sensors %d\n", logL, dm, nn, ns);
    return 0;
}
