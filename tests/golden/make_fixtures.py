"""Copies the reference's own golden data files (plain-text inputs and expected
outputs) into tests/golden/. These are data fixtures, not source: mesh, sensor
and parameter tables plus the recorded outputs the reference's regression
checks compare against.

Run once in a container that has the reference tree at /root/reference:
    python tests/golden/make_fixtures.py
"""
import os
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

FILES = {
    # dense GP known answer (statX/apps/reference/gaussianProcess/watchDog.0.mk:19 logL -4.837171259)
    "gp": ["statX/apps/reference/gaussianProcess/gpHps.dat",
           "statX/apps/reference/gaussianProcess/gpCoordinates.dat",
           "statX/apps/reference/gaussianProcess/predictionCoord.dat",
           "statX/apps/reference/gaussianProcess/watchDog.0.mk.predictionMean.dat",
           "statX/apps/reference/gaussianProcess/watchDog.0.mk.predictionVar.dat"],
    # 1D statFEM end-to-end (del2/apps/reference/laplaceRHSinfer/watchDog.0.mk:17-22)
    "laplaceRHSinfer": ["del2/apps/reference/laplaceRHSinfer/bar.smf",
                        "del2/apps/reference/laplaceRHSinfer/allParam.dat",
                        "del2/apps/reference/laplaceRHSinfer/sensorValues.dat",
                        "del2/apps/reference/laplaceRHSinfer/sensorCoordinates.dat",
                        "del2/apps/reference/laplaceRHSinfer/phiMat.dat",
                        "del2/apps/reference/laplaceRHSinfer/watchDog.0.mk.inferredFEcondData.dat"],
    # conductivity (LHS) inverse problem (del2/apps/reference/laplaceInverse/watchDog.0.mk:21-29)
    "laplaceInverse": ["del2/apps/reference/laplaceInverse/bar.smf",
                       "del2/apps/reference/laplaceInverse/input.dat",
                       "del2/apps/reference/laplaceInverse/inferenceInput.dat",
                       "del2/apps/reference/laplaceInverse/allHyperParam.dat",
                       "del2/apps/reference/laplaceInverse/gpCoordinates.dat",
                       "del2/apps/reference/laplaceInverse/phiMat.dat",
                       "del2/apps/reference/laplaceInverse/sensorCoordinates.dat",
                       "del2/apps/reference/laplaceInverse/sensorValuesRead5.dat",
                       "del2/apps/reference/laplaceInverse/hParMean.dat"],
    # RW-MH statistics (statX/apps/reference/mcmcGaussian)
    "mcmcGaussian": ["statX/apps/reference/mcmcGaussian/input.dat"],
}


def main():
    for sub, files in FILES.items():
        os.makedirs(os.path.join(HERE, sub), exist_ok=True)
        for f in files:
            shutil.copyfile(os.path.join(REF, f), os.path.join(HERE, sub, os.path.basename(f)))
    # summary statistics of the recorded 50,000-step chain (the 1.3 MB chain
    # files themselves are not copied)
    import numpy as np
    s = np.loadtxt(os.path.join(REF, "statX/apps/reference/mcmcGaussian/watchDog.0.mk.sample.dat"))
    l = np.loadtxt(os.path.join(REF, "statX/apps/reference/mcmcGaussian/watchDog.0.mk.likelihood.dat"))
    moves = float(np.mean(s[1:] != s[:-1]))
    with open(os.path.join(HERE, "mcmcGaussian", "chain_summary.txt"), "w") as fh:
        fh.write("# recorded reference chain (watchDog.0.mk.sample.dat / .likelihood.dat)\n")
        fh.write(f"n {s.size}\nmean {s.mean():.10g}\nsd {s.std():.10g}\nmove_fraction {moves:.10g}\n")
        fh.write("acceptance_reported 0.24592\n")
        # log-density column check: max |l - log N(s;1,0.25)|
        ln = np.log(1.0 / (np.sqrt(2 * np.pi) * 0.5) * np.exp(-(s - 1.0) ** 2 / (2 * 0.25)))
        fh.write(f"max_abs_logpdf_dev {np.max(np.abs(l - ln)):.3g}\n")
    # every 10th row of the recorded laplaceInverse chain (states after burn-in
    # and their log target; the 236 KB file itself is not copied)
    rows = open(os.path.join(REF, "del2/apps/reference/laplaceInverse/logMLsamples.dat")).read().splitlines()
    with open(os.path.join(HERE, "laplaceInverse", "logMLsamples_every10.dat"), "w") as fh:
        fh.write("\n".join(rows[::10]) + "\n")
    with open(os.path.join(HERE, "gp", "expected.txt"), "w") as fh:
        fh.write("# statX/apps/reference/gaussianProcess/watchDog.0.mk:19\nlogL -4.837171259\n")
    with open(os.path.join(HERE, "laplaceRHSinfer", "expected.txt"), "w") as fh:
        fh.write("# del2/apps/reference/laplaceRHSinfer/watchDog.0.mk:20-22 (MCMC posterior means)\n")
        fh.write("rho 0.804412\nmismatch.sigma 0.0191724\nmismatch.length 0.0740285\n")


if __name__ == "__main__":
    main()
