"""The laplaceInverse input deck (tests/golden/laplaceInverse) as arrays, for
the conductivity-uncertainty tests."""
import os

import numpy as np

from statfem_amd import io

from .conftest import GOLDEN

DECK = os.path.join(GOLDEN, "laplaceInverse")
IDS = [f"gaussianProcess.w{i}" for i in range(5)] + ["mismatch.sensorEps"]


def load():
    d = {}
    d["mesh"] = io.read_smf(os.path.join(DECK, "bar.smf"))
    hp = io.read_param_map(os.path.join(DECK, "allHyperParam.dat"))
    d["hp"] = hp
    d["anchors"] = io.read_matrix_data(os.path.join(DECK, "gpCoordinates.dat"))
    d["Y"] = io.read_matrix_data(os.path.join(DECK, "sensorValuesRead5.dat"))      # ns x no
    d["xs"] = io.read_coord_data(os.path.join(DECK, "sensorCoordinates.dat"), 1)
    d["cols"], d["vals"] = io.read_matrix_phi(os.path.join(DECK, "phiMat.dat"))
    d["bc"] = np.array([0, d["mesh"].nnodes - 1], np.int32)
    d["gp"] = (hp["gaussianProcess.sigma"], hp["gaussianProcess.length"], hp["gaussianProcess.eps"])
    d["w0"] = np.array([hp[k] for k in IDS[:-1]])
    d["pm"] = np.array([np.log(hp["priorMean." + k]) for k in IDS[:-1]] + [hp["priorMean." + IDS[-1]]])
    d["pv"] = np.array([hp["priorCov." + k] for k in IDS])
    d["init"] = np.log(np.array([hp["init." + k] for k in IDS]))
    d["propsd"] = np.array([hp["propStDev." + k] for k in IDS])
    d["golden"] = np.loadtxt(os.path.join(DECK, "logMLsamples_every10.dat"))
    d["hparmean"] = io.read_param_map(os.path.join(DECK, "hParMean.dat"))
    return d


def projection(d, n):
    P = np.zeros((len(d["cols"]), n))
    for i, (c, v) in enumerate(zip(d["cols"], d["vals"])):
        P[i, c] = v
    return P


def oracle_target(orc, d, theta, mean):
    hp = d["hp"]
    return orc.lhs_log_target(theta, mean, d["Y"].T.copy(), d["xs"], hp["rho"], hp["mismatch.sigma"],
                              hp["mismatch.length"], d["pm"], d["pv"])
