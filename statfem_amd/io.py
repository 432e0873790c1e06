"""Readers and writers of the plain-text files the reference's applications
take and produce, so that their input decks drive this library unchanged.

    read_properties   corlib::PropertiesParser "name value" lines, '#' comments
                      (corlib/PropertiesParser.hpp:178-223)
    read_param_map    readParamToMap (del2/apps/reference/laplaceInverse/ioHelpers.hpp:156-181)
    read_matrix_data  readMatrixData: "rows cols" then row-major values (ioHelpers.hpp:127-153)
    read_coord_data   readCoordData: one point per line (ioHelpers.hpp:53-80)
    read_matrix_phi   readMatrixPhi: "row col value" triplets, rows = sensors (ioHelpers.hpp:88-123)
    read_smf          SMF meshes (statfem_amd.mesh.read_smf)
"""
import numpy as np

from .mesh import read_smf  # noqa: F401  (re-exported: the deck's mesh format)


def read_properties(path):
    """name -> value (float where the token parses as a number, else the string)."""
    out = {}
    with open(path) as fh:
        for line in fh:
            t = line.split()
            if not t or t[0].startswith("#") or len(t) < 2:
                continue
            try:
                out[t[0]] = float(t[1])
            except ValueError:
                out[t[0]] = t[1]
    return out


def read_param_map(path):
    """key -> float for every 'key value' pair on non-comment lines."""
    out = {}
    with open(path) as fh:
        for line in fh:
            if not line.strip() or line.startswith("#"):
                continue
            t = line.split()
            for k in range(0, len(t) - 1, 2):
                out[t[k]] = float(t[k + 1])
    return out


def read_matrix_data(path):
    with open(path) as fh:
        tok = fh.read().split()
    r, c = int(tok[0]), int(tok[1])
    vals = np.array([float(x) for x in tok[2:]], np.float64)
    if vals.size != r * c:
        raise ValueError(f"{path}: {vals.size} values for a {r} x {c} matrix")
    return vals.reshape(r, c)


def read_coord_data(path, dim):
    pts = []
    with open(path) as fh:
        for line in fh:
            t = line.split()
            if not t:
                continue
            if len(t) != dim:
                raise ValueError(f"{path}: line with {len(t)} coordinates, expected {dim}")
            pts.append([float(x) for x in t])
    return np.array(pts, np.float64).reshape(-1, dim)


def read_matrix_phi(path):
    """Sensor projection as rows of (DOF columns, values): (cols, vals) lists,
    duplicates summed as Eigen's setFromTriplets does."""
    rows = {}
    nrow = 0
    with open(path) as fh:
        for line in fh:
            t = line.split()
            if len(t) < 3:
                continue
            r, c, v = int(t[0]), int(t[1]), float(t[2])
            rows.setdefault(r, {})
            rows[r][c] = rows[r].get(c, 0.0) + v
            nrow = max(nrow, r + 1)
    cols = [np.array(sorted(rows.get(r, {})), np.int32) for r in range(nrow)]
    vals = [np.array([rows[r][c] for c in sorted(rows.get(r, {}))], np.float64) for r in range(nrow)]
    return cols, vals


def write_param_means(path, names, values, header="# Output file containing the learned hyperparameters"):
    """hpsMeanFile of hyperParamLearning (laplaceInverse/hyperParamLearning.hpp:233-247)."""
    with open(path, "w") as fh:
        fh.write(header + "\n")
        for n, v in zip(names, values):
            fh.write(f"{n}  {v:.6g}\n")


def write_samples(path, natural, values):
    """sampleMCMCFile: each state's natural values then its log target
    (laplaceInverse/hyperParamLearning.hpp:236-246)."""
    with open(path, "w") as fh:
        for x, y in zip(natural, values):
            fh.write("".join(f"{v:.6g}  " for v in x) + f"{y:.6g}\n")
