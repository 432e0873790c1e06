"""Readers and writers of the plain-text files the reference's applications
take and produce, so that their input decks drive this library unchanged.

    read_properties   corlib::PropertiesParser "name value" lines, '#' comments
                      (corlib/PropertiesParser.hpp:178-223)
    read_param_map    readParamToMap (del2/apps/reference/laplaceInverse/ioHelpers.hpp:156-181)
    read_matrix_data  readMatrixData: "rows cols" then row-major values (ioHelpers.hpp:127-153)
    read_coord_data   readCoordData: one point per line (ioHelpers.hpp:53-80)
    read_matrix_phi   readMatrixPhi: "row col value" triplets, rows = sensors (ioHelpers.hpp:88-123)
    read_smf          SMF meshes (statfem_amd.mesh.read_smf)
    write_vtu         VTUwriter's ASCII unstructured grid with point and cell data
                      (corlib/VTUwriter.ipp:22-160, corlib/ShapeToVtk.hpp:30-41)
"""
import numpy as np

from .mesh import HEX8, LINE2, QUAD4, TRI3, read_smf  # noqa: F401  (read_smf: the deck's mesh format)

# VTK cell types of the linear elements (corlib/ShapeToVtk.hpp:31-40)
_VTK_CELL = {LINE2: 3, TRI3: 5, QUAD4: 9, HEX8: 12}


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


def _data_array(fh, vtype, name, values, ncomp=None):
    v = np.asarray(values)
    nc = 1 if v.ndim == 1 else v.shape[1]
    if ncomp is not None and ncomp > nc:       # pad to 3 components as the reference does for 2D data
        v = np.concatenate([v.reshape(len(v), nc), np.zeros((len(v), ncomp - nc))], axis=1)
        nc = ncomp
    head = f'      <DataArray type="{vtype}"'
    if nc > 1 or ncomp is not None:
        head += f' NumberOfComponents="{nc}"'
    fh.write(head + f' Name="{name}" format="ascii">\n')
    rows = v.reshape(len(v), -1)
    fmt = "{:d}" if vtype == "Int32" else "{:.17g}"
    for r in rows:
        fh.write("        " + " ".join(fmt.format(int(x) if vtype == "Int32" else float(x)) for x in r) + "\n")
    fh.write("      </DataArray>\n")


def write_vtu(path, mesh, point_data=None, cell_data=None):
    """VTUwriter (corlib/VTUwriter.ipp:22-160): ASCII UnstructuredGrid, points
    as 3-component Float64 "Coordinates", cells as Int32 connectivity (the
    element's node order), offsets and VTK types; point data with 2 components
    padded to 3 (as the reference does for ParaView), then cell data.
    point_data / cell_data: {name: (N,) or (N, c) arrays}."""
    X = np.asarray(mesh.coords, np.float64).reshape(mesh.nnodes, -1)
    conn = np.asarray(mesh.conn, np.int64)
    nen = conn.shape[1]
    with open(path, "w") as fh:
        fh.write('<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" byte_order="LittleEndian">\n')
        fh.write(f'  <UnstructuredGrid>\n  <Piece NumberOfPoints="{mesh.nnodes}"  NumberOfCells="{mesh.nel}">\n')
        fh.write("    <Points>\n")
        _data_array(fh, "Float64", "Coordinates", X, ncomp=3)
        fh.write("    </Points>\n    <Cells>\n")
        _data_array(fh, "Int32", "connectivity", conn.reshape(mesh.nel, nen))
        _data_array(fh, "Int32", "offsets", (np.arange(mesh.nel) + 1) * nen)
        _data_array(fh, "Int32", "types", np.full(mesh.nel, _VTK_CELL[mesh.kind]))
        fh.write("    </Cells>\n")
        if point_data:
            fh.write("    <PointData>\n")
            for name, v in point_data.items():
                v = np.asarray(v, np.float64)
                _data_array(fh, "Float64", name, v, ncomp=3 if (v.ndim == 2 and v.shape[1] == 2) else None)
            fh.write("    </PointData>\n")
        if cell_data:
            fh.write("    <CellData>\n")
            for name, v in cell_data.items():
                _data_array(fh, "Float64", name, np.asarray(v, np.float64))
            fh.write("    </CellData>\n")
        fh.write("  </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n")
