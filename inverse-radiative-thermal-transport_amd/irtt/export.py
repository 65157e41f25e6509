"""Result exports in the reference's formats
(src/implementations/thermal_renderer/thermal_export.cpp:16-222).

* ``write_vtk``: legacy ASCII VTK unstructured grid of the scene's triangles
  with the per-vertex temperature as ``SCALARS kelvin``; points written as
  (x, -z, y) — the reference swaps Y and Z for its Y-up glTF scenes (:99).
* ``write_matlab``: ``co`` / ``el`` arrays of a MATLAB script, one-based (:158-222).
* ``write_point_values_csv``, ``write_object_averages_csv``,
  ``write_float_vec_csv``: the three CSV writers (:16-49), 5 decimals.
"""
import numpy as np


def write_vtk(path, xyz, idx, kelvin, swap_yz=True):
    xyz = np.asarray(xyz, np.float64)
    idx = np.asarray(idx, np.int64)
    kelvin = np.asarray(kelvin, np.float64)
    if kelvin.shape != (xyz.shape[0],):
        raise ValueError("one temperature per vertex")
    pts = np.stack([xyz[:, 0], -xyz[:, 2], xyz[:, 1]], 1) if swap_yz else xyz
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 2.0\nVTK exported mesh\nASCII\n")
        f.write("DATASET UNSTRUCTURED_GRID\n")
        f.write(f"POINTS {pts.shape[0]} double\n")
        np.savetxt(f, pts, fmt="%.9g")
        f.write(f"CELLS {idx.shape[0]} {4 * idx.shape[0]}\n")
        np.savetxt(f, np.hstack([np.full((idx.shape[0], 1), 3), idx]), fmt="%d")
        f.write(f"CELL_TYPES {idx.shape[0]}\n")
        f.write("5\n" * idx.shape[0])
        f.write(f"POINT_DATA {pts.shape[0]}\nSCALARS kelvin double 1\nLOOKUP_TABLE default\n")
        np.savetxt(f, kelvin, fmt="%.9g")


def write_matlab(path, xyz, idx):
    xyz = np.asarray(xyz, np.float64)
    idx = np.asarray(idx, np.int64) + 1
    with open(path, "w") as f:
        f.write(f"co = zeros( {xyz.shape[0]},3);\nel = zeros( {idx.shape[0]},3);\n")
        for i, p in enumerate(xyz, 1):
            f.write(f"co({i},:)=[{p[0]:.12g} {p[1]:.12g} {p[2]:.12g}];\n")
        for i, t in enumerate(idx, 1):
            f.write(f"el({i},:)=[ {t[0]} {t[1]} {t[2]}];\n")


def write_point_values_csv(path, values):
    with open(path, "w") as f:
        f.write('"temperature", "Points:0", "Points:1", "Points:2"\n')
        for v in np.asarray(values, np.float32):
            f.write(f"{v:.5f}\n")


def write_object_averages_csv(path, names, averages):
    with open(path, "w") as f:
        for n, a in zip(names, averages):
            f.write(f"{n}; {a:.5f}\n")


def write_float_vec_csv(path, values):
    with open(path, "w") as f:
        for v in np.asarray(values, np.float32):
            f.write(f"{v:.5f}; ")
