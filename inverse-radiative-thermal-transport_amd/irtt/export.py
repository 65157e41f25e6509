"""Result exports in the reference's formats
(src/implementations/thermal_renderer/thermal_export.cpp:16-222).

* ``write_vtk``: legacy ASCII VTK unstructured grid of the scene's triangles
  with the per-vertex temperature as ``SCALARS kelvin``; points written as
  (x, -z, y) — the reference swaps Y and Z for its Y-up glTF scenes (:99).
* ``write_matlab``: ``co`` / ``el`` arrays of a MATLAB script, one-based (:158-222).
* ``write_point_values_csv``, ``write_object_averages_csv``,
  ``write_float_vec_csv``: the three CSV writers (:16-49), 5 decimals.
* ``write_transport_view``: one row or column of the operator as a per-vertex
  field (the reference's "Emission to all" / "Absorption from all" views,
  thermal_gui.cpp:48, simple_ray.rgen:413-420), written as VTK.
"""
import numpy as np


def write_vtk(path, xyz, idx, kelvin, swap_yz=True, name="kelvin"):
    xyz = np.asarray(xyz, np.float64)
    idx = np.asarray(idx, np.int64)
    kelvin = np.asarray(kelvin, np.float64)
    if kelvin.shape != (xyz.shape[0],):
        raise ValueError("one value per vertex")
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
        f.write(f"POINT_DATA {pts.shape[0]}\nSCALARS {name} double 1\nLOOKUP_TABLE default\n")
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


def transport_view(scene, vertex, mode="emission"):
    """Per-bin values of one line of the scene's operator: mode "emission"
    gives T[:, vertex] (where the selected bin's emission is absorbed, the
    reference's visualization 0), "absorption" gives T[vertex, :] (where the
    power the selected bin absorbs comes from, visualization 1)."""
    if mode not in ("emission", "absorption"):
        raise ValueError("mode is 'emission' or 'absorption'")
    return scene.transport_line(vertex, column=(mode == "emission")).astype(np.float64)


def write_transport_view(path, xyz, idx, scene, vertex, mode="emission", swap_yz=True):
    """transport_view of a vertex-binned scene written as a VTK scalar field
    named after the mode."""
    v = transport_view(scene, vertex, mode)
    write_vtk(path, xyz, idx, v, swap_yz=swap_yz, name=f"transport_{mode}_{int(vertex)}")
    return v
