"""Exports in the reference's formats (irtt/export.py, thermal_export.cpp:16-222),
written for a small scene and parsed back."""
import numpy as np


def test_vtk_matlab_csv_round_trip(tmp_path):
    from irtt import export, scenes
    sd = scenes.closed_box(2)
    k = np.linspace(280.0, 320.0, sd.nv)
    export.write_vtk(tmp_path / "t.vtk", sd.xyz, sd.idx, k)
    lines = (tmp_path / "t.vtk").read_text().splitlines()
    assert lines[0] == "# vtk DataFile Version 2.0" and lines[3] == "DATASET UNSTRUCTURED_GRID"
    assert lines[4] == f"POINTS {sd.nv} double"
    pts = np.loadtxt(lines[5:5 + sd.nv])
    x = sd.xyz.astype(np.float64)
    assert np.allclose(pts, np.stack([x[:, 0], -x[:, 2], x[:, 1]], 1))   # Y and Z swapped
    c0 = 5 + sd.nv
    assert lines[c0] == f"CELLS {sd.nt} {4 * sd.nt}"
    cells = np.loadtxt(lines[c0 + 1:c0 + 1 + sd.nt], dtype=np.int64)
    assert np.all(cells[:, 0] == 3) and np.array_equal(cells[:, 1:], sd.idx)
    t0 = c0 + 1 + sd.nt
    assert lines[t0] == f"CELL_TYPES {sd.nt}" and set(lines[t0 + 1:t0 + 1 + sd.nt]) == {"5"}
    p0 = t0 + 1 + sd.nt
    assert lines[p0:p0 + 3] == [f"POINT_DATA {sd.nv}", "SCALARS kelvin double 1",
                                "LOOKUP_TABLE default"]
    assert np.allclose(np.loadtxt(lines[p0 + 3:]), k)

    export.write_matlab(tmp_path / "g.m", sd.xyz, sd.idx)
    m = (tmp_path / "g.m").read_text().splitlines()
    assert m[0] == f"co = zeros( {sd.nv},3);" and m[1] == f"el = zeros( {sd.nt},3);"
    assert m[2 + sd.nv].startswith("el(1,:)=[ ") and m[2 + sd.nv].endswith(
        f"{sd.idx[0][0] + 1} {sd.idx[0][1] + 1} {sd.idx[0][2] + 1}];")

    export.write_point_values_csv(tmp_path / "d.csv", k)
    d = (tmp_path / "d.csv").read_text().splitlines()
    assert d[0] == '"temperature", "Points:0", "Points:1", "Points:2"'
    assert np.allclose([float(v) for v in d[1:]], k, atol=5e-6)
    export.write_object_averages_csv(tmp_path / "o.csv", ["a", "b"], [1.0, 2.5])
    assert (tmp_path / "o.csv").read_text() == "a; 1.00000\nb; 2.50000\n"
    export.write_float_vec_csv(tmp_path / "v.csv", [0.5, 1.25])
    assert (tmp_path / "v.csv").read_text() == "0.50000; 1.25000; "
