"""Rows and columns of the stored operator (irtt_get_transport_line): the
reference's "Emission to all" (T[:, v]) and "Absorption from all" (T[v, :])
views, equal entry for entry to the downloaded operator in both layouts, and
written as a VTK field that parses back."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("layout", [0, 1])
def test_transport_line_equals_operator(irtt, layout, tmp_path):
    from irtt import export, scenes
    sd = scenes.city_block(seed=3, nx=2, ny=2, target_tris=2000)
    print(f"[view] city 2x2 layout={layout}", flush=True)
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        s.build_transport(threads_per_tri=32, rays_per_thread=4, batch_seed=5, layout=layout)
        T = s.transport()
        for v in (0, 17, s.n_bins // 2, s.n_bins - 1):
            assert np.array_equal(s.transport_line(v, column=True), T[:, v])
            assert np.array_equal(s.transport_line(v, column=False), T[v, :])
        v = 17
        got = export.write_transport_view(tmp_path / "view.vtk", sd.xyz, sd.idx, s, v,
                                          mode="emission")
        with pytest.raises(irtt.IrttError):
            s.transport_line(s.n_bins)
    assert np.array_equal(got, T[:, v].astype(np.float64))
    lines = (tmp_path / "view.vtk").read_text().splitlines()
    p0 = lines.index(f"POINT_DATA {sd.nv}")
    assert lines[p0 + 1] == f"SCALARS transport_emission_{v} double 1"
    assert np.allclose(np.loadtxt(lines[p0 + 3:]), T[:, v], rtol=1e-8, atol=0)
