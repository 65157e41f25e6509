"""The sky model (irtt/sky.py): the reference's hour-of-year, WEA filter and
sky-matrix parser (sky/skymatrix.hpp), and the Tregenza dome of
thermal_sky.cpp, checked against dates, hand-made files and the dome's
geometry (patch counts per ring, inclinations, inward normals, distances)."""
import datetime as dt

import numpy as np
import pytest

from irtt import scenes, sky


def test_hour_of_year():
    assert sky.datetime_to_hoy("2022-01-01T00:00:00+00:00") == 0
    assert sky.datetime_to_hoy("2022-06-01T12:00:00+00:00") == 151 * 24 + 12
    assert sky.datetime_to_hoy("2024-12-31T23:59:59+02:00") == 365 * 24 + 23  # leap year
    for d in (dt.datetime(2023, 3, 7, 5, 30), dt.datetime(2023, 11, 2, 17, 0)):
        s = d.strftime("%Y-%m-%dT%H:%M:%S+00:00")
        assert sky.datetime_to_hoy(s) == int((d - dt.datetime(d.year, 1, 1)).total_seconds() // 3600)


def test_filter_wea_keeps_header_and_hour_range():
    lines = [f"h{i}" for i in range(6)] + [f"d{k}" for k in range(48)]
    out = sky.filter_wea(lines, 10, 14)
    assert out == [f"h{i}" for i in range(6)] + ["d10", "d11", "d12", "d13"]
    assert sky.filter_wea(lines, 5, 5) == lines[:6]


def test_parse_sky_mtx_order_and_broadband():
    assert abs(sum(sky.BROADBAND) - 1.0) < 1e-8
    rng = np.random.default_rng(0)
    rgb = rng.uniform(0, 500, (145, 3))
    text = "\n".join([f"header {i}" for i in range(8)] + ["0 0 0"] +
                     [f"{r:.6f} {g:.6f} {b:.6f}" for r, g, b in rgb]) + "\n"
    v = sky.parse_sky_mtx(text)
    want = 0.01 * (0.265074126 * rgb[:, 0] + 0.670114631 * rgb[:, 1] + 0.064811243 * rgb[:, 2])
    assert v.shape == (145,)
    assert np.allclose(v, want, rtol=1e-6)
    # a truncated matrix leaves zeros after its last line
    short = sky.parse_sky_mtx("\n".join(text.splitlines()[:9 + 40]))
    assert np.allclose(short[:40], want[:40], rtol=1e-6) and not short[40:].any()


def test_sky_value_to_kelvin():
    sig = 5.670374419e-8
    assert np.isclose(sky.sky_value_to_kelvin(800.0), (800.0 / sig) ** 0.25)
    assert sky.sky_value_to_kelvin(0.0) == sky.MINIMAL_KELVIN
    assert sky.sky_value_to_kelvin(1e-3) == sky.MINIMAL_KELVIN


@pytest.mark.parametrize("up", ["y", "z"])
def test_tregenza_dome_geometry(up):
    c = np.array([3.0, -2.0, 5.0])
    r = 10.0
    xyz, tri, dirs = sky.tregenza_dome(c, r, up)
    assert xyz.shape == (580, 3) and tri.shape == (290, 3) and dirs.shape == (145, 3)
    upv = np.array([0.0, 1.0, 0.0]) if up == "y" else np.array([0.0, 0.0, 1.0])
    # zenith patch first, then rings of 6, 12, 18, 24, 24, 30, 30 at 12 degree steps
    incl = np.degrees(np.arccos(np.clip(dirs @ upv, -1, 1)))
    k = 0
    for n, want in zip((1, 6, 12, 18, 24, 24, 30, 30), range(0, 96, 12)):
        assert np.allclose(incl[k:k + n], want, atol=1e-9)
        if n > 1:   # evenly spread azimuths
            az = np.sort(np.angle(dirs[k:k + n] @ np.array([1.0, 0, 0]) +
                                  1j * (dirs[k:k + n] @ np.cross(upv, [1.0, 0, 0]))))
            assert np.allclose(np.diff(az), 2 * np.pi / n, atol=1e-9)
        k += n
    assert k == 145
    for i in range(145):
        q = xyz[4 * i:4 * i + 4]
        centre = q.mean(0)
        assert np.isclose(np.linalg.norm(centre - c), 2.2 * r)
        assert np.allclose((centre - c) / np.linalg.norm(centre - c), dirs[i], atol=1e-12)
        # each corner 2.2 r from the patch centre, the quad in the patch's tangent plane
        assert np.allclose(np.linalg.norm(q - centre, axis=1), 2.2 * r)
        assert np.allclose((q - centre) @ dirs[i], 0.0, atol=1e-9)
        for t in tri[2 * i:2 * i + 2]:
            nrm = np.cross(xyz[t[1]] - xyz[t[0]], xyz[t[2]] - xyz[t[0]])
            assert nrm @ (c - centre) > 0   # emits toward the scene


def test_add_sky_appends_fixed_collimated_emitters():
    sd = scenes.parallel_plates(n=4)
    vals = np.linspace(0.0, 300.0, 145)
    out = sky.add_sky(sd, vals)
    assert out.nv == sd.nv + 580 and out.nt == sd.nt + 290
    s = out.meta["sky"]
    assert np.all(out.fixed[s["vertices"]] == 1)
    t0, t1 = s["triangles"]
    assert np.all(out.flags[t0:t1] == 0)          # not traceable, not diffuse
    assert np.all(out.rho_d[t0:t1] == 0.0)         # emissivity 1
    kel = out.kelvin[s["vertices"]].reshape(145, 4)
    assert np.allclose(kel, sky.sky_value_to_kelvin(vals)[:, None], rtol=1e-6)
    v, k = sky.sky_vertex_kelvin(out, vals[::-1])
    assert np.array_equal(v, s["vertices"]) and np.allclose(k[::4], sky.sky_value_to_kelvin(vals[::-1]), rtol=1e-6)
    with pytest.raises(ValueError):
        sky.add_sky(sd, vals[:10])
