"""glTF scene ingest (irtt/gltf.py), the reference's on-disk scene format.

* A synthetic glTF written here (node T R S and a child with a column-major
  matrix, uint16 and uint32 indices, hyphenated thermal extras, a 145-patch
  sky with two hours of W/m^2 values): positions, materials, fixed flags and
  every sky vertex's kelvin are checked against values computed by hand.
* The reference's own city-block scene, when the reference repository is
  present: 5 014 vertices / 8 964 triangles / 6 objects, sky maximum
  226.8 W/m^2, and equality with the committed fixture the GPU tests use.
"""
import json
import os

import numpy as np
import pytest

SIGMA = 5.670374419e-8
REF_SCENE = "/root/reference/assets/scenes/city-block.gltf"
FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "city_block_gltf.npz")


def _write_gltf(tmp_path, rng):
    blobs, views, accessors = [], [], []

    def add(arr, ctype, typ):
        data = np.ascontiguousarray(arr).tobytes()
        off = sum(len(b) for b in blobs)
        blobs.append(data + b"\0" * (-len(data) % 4))
        views.append({"buffer": 0, "byteOffset": off, "byteLength": len(data)})
        n = arr.shape[0]
        accessors.append({"bufferView": len(views) - 1, "componentType": ctype, "count": n,
                          "type": typ})
        return len(accessors) - 1

    quad = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    qidx = np.array([0, 1, 2, 0, 2, 3], np.uint16)
    # sky: 145 separate quads on a hemisphere-ish arc
    sky_p = rng.normal(size=(145 * 4, 3)).astype(np.float32)
    sky_i = np.arange(145 * 4, dtype=np.uint32).reshape(145, 4)
    sky_idx = np.stack([sky_i[:, 0], sky_i[:, 1], sky_i[:, 2], sky_i[:, 0], sky_i[:, 2],
                        sky_i[:, 3]], 1).ravel().astype(np.uint32)
    values = rng.uniform(0.0, 400.0, (2, 145)).astype(np.float32)
    a_pos, a_idx = add(quad, 5126, "VEC3"), add(qidx, 5123, "SCALAR")
    a_spos, a_sidx = add(sky_p, 5126, "VEC3"), add(sky_idx, 5125, "SCALAR")
    ex_wall = {"kelvin": 300.0, "temperature-fixed": 0, "thickness": 0.25, "density": 2000.0,
               "heat-capacity": 900.0, "heat-conductivity": 1.5, "diffuse-reflectance": 0.3,
               "specular-reflectance": 0.2, "diffuse-emission": 1, "traceable": 1}
    ex_sky = {"kelvin": 1.0, "temperature-fixed": 1, "diffuse-emission": 0, "traceable": 0,
              "sky-values": values.ravel().tolist(), "sky-hours": [10, 11], "sky-hour": 10}
    g = {"asset": {"version": "2.0"}, "scene": 0,
         "scenes": [{"nodes": [0, 2]}],
         "nodes": [{"mesh": 0, "translation": [1.0, 2.0, 3.0],
                    "rotation": [0.0, np.sqrt(0.5), 0.0, np.sqrt(0.5)], "scale": [2.0, 2.0, 2.0],
                    "children": [1]},
                   {"mesh": 0, "matrix": [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 5, 1]},
                   {"mesh": 1}],
         "meshes": [{"name": "wall", "extras": ex_wall,
                     "primitives": [{"attributes": {"POSITION": a_pos}, "indices": a_idx}]},
                    {"name": "sky", "extras": ex_sky,
                     "primitives": [{"attributes": {"POSITION": a_spos}, "indices": a_sidx}]}],
         "accessors": accessors, "bufferViews": views,
         "buffers": [{"uri": "scene.bin", "byteLength": sum(len(b) for b in blobs)}]}
    (tmp_path / "scene.bin").write_bytes(b"".join(blobs))
    path = tmp_path / "scene.gltf"
    path.write_text(json.dumps(g))
    return path, quad, values


def test_synthetic_gltf(tmp_path):
    from irtt import MAT_DIFFUSE_EMISSION, MAT_TRACEABLE, gltf
    rng = np.random.default_rng(0)
    path, quad, values = _write_gltf(tmp_path, rng)
    sd = gltf.load(str(path))
    assert (sd.nv, sd.nt) == (4 + 4 + 580, 2 + 2 + 290)
    # node 0: x' = R_y(90) (2 p) + (1, 2, 3); R_y(90) maps (x, y, z) -> (z, y, -x)
    p = 2.0 * quad.astype(np.float64)
    want0 = np.stack([p[:, 2], p[:, 1], -p[:, 0]], 1) + [1.0, 2.0, 3.0]
    assert np.allclose(sd.xyz[:4], want0, atol=1e-6)
    # node 1 (child of node 0): parent transform after a +5 z translation
    q = 2.0 * (quad.astype(np.float64) + [0.0, 0.0, 5.0])
    want1 = np.stack([q[:, 2], q[:, 1], -q[:, 0]], 1) + [1.0, 2.0, 3.0]
    assert np.allclose(sd.xyz[4:8], want1, atol=1e-5)
    assert np.array_equal(sd.idx[:2], [[0, 1, 2], [0, 2, 3]])
    assert np.array_equal(sd.idx[2:4], [[4, 5, 6], [4, 6, 7]])
    assert np.allclose(sd.rho_d[:4], 0.3) and np.allclose(sd.rho_s[:4], 0.2)
    assert np.allclose(sd.thickness[:4], 0.25) and np.allclose(sd.conductivity[:4], 1.5)
    assert np.all(sd.flags[:4] == (MAT_DIFFUSE_EMISSION | MAT_TRACEABLE))
    assert np.all(sd.flags[4:] == 0)
    assert np.all(sd.fixed[:8] == 0) and np.all(sd.fixed[8:] == 1)
    assert np.allclose(sd.kelvin[:8], 300.0)
    # sky, hour-averaged: each patch's 4 vertices carry max((v / sigma)^1/4, 111.575)
    # of the mean over hours of max(v, sigma 111.575^4)
    vmean = np.maximum(values.astype(np.float64), SIGMA * 111.575 ** 4).mean(0)
    k = np.maximum((vmean / SIGMA) ** 0.25, 111.575)
    assert np.allclose(sd.kelvin[8:].reshape(145, 4), k[:, None], rtol=1e-6)
    # one hour between rows: linear in the values
    sd2 = gltf.load(str(path), hour=0.25)
    v = values[0].astype(np.float64) * 0.75 + values[1].astype(np.float64) * 0.25
    k2 = np.maximum((v / SIGMA) ** 0.25, 111.575)
    assert np.allclose(sd2.kelvin[8:].reshape(145, 4), k2[:, None], rtol=1e-6)


def _fixture_arrays():
    return dict(np.load(FIXTURE, allow_pickle=False))


@pytest.mark.skipif(not os.path.exists(REF_SCENE), reason="reference repository not present")
def test_reference_city_block_scene():
    from irtt import gltf
    sd = gltf.load(REF_SCENE)
    assert (sd.nv, sd.nt) == (5014, 8964)
    objs = sd.meta["objects"]
    assert [o["name"] for o in objs][:2] == ["Building", "Ground"] and len(objs) == 6
    sky = sd.meta["sky"]
    assert len(sky["hours"]) == 24 and abs(sky["values"].max() - 226.807) < 1e-3
    assert sd.fixed.sum() == 580                      # the sky dome is fixed
    fx = _fixture_arrays()
    for k in ("xyz", "idx", "rho_d", "rho_s", "thickness", "density", "heat_capacity",
              "conductivity", "flags", "kelvin", "fixed"):
        assert np.array_equal(getattr(sd, k), fx[k]), k
    assert np.array_equal(sky["kelvin"], fx["sky_kelvin"])


def test_fixture_is_consistent():
    fx = _fixture_arrays()
    assert fx["xyz"].shape == (5014, 3) and fx["idx"].shape == (8964, 3)
    assert fx["idx"].max() == 5013
    objs = json.loads(str(fx["objects"]))
    assert sum(o["triangle_count"] for o in objs) == 8964
    # sky vertices carry their patch's kelvin (the last patch wins on shared vertices)
    k = fx["kelvin"].astype(np.float64)
    pv = fx["sky_patch_vertices"]
    last = {}
    for q in range(145):
        for v in pv[q]:
            last[int(v)] = fx["sky_kelvin"][q]
    assert all(abs(k[v] - val) < 1e-4 for v, val in last.items())
