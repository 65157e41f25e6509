"""Scene ingest in the reference's on-disk format: glTF 2.0 (.gltf + .bin).

The reference loads scenes with tinygltf (src/engine/importer/gltf.cpp) and
reads each model's thermal properties from the glTF ``extras`` of its mesh,
with hyphenated keys (gltf.cpp:70-103 copies the extras into custom
properties; thermal_scene.cpp:76-108 maps them onto ObjectProperties):

    kelvin, temperature-fixed, thickness, density, heat-capacity,
    heat-conductivity, diffuse-reflectance, specular-reflectance,
    diffuse-emission, traceable

A missing key reads as 0, as the reference's custom-property lookup does.
Vertices and triangles are numbered globally in scene-model order, mesh by
mesh (ThermalData::load, thermal_data.cpp:83-173), with the node transforms
applied.  A model whose extras carry ``sky-values`` is a sky dome of 145
Tregenza patches (6 indices each) with ``sky-hours`` rows of W/m^2 values: its
vertices take the kelvin of their patch, sky_value_to_kelvin(v) =
max((v / sigma)^(1/4), 111.575 K) of the hour-average of max(v, sigma 111.575^4)
per patch (SkyData::setHourByTime, thermal_data.hpp:19-176, with its default
``averaged``), or of one hour's row (linear between rows) when ``hour`` is
given.  Units: the reference's kelvin and seconds unit factors are 1.
"""
import json
import os

import numpy as np

from . import MAT_DIFFUSE_EMISSION, MAT_TRACEABLE
from . import scenes as _scenes

SIGMA = 5.670374419e-8           # thermal_common.hpp:39
SKY_VALUE_COUNT = 145            # Tregenza patches, thermal_data.hpp:27
SKY_MINIMAL_KELVIN = 111.575     # thermal_data.hpp:29

_COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16,
              5125: np.uint32, 5126: np.float32}
_WIDTH = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT4": 16}


def sky_value_to_kelvin(value, minimum=SKY_MINIMAL_KELVIN):
    """thermal_common.cpp:12-18: (value / sigma)^(1/4), clamped below."""
    return np.maximum(np.power(np.asarray(value, np.float64) / SIGMA, 0.25), minimum)


def _quat_matrix(q):
    x, y, z, w = (float(c) for c in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def node_matrix(node):
    """Local 4x4 transform of a glTF node: ``matrix`` (column-major) or T R S."""
    if "matrix" in node:
        return np.array(node["matrix"], np.float64).reshape(4, 4).T
    m = np.eye(4)
    s = np.array(node.get("scale", [1.0, 1.0, 1.0]), np.float64)
    m[:3, :3] = _quat_matrix(node.get("rotation", [0.0, 0.0, 0.0, 1.0])) * s[None, :]
    m[:3, 3] = node.get("translation", [0.0, 0.0, 0.0])
    return m


def _accessor(g, buffers, i):
    acc = g["accessors"][i]
    view = g["bufferViews"][acc["bufferView"]]
    dt = np.dtype(_COMPONENT[acc["componentType"]])
    width = _WIDTH[acc["type"]]
    count = acc["count"]
    data = buffers[view["buffer"]]
    start = view.get("byteOffset", 0) + acc.get("byteOffset", 0)
    stride = view.get("byteStride", 0) or dt.itemsize * width
    if stride == dt.itemsize * width:
        a = np.frombuffer(data, dt, count * width, start)
    else:
        rows = np.lib.stride_tricks.as_strided(
            np.frombuffer(data, np.uint8, len(data) - start, start),
            (count, dt.itemsize * width), (stride, 1))
        a = np.ascontiguousarray(rows).view(dt).ravel()
    return a.reshape(count, width) if width > 1 else a


def _models(g):
    """(node, world matrix) of every mesh node, depth first in scene order."""
    scene = g["scenes"][g.get("scene", 0)]
    out = []

    def visit(i, parent):
        node = g["nodes"][i]
        m = parent @ node_matrix(node)
        if "mesh" in node:
            out.append((node, m))
        for c in node.get("children", []):
            visit(c, m)

    for i in scene.get("nodes", []):
        visit(i, np.eye(4))
    return out


def _prop(extras, key):
    v = extras.get(key, 0)
    return float(v) if isinstance(v, (int, float)) else 0.0


def load(path, hour=None, sky_minimum_kelvin=SKY_MINIMAL_KELVIN):
    """SceneData of a glTF scene (per-triangle materials, per-vertex initial
    and fixed temperatures).  meta: ``objects`` (name, first vertex, vertex
    count, first triangle, triangle count, properties) and, for a sky dome,
    ``sky`` (hours, values [hours x 145], patch vertex indices [145 x 6],
    kelvin per patch)."""
    with open(path) as f:
        g = json.load(f)
    base = os.path.dirname(os.path.abspath(path))
    buffers = []
    for b in g.get("buffers", []):
        with open(os.path.join(base, b["uri"]), "rb") as f:
            buffers.append(f.read())
    xs, ts, mats, kelvin, fixed, objects = [], [], [], [], [], []
    sky = None
    nv = nt = 0
    for node, m in _models(g):
        mesh = g["meshes"][node["mesh"]]
        ex = mesh.get("extras", {})
        props = {k: _prop(ex, k) for k in (
            "kelvin", "temperature-fixed", "thickness", "density", "heat-capacity",
            "heat-conductivity", "diffuse-reflectance", "specular-reflectance",
            "diffuse-emission", "traceable")}
        v0, t0 = nv, nt
        for prim in mesh["primitives"]:
            if prim.get("mode", 4) != 4:
                raise ValueError(f"{mesh.get('name')}: only triangle primitives are supported")
            p = _accessor(g, buffers, prim["attributes"]["POSITION"]).astype(np.float64)
            p = p @ m[:3, :3].T + m[:3, 3]
            if "indices" in prim:
                idx = _accessor(g, buffers, prim["indices"]).astype(np.int64)
            else:
                idx = np.arange(p.shape[0], dtype=np.int64)
            tri = idx.reshape(-1, 3) + nv
            n_t = tri.shape[0]
            mat = {"rho_d": props["diffuse-reflectance"], "rho_s": props["specular-reflectance"],
                   "thickness": props["thickness"], "density": props["density"],
                   "heat_capacity": props["heat-capacity"],
                   "conductivity": props["heat-conductivity"],
                   "flags": (MAT_DIFFUSE_EMISSION if props["diffuse-emission"] else 0) |
                            (MAT_TRACEABLE if props["traceable"] else 0)}
            mats.append((n_t, mat))
            k = np.full(p.shape[0], props["kelvin"])
            values = ex.get("sky-values") or []
            if len(values):
                sky = _sky(ex, idx, nv, hour, sky_minimum_kelvin)
                # patch by patch, a shared vertex keeps the last patch's value
                # (SkyData::setVector)
                for q in range(SKY_VALUE_COUNT):
                    k[sky["patch_vertices"][q] - nv] = sky["kelvin"][q]
            xs.append(p)
            ts.append(tri)
            kelvin.append(k)
            fixed.append(np.full(p.shape[0], 1 if props["temperature-fixed"] else 0, np.uint8))
            nv += p.shape[0]
            nt += n_t
        objects.append({"name": node.get("name", mesh.get("name", "")), "first_vertex": v0,
                        "vertex_count": nv - v0, "first_triangle": t0,
                        "triangle_count": nt - t0, "properties": props})
    xyz = np.concatenate(xs)
    idx = np.concatenate(ts)
    sd = _scenes._uniform_scene(os.path.splitext(os.path.basename(path))[0], xyz, idx)
    off = 0
    for n_t, mat in mats:
        sl = slice(off, off + n_t)
        sd.rho_d[sl] = mat["rho_d"]
        sd.rho_s[sl] = mat["rho_s"]
        sd.thickness[sl] = mat["thickness"]
        sd.density[sl] = mat["density"]
        sd.heat_capacity[sl] = mat["heat_capacity"]
        sd.conductivity[sl] = mat["conductivity"]
        sd.flags[sl] = mat["flags"]
        off += n_t
    sd.heat_flux[:] = 0.0
    sd.kelvin = np.concatenate(kelvin).astype(np.float32)
    sd.fixed = np.concatenate(fixed)
    sd.meta.update(objects=objects, sky=sky, source=os.path.basename(path))
    return sd


def _sky(ex, idx, v0, hour, minimum):
    values = np.asarray(ex["sky-values"], np.float32)
    hours = [int(h) for h in (ex.get("sky-hours") or [0])]
    if values.size != SKY_VALUE_COUNT * len(hours):
        raise ValueError(f"sky: {values.size} values for {len(hours)} hours")
    if idx.size != SKY_VALUE_COUNT * 6:
        raise ValueError(f"sky: {idx.size} indices, expected {SKY_VALUE_COUNT * 6}")
    mat = values.reshape(len(hours), SKY_VALUE_COUNT).astype(np.float64)
    if hour is None:
        # averaged: hour-mean of the W/m^2 values clipped below at sigma Tmin^4
        v = np.maximum(mat, SIGMA * minimum ** 4).mean(axis=0)
    else:
        h = float(hour) % len(hours)
        a = int(h)
        f = h - a
        v = mat[a] * (1.0 - f) + mat[(a + 1) % len(hours)] * f
    return {"hours": hours, "values": mat, "patch_vertices": idx.reshape(SKY_VALUE_COUNT, 6) + v0,
            "kelvin": sky_value_to_kelvin(v, minimum), "mean_value": float(v.mean())}
