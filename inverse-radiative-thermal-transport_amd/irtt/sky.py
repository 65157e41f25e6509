"""The sky as 145 Tregenza patches: weather-file preparation, sky-matrix
parsing and the dome of fixed-temperature emitters.

Restates the reference's own sky code
(src/implementations/thermal_renderer/sky/skymatrix.hpp:30-409,
thermal_sky.cpp:8-144, thermal_data.hpp:19-140, thermal_common.cpp:12-18):

* ``datetime_to_hoy``: "YYYY-MM-DDTHH:MM:SS+hh:mm" -> hour of the year
  (skymatrix.hpp:69-93; the zone offset is read and not applied, the hour is
  truncated to an integer);
* ``filter_wea``: the WEA file's 6 header lines plus the data lines of hours
  [begin, end) (skymatrix.hpp:129-160);
* ``parse_sky_mtx``: a one-hour sky matrix -> 145 broadband values, patches
  in the matrix's order (row 0 = the 30 horizon patches ... row 7 = zenith),
  the ground patch and the header (9 lines) skipped, each value
  0.01 (0.265074126 R + 0.670114631 G + 0.064811243 B) W/m^2
  (skymatrix.hpp:96-106,165-188);
* ``tregenza_dome``: 145 quads on a sphere of 2.2 bounding-sphere radii,
  zenith patch first, then rings of 6, 12, 18, 24, 24, 30, 30 patches at
  12 degree steps of inclination, each quad spanning 2.2 radii from its
  centre toward its four corners, two triangles whose normal faces the
  centre (thermal_sky.cpp:10-90);
* ``add_sky``: the dome appended to a scene as fixed-temperature, non-diffuse
  (collimated along the quad normal), non-traceable emitters, value i on quad
  i (SkyData::setVector, thermal_data.hpp:79-99) at
  kelvin = max((v / sigma)^(1/4), 111.575 K) (sky_value_to_kelvin,
  thermal_common.cpp:12-18; SkyData::MINIMAL_KELVIN).

Producing the sky matrix itself from a weather file is Radiance's
``epw2wea`` / ``gendaymtx`` (third-party programs the reference runs with
``system()``); they are not part of this package: ``parse_sky_mtx`` reads
their output.
"""
import datetime as _dt

import numpy as np

SIGMA = 5.670374419e-8
NUM_SKY_VALUES = 145
PATCHES_PER_ROW = (30, 30, 24, 24, 18, 12, 6, 1)   # horizon row first
DEGREES_PER_ROW = 12.0
MINIMAL_KELVIN = 111.575
BROADBAND = (0.265074126, 0.670114631, 0.064811243)


def datetime_to_hoy(text):
    """Hour of the year of "YYYY-MM-DDTHH:MM:SS+hh:mm" (zone not applied)."""
    date, _, rest = text.partition("T")
    y, m, d = (int(x) for x in date.split("-"))
    clock = rest.split("+")[0].split("-")[0]
    hh, mm, ss = (int(x) for x in clock.split(":"))
    t = _dt.datetime(y, m, d) + _dt.timedelta(hours=hh, minutes=mm, seconds=ss)
    return int((t - _dt.datetime(y, 1, 1)).total_seconds() // 3600)


def filter_wea(lines, begin_hoy, end_hoy):
    """The 6 header lines and the data lines of hours [begin_hoy, end_hoy)."""
    out = []
    for k, line in enumerate(lines):
        if k < 6 or begin_hoy <= k - 6 < end_hoy:
            out.append(line)
    return out


def broadband(r, g, b):
    """W/m^2 of one patch from the sky matrix's three channels."""
    return 0.01 * (BROADBAND[0] * r + BROADBAND[1] * g + BROADBAND[2] * b)


def parse_sky_mtx(text):
    """145 values (matrix order: horizon row first, zenith last) of a
    one-hour sky matrix; missing lines leave zeros."""
    lines = text.splitlines()[9:]
    out = np.zeros(NUM_SKY_VALUES, np.float32)
    k = 0
    for row, n in enumerate(PATCHES_PER_ROW):
        for _ in range(n):
            if k >= len(lines):
                return out
            r, g, b = (float(x) for x in lines[k].split()[:3])
            out[k] = np.float32(broadband(r, g, b))
            k += 1
    return out


def sky_value_to_kelvin(value, minimum=MINIMAL_KELVIN):
    v = np.maximum(np.asarray(value, np.float64), 0.0)
    return np.maximum(np.power(v / SIGMA, 0.25), minimum)


def _ring_order():
    """(patches, inclination in degrees) of each ring, zenith first."""
    n = len(PATCHES_PER_ROW)
    return [(PATCHES_PER_ROW[n - 1 - ci], ci * DEGREES_PER_ROW) for ci in range(n)]


def tregenza_dome(center, radius, up="z"):
    """(xyz (580 x 3), triangles (290 x 3), patch centre directions
    (145 x 3)) of the dome around `center` for a bounding-sphere `radius`."""
    center = np.asarray(center, np.float64)
    R = 2.2 * radius
    xyz, tri, dirs = [], [], []
    for c, incl in _ring_order():
        for i in range(c):
            az = np.radians(i * 360.0 / c)
            th = np.radians(incl)
            # the reference's sph2cart_Yup: (r sin i cos a, r cos i, r sin i sin a)
            ny = np.array([np.sin(th) * np.cos(az), np.cos(th), np.sin(th) * np.sin(az)])
            up_y = np.array([0.0, 1.0, 0.0])
            t = np.cross(ny, up_y)
            if c == 1:
                t = np.cross(ny, np.array([1.0, 0.0, 0.0]))
            t /= np.linalg.norm(t)
            b = np.cross(ny, t)
            b /= np.linalg.norm(b)
            corners = [t + b, t - b, -t - b, -t + b]
            q = [R * ny + R * k / np.linalg.norm(k) for k in corners]
            if up == "z":   # (x, y, z)_yup -> (x, -z, y)_zup keeps the handedness
                q = [np.array([p[0], -p[2], p[1]]) for p in q]
                ny = np.array([ny[0], -ny[2], ny[1]])
            base = len(xyz)
            xyz.extend(center + p for p in q)
            tri.append([base, base + 1, base + 2])
            tri.append([base, base + 2, base + 3])
            dirs.append(ny)
    return (np.array(xyz, np.float64), np.array(tri, np.int64), np.array(dirs, np.float64))


def add_sky(sd, values, up="z", minimum_kelvin=MINIMAL_KELVIN):
    """A new SceneData with the Tregenza dome appended: 145 fixed-temperature,
    collimated, non-traceable emitter quads around the scene's bounding sphere,
    quad i at sky_value_to_kelvin(values[i]).  meta["sky"]: ``vertices``
    (580 ids, 4 per patch), ``triangles`` (first, end), ``directions``."""
    from .scenes import SceneData
    values = np.asarray(values, np.float64)
    if values.shape != (NUM_SKY_VALUES,):
        raise ValueError("145 sky values, one per Tregenza patch")
    x = sd.xyz.astype(np.float64)
    lo, hi = x.min(0), x.max(0)
    center = 0.5 * (lo + hi)
    radius = 0.5 * float(np.linalg.norm(hi - lo))
    q, t, dirs = tregenza_dome(center, radius, up)
    nv0, nt0 = sd.nv, sd.nt
    nts = t.shape[0]

    def cat(a, val, n):
        a = np.asarray(a)
        return np.concatenate([a, np.full((n,) + a.shape[1:], val, a.dtype)])

    kel = np.repeat(sky_value_to_kelvin(values, minimum_kelvin), 4).astype(np.float32)
    out = SceneData(sd.name, np.concatenate([sd.xyz, q.astype(np.float32)]),
                    np.concatenate([sd.idx, (t + nv0).astype(np.uint32)]),
                    cat(sd.rho_d, 0.0, nts), cat(sd.rho_s, 0.0, nts), cat(sd.flags, 0, nts),
                    cat(sd.thickness, 1.0, nts), cat(sd.density, 1.0, nts),
                    cat(sd.heat_capacity, 1.0, nts), cat(sd.conductivity, 0.0, nts),
                    cat(sd.heat_flux, 0.0, nts), np.concatenate([sd.kelvin, kel]),
                    cat(sd.fixed, 1, q.shape[0]), dict(sd.meta))
    meta = out.meta
    for key, val in (("ground_vertex", False), ("vertex_building", -1)):
        if key in meta:
            meta[key] = cat(meta[key], val, q.shape[0])
    if "base_xyz" in meta:
        meta["base_xyz"] = np.concatenate([np.asarray(meta["base_xyz"], np.float64), q])
    if "alpha" in meta:
        meta["alpha"] = cat(meta["alpha"], 0.0, nts)
    meta["sky"] = {"vertices": np.arange(nv0, nv0 + q.shape[0], dtype=np.uint32),
                   "triangles": (nt0, nt0 + nts), "directions": dirs}
    meta.setdefault("scene_vertices", nv0)
    meta.setdefault("scene_triangles", nt0)
    return out


def sky_vertex_kelvin(sd, values, minimum_kelvin=MINIMAL_KELVIN):
    """(vertex ids, kelvin) that set the dome of `sd` (add_sky) to another
    hour's 145 values: a state update between time steps."""
    v = sd.meta["sky"]["vertices"]
    k = np.repeat(sky_value_to_kelvin(np.asarray(values, np.float64), minimum_kelvin), 4)
    return v, k.astype(np.float32)
