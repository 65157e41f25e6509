"""The sun as a collimated emitter in the transport operator, and the solar
position that places it.

The reference adds the sun to a scene as one more model
(src/implementations/thermal_renderer/main.cpp:91-174, load_sun): a quad
beyond the scene's bounding sphere, facing the scene, traceable, non-diffuse
emission (photons leave along the quad's normal,
thermal_tracing_unified.glsl:153-158), fixed at the temperature whose
black-body emission is the solar flux (sky-value -> kelvin,
thermal_common.cpp:12-18).  Direct sun then enters T like every other
emitter, so buildings shade each other in the forward solve and in the
adjoint trace.

Here (``add_suns``) each sun is such a quad, with these differences:
* it spans the scene's projection onto the plane normal to the sun direction
  (a small margin around it) instead of 2.4 bounding-sphere radii per side,
  so most of its photons reach the scene;
* its two triangles are repeated ``copies`` times (coincident copies, each
  with its own photons) and the quad's temperature is lowered to
  (flux / (sigma copies))^(1/4), which keeps the emitted power at flux per
  m^2 of quad while sampling the sun with copies x as many photons as one
  scene triangle gets;
* it is not traceable: photons leaving the scene upwards escape to the
  ambient instead of being absorbed by the sun quad (a quad covering much of
  the sky would otherwise act as a black surface at the sun's temperature);
* the direction comes from solar elevation and azimuth (``solar_position``)
  in a z-up frame with x east and y north.  The reference ignores the
  azimuth, rotates (0, 0, -1) about x by the zenith angle and scales the
  solar constant by 1 / cos(zenith) (main.cpp:103-107).

Several suns can sit in one scene (one per hour of a day): each has its own
four bins, so a diurnal sweep changes only fixed temperatures, not the
operator.
"""
import datetime as _dt

import numpy as np

SIGMA = 5.670374419e-8
SOLAR_CONSTANT = 1362.0   # W/m^2 (main.cpp:107)


def solar_position(latitude_deg, longitude_deg, utc_offset_h, year, month, day, hour,
                   minute=0):
    """Solar elevation, compass azimuth (degrees from north, clockwise) and
    zenith angle in degrees at local clock time (hour:minute, utc_offset_h
    hours east of UTC; longitude east-positive).

    Declination and equation of time from Spencer's Fourier series
    (Spencer 1971, as given by Iqbal, An Introduction to Solar Radiation,
    1983, eqs. 1.3.1 and 1.4.1) with the fractional year taken at the start
    of the day; local solar time from the longitude offset to the time zone's
    meridian; elevation from the spherical-astronomy relation
    sin(el) = sin(lat) sin(dec) + cos(lat) cos(dec) cos(hour angle).  The
    reference's sunAngle (gpl/solar_position.cpp:50-90) evaluates the same
    relations with west-positive longitude and time zone, i.e.
    sunAngle(lat, -lon, -utc) = zenith here, and returns the zenith only."""
    n = _dt.date(year, month, day).timetuple().tm_yday
    g = 2.0 * np.pi * (n - 1) / 365.0
    dec = (0.006918 - 0.399912 * np.cos(g) + 0.070257 * np.sin(g) - 0.006758 * np.cos(2 * g)
           + 0.000907 * np.sin(2 * g) - 0.002697 * np.cos(3 * g) + 0.00148 * np.sin(3 * g))
    eot_min = 229.18 * (0.000075 + 0.001868 * np.cos(g) - 0.032077 * np.sin(g)
                        - 0.014615 * np.cos(2 * g) - 0.04089 * np.sin(2 * g))
    meridian = 15.0 * utc_offset_h
    solar_h = hour + minute / 60.0 + (4.0 * (longitude_deg - meridian) + eot_min) / 60.0
    w = np.radians(15.0 * (solar_h - 12.0))
    lat = np.radians(latitude_deg)
    s = np.sin(lat) * np.sin(dec) + np.cos(lat) * np.cos(dec) * np.cos(w)
    el = np.arcsin(np.clip(s, -1.0, 1.0))
    den = np.cos(el) * np.cos(lat)
    if abs(den) < 1e-12:
        az = 180.0
    else:
        ca = np.clip((np.sin(dec) - np.sin(el) * np.sin(lat)) / den, -1.0, 1.0)
        az = np.degrees(np.arccos(ca))
        if w > 0.0:          # afternoon: west of the meridian
            az = 360.0 - az
    eld = float(np.degrees(el))
    return eld, float(az), 90.0 - eld


def sun_direction(elevation_deg, azimuth_deg, compass=False):
    """Unit vector towards the sun, z up.  azimuth: counter-clockwise from +x
    (default) or, with compass=True, clockwise from north (+y) with +x east."""
    el = np.radians(elevation_deg)
    az = np.radians(90.0 - azimuth_deg if compass else azimuth_deg)
    return np.array([np.cos(el) * np.cos(az), np.cos(el) * np.sin(az), np.sin(el)])


def clear_sky_dni(zenith_deg, solar_constant=SOLAR_CONSTANT):
    """Direct normal irradiance of a clear sky, W/m^2: S0 0.7^(AM^0.678) with
    the plane-parallel air mass AM = 1 / cos(zenith) (Meinel and Meinel,
    Applied Solar Energy, 1976); 0 with the sun below the horizon."""
    c = np.cos(np.radians(zenith_deg))
    if c <= 0.0:
        return 0.0
    return float(solar_constant * 0.7 ** ((1.0 / c) ** 0.678))


def sun_kelvin(flux, copies):
    """Fixed temperature of a sun quad whose two triangles are repeated
    `copies` times, so that it emits `flux` W per m^2 of quad."""
    flux = np.asarray(flux, np.float64)
    return np.power(np.maximum(flux, 0.0) / (SIGMA * copies), 0.25)


def _basis(d):
    a = np.array([0.0, 0.0, 1.0]) if abs(d[2]) < 0.9 else np.array([1.0, 0.0, 0.0])
    u = np.cross(a, d)
    u /= np.linalg.norm(u)
    return u, np.cross(d, u)      # u x v = d


def add_suns(sd, suns, triangles_per_sun=None, margin=0.02, distance=1.2):
    """A new SceneData with one collimated sun quad per (direction, flux) in
    `suns` appended after the scene's vertices and triangles.

    Each quad lies `distance` bounding-sphere radii from the scene's
    bounding-box centre along the direction, faces the scene and covers the
    scene's projection plus `margin` of its extent; its two triangles are
    repeated triangles_per_sun / 2 times (default: about a quarter of the
    scene's triangle count, split over the suns).  Sun vertices are fixed at
    sun_kelvin(flux, copies); emissivity 1, no conduction.

    meta["suns"]: per sun, ``vertices`` (4 ids), ``triangles`` (first,
    end), ``direction``, ``flux``, ``copies``; meta["scene_vertices"] and
    meta["scene_triangles"] count the scene without the suns."""
    from .scenes import SceneData
    suns = [(np.asarray(d, np.float64) / np.linalg.norm(d), float(f)) for d, f in suns]
    if not suns:
        return sd
    nv0, nt0 = sd.nv, sd.nt
    if triangles_per_sun is None:
        triangles_per_sun = max(2, 2 * int(round(nt0 / (8 * len(suns)))))
    if triangles_per_sun < 2 or triangles_per_sun % 2:
        raise ValueError("triangles_per_sun must be even and at least 2")
    k = triangles_per_sun // 2
    X = sd.xyz.astype(np.float64)
    lo, hi = X.min(0), X.max(0)
    c = 0.5 * (lo + hi)
    r = float(np.linalg.norm(X - c, axis=1).max())
    xs, ts, info = [], [], []
    for i, (d, flux) in enumerate(suns):
        u, v = _basis(d)
        pu, pv = (X - c) @ u, (X - c) @ v
        ext = max(pu.max() - pu.min(), pv.max() - pv.min())
        m = margin * ext + 1e-3 * max(r, 1.0)
        u0, u1, v0, v1 = pu.min() - m, pu.max() + m, pv.min() - m, pv.max() + m
        o = c + distance * r * d
        q = np.array([o + u0 * u + v0 * v, o + u1 * u + v0 * v, o + u1 * u + v1 * v,
                      o + u0 * u + v1 * v])
        tri = np.array([[0, 2, 1], [0, 3, 2]])     # normal (q2-q0)x(q1-q0) = -d
        n = np.cross(q[tri[0, 1]] - q[tri[0, 0]], q[tri[0, 2]] - q[tri[0, 0]])
        assert n @ d < 0.0
        base = nv0 + 4 * i
        xs.append(q)
        ts.append(np.tile(tri + base, (k, 1)))
        t0 = nt0 + i * triangles_per_sun
        info.append({"vertices": np.arange(base, base + 4, dtype=np.uint32),
                     "triangles": (t0, t0 + triangles_per_sun), "direction": d.tolist(),
                     "flux": flux, "copies": k})
    ns = len(suns)
    nts = ns * triangles_per_sun

    def cat(a, val, n):
        a = np.asarray(a)
        return np.concatenate([a, np.full((n,) + a.shape[1:], val, a.dtype)])

    out = SceneData(sd.name, np.concatenate([sd.xyz, np.concatenate(xs).astype(np.float32)]),
                    np.concatenate([sd.idx, np.concatenate(ts).astype(np.uint32)]),
                    cat(sd.rho_d, 0.0, nts), cat(sd.rho_s, 0.0, nts), cat(sd.flags, 0, nts),
                    cat(sd.thickness, 1.0, nts), cat(sd.density, 1.0, nts),
                    cat(sd.heat_capacity, 1.0, nts), cat(sd.conductivity, 0.0, nts),
                    cat(sd.heat_flux, 0.0, nts),
                    np.concatenate([sd.kelvin, np.repeat(
                        sun_kelvin([f for _, f in suns], k), 4).astype(np.float32)]),
                    cat(sd.fixed, 1, 4 * ns), dict(sd.meta))
    meta = out.meta
    for key, val in (("ground_vertex", False), ("vertex_building", -1)):
        if key in meta:
            meta[key] = cat(meta[key], val, 4 * ns)
    if "base_xyz" in meta:
        meta["base_xyz"] = np.concatenate([np.asarray(meta["base_xyz"], np.float64),
                                           np.concatenate(xs)])
    if "alpha" in meta:
        meta["alpha"] = cat(meta["alpha"], 0.0, nts)
    meta["suns"] = list(meta.get("suns", [])) + info
    meta.setdefault("scene_vertices", nv0)
    meta.setdefault("scene_triangles", nt0)
    return out


def sun_vertex_kelvin(sd, fluxes):
    """(vertex ids, kelvin) that set the suns of `sd` (add_suns) to the given
    fluxes (W/m^2 of quad, one per sun): for a state update between time
    steps or designs."""
    suns = sd.meta["suns"]
    if len(fluxes) != len(suns):
        raise ValueError("one flux per sun")
    v = np.concatenate([s["vertices"] for s in suns])
    k = np.repeat([sun_kelvin(f, s["copies"]) for s, f in zip(suns, fluxes)], 4)
    return v, k.astype(np.float32)
