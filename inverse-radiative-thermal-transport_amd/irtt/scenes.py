"""Synthetic scene generators (numpy, seeded).

* ``parallel_plates`` -- two coaxial 1 m x 1 m plates 1 m apart, each a shared
  17 x 17 vertex grid of 16 x 16 quads (512 triangles), facing each other.
* ``city_block`` -- 8 x 8 box buildings (footprints 10-20 m, heights 6-40 m,
  >= 5 m streets) on a ground plane, every surface subdivided to roughly equal
  triangles (about 50k in total by default); per-patch emissivity, solar
  absorptivity and areal heat capacity; direct sun at a given elevation.
* ``closed_box`` -- the inside of a closed unit cube (every photon is absorbed
  somewhere), used for conservation checks.
* ``blocker_plates`` -- an emitter plate, an opaque strip and a receiver
  plate: a straight shadow edge, for checking the edge-sampling adjoint.
* ``fin_plates`` -- the same with a vertical fin that the emitter's rays
  reach at grazing incidence.
* ``balcony_wall`` -- a wall with one balcony under a collimated sun: the
  balcony's shadow on the wall, for the adjoint of collimated emitters.

Each returns a ``SceneData`` with per-triangle materials and per-vertex
initial / fixed temperatures, the per-object inputs of the reference's
``ObjectProperties`` (src/implementations/thermal_renderer/thermal_api.cpp:25-45)
flattened to triangles.
"""
from dataclasses import dataclass, field

import numpy as np

from . import MATERIAL_DTYPE, MAT_DIFFUSE_EMISSION, MAT_TRACEABLE


@dataclass
class SceneData:
    name: str
    xyz: np.ndarray          # (nv, 3) float32
    idx: np.ndarray          # (nt, 3) uint32
    rho_d: np.ndarray        # (nt,) float32
    rho_s: np.ndarray        # (nt,) float32
    flags: np.ndarray        # (nt,) uint8
    thickness: np.ndarray    # (nt,)
    density: np.ndarray
    heat_capacity: np.ndarray
    conductivity: np.ndarray
    heat_flux: np.ndarray    # absorbed external flux, W/m^2
    kelvin: np.ndarray       # (nv,) initial / fixed temperature
    fixed: np.ndarray        # (nv,) uint8
    meta: dict = field(default_factory=dict)

    @property
    def nv(self):
        return self.xyz.shape[0]

    @property
    def nt(self):
        return self.idx.shape[0]

    def materials(self):
        m = np.zeros(self.nt, MATERIAL_DTYPE)
        m["diffuse_reflectance"] = self.rho_d
        m["specular_reflectance"] = self.rho_s
        m["thickness"] = self.thickness
        m["density"] = self.density
        m["heat_capacity"] = self.heat_capacity
        m["heat_conductivity"] = self.conductivity
        m["heat_flux"] = self.heat_flux
        m["flags"] = self.flags
        return m

    def triangle_areas(self):
        p = self.xyz.astype(np.float64)[self.idx]
        c = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
        return 0.5 * np.linalg.norm(c, axis=1)

    def normals(self):
        p = self.xyz.astype(np.float64)[self.idx]
        c = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
        return c / np.linalg.norm(c, axis=1)[:, None]


def _uniform_scene(name, xyz, idx, rho_d=0.0, rho_s=0.0, flags=MAT_DIFFUSE_EMISSION | MAT_TRACEABLE,
                   thickness=1.0, density=1.0, heat_capacity=1.0, conductivity=0.0,
                   heat_flux=0.0, kelvin=0.0):
    nt = idx.shape[0]
    nv = xyz.shape[0]

    def f(v, dt=np.float32, n=nt):
        # always a writable copy (a broadcast view of a full-size array is read-only)
        return np.array(np.broadcast_to(np.asarray(v, dt), (n,)), dtype=dt)

    return SceneData(name, np.ascontiguousarray(xyz, np.float32),
                     np.ascontiguousarray(idx, np.uint32), f(rho_d), f(rho_s), f(flags, np.uint8),
                     f(thickness), f(density), f(heat_capacity), f(conductivity), f(heat_flux),
                     f(kelvin, np.float32, nv), np.zeros(nv, np.uint8))


def grid_patch(origin, u, v, nu, nv_):
    """Vertices and triangles of a planar patch origin + a*u + b*v, a,b in
    [0,1], subdivided nu x nv_; normal = u x v."""
    origin, u, v = (np.asarray(a, np.float64) for a in (origin, u, v))
    a = np.linspace(0.0, 1.0, nu + 1)
    b = np.linspace(0.0, 1.0, nv_ + 1)
    A, B = np.meshgrid(a, b, indexing="ij")
    P = origin + A[..., None] * u + B[..., None] * v
    xyz = P.reshape(-1, 3)
    i = np.arange(nu)[:, None]
    j = np.arange(nv_)[None, :]
    v00 = (i * (nv_ + 1) + j).ravel()
    v10 = ((i + 1) * (nv_ + 1) + j).ravel()
    v11 = ((i + 1) * (nv_ + 1) + j + 1).ravel()
    v01 = (i * (nv_ + 1) + j + 1).ravel()
    tri = np.concatenate([np.stack([v00, v10, v11], 1), np.stack([v00, v11, v01], 1)])
    # interleave so the two triangles of a quad are adjacent (quad q -> 2q, 2q+1)
    nq = v00.size
    order = np.empty(2 * nq, np.int64)
    order[0::2] = np.arange(nq)
    order[1::2] = np.arange(nq) + nq
    return xyz, tri[order]


def parallel_plates(n=16, gap=1.0, emissivity=0.9, heat_flux_a=500.0, kelvin=0.0):
    """Plate A at z=0 facing +z, plate B at z=gap facing -z, both [0,1]^2."""
    xa, ta = grid_patch([0, 0, 0], [1, 0, 0], [0, 1, 0], n, n)
    xb, tb = grid_patch([0, 0, gap], [0, 1, 0], [1, 0, 0], n, n)  # swapped axes: normal -z
    xyz = np.concatenate([xa, xb])
    idx = np.concatenate([ta, tb + xa.shape[0]])
    nt = idx.shape[0]
    flux = np.zeros(nt, np.float32)
    flux[: ta.shape[0]] = heat_flux_a
    sd = _uniform_scene("parallel_plates", xyz, idx, rho_d=1.0 - emissivity, heat_flux=flux,
                        kelvin=kelvin)
    sd.meta.update(n=n, gap=gap, emissivity=emissivity, plate_a_tris=ta.shape[0],
                   plate_a_verts=xa.shape[0])
    return sd


def blocker_plates(n=16, gap=1.0, blocker_z=0.5, blocker_x=(0.3, 0.6), blocker_y=(-0.5, 1.5),
                   blocker_cells=(6, 20), receiver=True, emissivity=1.0):
    """Emitter plate [0,1]^2 at z=0 facing +z, an opaque strip at z=blocker_z
    spanning blocker_x x blocker_y facing the emitter (-z), and (receiver)
    plate [0,1]^2 at z=gap facing -z.  The strip's edge at x = blocker_x[1]
    casts a straight shadow across the receiver: moving that edge is a pure
    visibility change, the case the adjoint's edge term is for.
    meta: ``emitter_verts``, ``blocker_verts`` (vertex ranges), ``edge_vertex``
    (the strip's vertices on x = blocker_x[1])."""
    xe, te = grid_patch([0, 0, 0], [1, 0, 0], [0, 1, 0], n, n)
    x0, x1 = blocker_x
    y0, y1 = blocker_y
    # normal = u x v = (0, y, 0) x (x, 0, 0) = -z
    xs_, ts_ = grid_patch([x0, y0, blocker_z], [0, y1 - y0, 0], [x1 - x0, 0, 0],
                          blocker_cells[1], blocker_cells[0])
    parts = [(xe, te), (xs_, ts_)]
    if receiver:
        parts.append(grid_patch([0, 0, gap], [0, 1, 0], [1, 0, 0], n, n))
    xyz, idx, off = [], [], 0
    for x, t in parts:
        xyz.append(x)
        idx.append(t + off)
        off += x.shape[0]
    xyz = np.concatenate(xyz)
    sd = _uniform_scene("blocker_plates", xyz, np.concatenate(idx), rho_d=1.0 - emissivity)
    nve, nvb = xe.shape[0], xs_.shape[0]
    edge = nve + np.nonzero(np.isclose(xs_[:, 0], x1))[0]
    sd.meta.update(n=n, emitter_verts=(0, nve), blocker_verts=(nve, nve + nvb),
                   emitter_tris=te.shape[0], edge_vertex=edge.astype(np.uint32),
                   blocker_x=blocker_x, receiver=receiver)
    return sd


def fin_plates(n=16, fin_x=0.5, fin_z=(0.05, 0.3), fin_y=(-0.5, 1.5), fin_cells=(4, 20),
               emitter_x=(0.0, 0.5), gap=1.0, emissivity=1.0):
    """Emitter plate emitter_x x [0,1] at z=0 facing +z, a vertical opaque fin
    in the plane x = fin_x spanning fin_y x fin_z facing -x (towards the
    emitter), and a receiver plate [0,1]^2 at z=gap facing -z.  Rays reach
    the fin's top edge at grazing incidence: the case where the edge term's
    projection factor matters.  meta: ``emitter_verts``, ``fin_verts``,
    ``edge_vertex`` (the fin's vertices on z = fin_z[1])."""
    e0, e1 = emitter_x
    ne = max(1, int(round(n * (e1 - e0))))
    xe, te = grid_patch([e0, 0, 0], [e1 - e0, 0, 0], [0, 1, 0], ne, n)
    z0, z1 = fin_z
    y0, y1 = fin_y
    # normal = u x v = (0, 0, z) x (0, y, 0) = -x
    xf, tf = grid_patch([fin_x, y0, z0], [0, 0, z1 - z0], [0, y1 - y0, 0], fin_cells[0],
                        fin_cells[1])
    xr, tr = grid_patch([0, 0, gap], [0, 1, 0], [1, 0, 0], n, n)
    xyz = np.concatenate([xe, xf, xr])
    idx = np.concatenate([te, tf + xe.shape[0], tr + xe.shape[0] + xf.shape[0]])
    sd = _uniform_scene("fin_plates", xyz, idx, rho_d=1.0 - emissivity)
    nve, nvf = xe.shape[0], xf.shape[0]
    edge = nve + np.nonzero(np.isclose(xf[:, 2], z1))[0]
    sd.meta.update(n=n, emitter_verts=(0, nve), fin_verts=(nve, nve + nvf),
                   edge_vertex=edge.astype(np.uint32), fin_z=fin_z)
    return sd


def balcony_wall(depth=1.0, sun_elevation_deg=45.0, sun_azimuth_deg=0.0, sun_flux=800.0,
                 sun_triangles=1024, n=16):
    """A sunlit wall with one balcony: ground [0,4] x [-2,2] at z=0 facing
    +z, a wall in the plane x=0 facing +x (4 m wide, 4 m high), and a
    balcony slab (z 2-2.3 m, y -1..1 m, x 0..depth) as an open box of top,
    bottom, front and side faces, under a collimated sun quad (azimuth
    counter-clockwise from +x, so 0 faces the wall).  Emissivity 1.  Growing
    the depth moves the balcony's shadow down the wall: the balcony gains what
    the wall loses.  meta: ``owner`` per vertex (0 ground, 1 wall, 2
    balcony, -1 sun), ``depth_dir`` (d x / d depth per vertex)."""
    from .sun import add_suns, sun_direction
    z0, z1, y0, y1 = 2.0, 2.3, -1.0, 1.0
    ex, ez = np.array([1.0, 0, 0]), np.array([0, 0, 1.0])
    parts = [(grid_patch([0, -2, 0], [4, 0, 0], [0, 4, 0], n, n), 0),
             (_oriented_patch([0, -2, 0], [0, 4, 0], [0, 0, 4], n, n, ex), 1),
             (_oriented_patch([0, y0, z1], [depth, 0, 0], [0, y1 - y0, 0], 4, 8, ez), 2),
             (_oriented_patch([0, y0, z0], [depth, 0, 0], [0, y1 - y0, 0], 4, 8, -ez), 2),
             (_oriented_patch([depth, y0, z0], [0, y1 - y0, 0], [0, 0, z1 - z0], 8, 2, ex), 2),
             (_oriented_patch([0, y0, z0], [depth, 0, 0], [0, 0, z1 - z0], 4, 2,
                              np.array([0, -1.0, 0])), 2),
             (_oriented_patch([0, y1, z0], [depth, 0, 0], [0, 0, z1 - z0], 4, 2,
                              np.array([0, 1.0, 0])), 2)]
    xyz, idx, owner, off = [], [], [], 0
    for (x, t), o in parts:
        xyz.append(x)
        idx.append(t + off)
        owner.append(np.full(x.shape[0], o, np.int32))
        off += x.shape[0]
    xyz = np.concatenate(xyz)
    owner = np.concatenate(owner)
    sd = _uniform_scene("balcony_wall", xyz, np.concatenate(idx))
    ddir = np.zeros((xyz.shape[0], 3), np.float32)
    bal = owner == 2
    ddir[bal, 0] = xyz[bal, 0] / depth
    sd.meta.update(owner=owner, depth_dir=ddir, depth=depth)
    sd = add_suns(sd, [(sun_direction(sun_elevation_deg, sun_azimuth_deg), sun_flux)],
                  triangles_per_sun=sun_triangles)
    sd.meta["owner"] = np.concatenate([owner, np.full(4, -1, np.int32)])
    sd.meta["depth_dir"] = np.concatenate([ddir, np.zeros((4, 3), np.float32)])
    return sd


def closed_box(n=4, emissivity=1.0):
    """Inside of the unit cube, every face subdivided n x n, normals inward."""
    faces = [  # normal = u x v
        ([0, 0, 0], [1, 0, 0], [0, 1, 0]),  # z=0, normal +z
        ([0, 0, 1], [0, 1, 0], [1, 0, 0]),  # z=1, normal -z
        ([0, 0, 0], [0, 0, 1], [1, 0, 0]),  # y=0, normal +y
        ([0, 1, 0], [1, 0, 0], [0, 0, 1]),  # y=1, normal -y
        ([0, 0, 0], [0, 1, 0], [0, 0, 1]),  # x=0, normal +x
        ([1, 0, 0], [0, 0, 1], [0, 1, 0]),  # x=1, normal -x
    ]
    xs, ts, off = [], [], 0
    for o, u, v in faces:
        x, t = grid_patch(o, u, v, n, n)
        xs.append(x)
        ts.append(t + off)
        off += x.shape[0]
    return _uniform_scene("closed_box", np.concatenate(xs), np.concatenate(ts),
                          rho_d=1.0 - emissivity)


def city_block(seed=0, nx=8, ny=8, target_tris=50000, street=5.0, fmin=10.0, fmax=20.0,
               hmin=6.0, hmax=40.0, ground_cell=2.5, sun_elevation_deg=45.0,
               sun_azimuth_deg=135.0, sun_flux=800.0, eps_range=(0.85, 0.95),
               alpha_range=(0.3, 0.8), capacity_range=(1e4, 1e5), kelvin=293.0,
               sun_model="flux", sun_triangles=None):
    """Synthetic city block; returns SceneData with meta on the layout.

    Footprints are multiples of ``ground_cell`` so walls meet the ground grid
    exactly.  Patch = one quad (two triangles).  Direct sun (sun_model):
    "emitter" -- a collimated sun quad traced through the scene (sun.add_suns,
    sun_triangles repeated triangles; buildings shade each other and the
    ground); "flux" -- the unshadowed absorbed flux alpha * flux * max(0, n.s)
    per patch as an external source; None -- no sun."""
    rng = np.random.default_rng(seed)
    pitch = fmax + street
    gcells_per_pitch = int(round(pitch / ground_cell))
    ncx, ncy = nx * gcells_per_pitch, ny * gcells_per_pitch
    occupied = np.zeros((ncx, ncy), bool)
    bld = []
    for bi in range(nx):
        for bj in range(ny):
            wc = int(round(rng.uniform(fmin, fmax) / ground_cell))
            dc = int(round(rng.uniform(fmin, fmax) / ground_cell))
            h = float(rng.uniform(hmin, hmax))
            ox = bi * gcells_per_pitch + (gcells_per_pitch - wc) // 2
            oy = bj * gcells_per_pitch + (gcells_per_pitch - dc) // 2
            occupied[ox:ox + wc, oy:oy + dc] = True
            bld.append((ox, oy, wc, dc, h))
    # total surface to size the wall/roof subdivision
    ground_area = (~occupied).sum() * ground_cell ** 2
    wall_area = sum(2 * (wc + dc) * ground_cell * h for _, _, wc, dc, h in bld)
    roof_area = sum(wc * dc * ground_cell ** 2 for _, _, wc, dc, _ in bld)
    ground_tris = 2 * int((~occupied).sum())
    rest = max(target_tris - ground_tris, 1000)
    edge = np.sqrt(2.0 * (wall_area + roof_area) / rest)
    xs, ts, off = [], [], 0
    owner = []   # building index per vertex, -1 on the ground
    current = [-1]

    def add(o, u, v, nu, nv_):
        nonlocal off
        x, t = grid_patch(o, u, v, nu, nv_)
        owner.append(np.full(x.shape[0], current[0], np.int32))
        xs.append(x)
        ts.append(t + off)
        off += x.shape[0]

    # ground: one patch per maximal run of free cells along x in each row,
    # subdivided per cell so vertex spacing matches the walls' feet
    g = ground_cell
    for cy in range(ncy):
        cx = 0
        while cx < ncx:
            if occupied[cx, cy]:
                cx += 1
                continue
            c0 = cx
            while cx < ncx and not occupied[cx, cy]:
                cx += 1
            add([c0 * g, cy * g, 0.0], [(cx - c0) * g, 0, 0], [0, g, 0], cx - c0, 1)
    for b, (ox, oy, wc, dc, h) in enumerate(bld):
        current[0] = b
        x0, y0, x1, y1 = ox * g, oy * g, (ox + wc) * g, (oy + dc) * g
        nh = max(1, int(round(h / edge)))
        nwx = max(1, int(round((x1 - x0) / edge)))
        nwy = max(1, int(round((y1 - y0) / edge)))
        # walls, outward normals (u x v points out)
        add([x0, y0, 0], [x1 - x0, 0, 0], [0, 0, h], nwx, nh)   # y=y0, normal -y
        add([x1, y1, 0], [x0 - x1, 0, 0], [0, 0, h], nwx, nh)   # y=y1, normal +y
        add([x1, y0, 0], [0, y1 - y0, 0], [0, 0, h], nwy, nh)   # x=x1, normal +x
        add([x0, y1, 0], [0, y0 - y1, 0], [0, 0, h], nwy, nh)   # x=x0, normal -x
        add([x0, y0, h], [x1 - x0, 0, 0], [0, y1 - y0, 0], nwx, nwy)  # roof, +z
    xyz = np.concatenate(xs)
    idx = np.concatenate(ts)
    nt = idx.shape[0]
    nq = nt // 2
    eps = np.repeat(rng.uniform(*eps_range, nq), 2).astype(np.float32)
    alpha = np.repeat(rng.uniform(*alpha_range, nq), 2)
    cap = np.repeat(rng.uniform(*capacity_range, nq), 2).astype(np.float32)
    el, az = np.radians(sun_elevation_deg), np.radians(sun_azimuth_deg)
    sun = np.array([np.cos(el) * np.cos(az), np.cos(el) * np.sin(az), np.sin(el)])
    sd = _uniform_scene("city_block", xyz, idx, rho_d=1.0 - eps, heat_capacity=cap,
                        kelvin=kelvin)
    if sun_model == "flux":
        cosi = np.clip(sd.normals() @ sun, 0.0, None)
        sd.heat_flux = (alpha * sun_flux * cosi).astype(np.float32)
    areas = sd.triangle_areas()
    sd.meta.update(seed=seed, buildings=len(bld), edge=float(edge), sun=sun.tolist(),
                   mean_triangle_area=float(areas.mean()), total_area=float(areas.sum()),
                   alpha=alpha.astype(np.float32), vertex_building=np.concatenate(owner),
                   heights=np.array([b[4] for b in bld]), sun_model=sun_model)
    return _with_sun(sd, sun_model, sun, sun_flux, sun_triangles)


def _with_sun(sd, sun_model, direction, flux, triangles):
    if sun_model == "emitter":
        from .sun import add_suns
        return add_suns(sd, [(direction, flux)], triangles_per_sun=triangles)
    if sun_model not in ("flux", None):
        raise ValueError(f"unknown sun model {sun_model!r}")
    return sd


def _oriented_patch(origin, u, v, nu, nv_, normal):
    """grid_patch whose normal (u x v) points along `normal` (u, v swapped if not)."""
    u, v = np.asarray(u, np.float64), np.asarray(v, np.float64)
    if np.dot(np.cross(u, v), normal) < 0:
        return grid_patch(origin, v, u, nv_, nu)
    return grid_patch(origin, u, v, nu, nv_)


def facade(seed=0, nx=5, ny=4, target_tris=400_000, street=12.0, width=24.0, depth=16.0,
           hmin=20.0, hmax=45.0, floor=3.0, bay=4.0, slab=0.3, fin_width=0.3, n_fin_buildings=16,
           depth_range=(0.3, 1.2), balcony_cells=6, depth_cells=2, fin_cell=1.0,
           eps_range=(0.85, 0.95), alpha_range=(0.3, 0.8), capacity_range=(1e4, 1e5),
           sun_elevation_deg=45.0, sun_azimuth_deg=135.0, sun_flux=800.0, kelvin=293.0,
           weld=False, sun_model="flux", sun_triangles=None):
    """Procedural facade scene: nx x ny box buildings (footprint width x
    depth, heights U[hmin, hmax]) on a ground plane of 1 m cells; every
    facade carries balconies (slabs of thickness `slab`, one per bay and
    floor, 1 m narrower than the bay) and the -y facades of the first
    `n_fin_buildings` buildings carry vertical fins at the bay boundaries.
    Each extrusion is an open box (front, sides, and top/bottom) standing on
    the wall; its depth along the wall normal is a design parameter shared by
    a group: the balconies of one height third of one facade (3 x 4 x 20 =
    240 groups) and the fins of one building (16) = 256 parameters by
    default, initial depths U[depth_range].  Element subdivision is fixed
    (balcony_cells along a balcony, depth_cells along the depth, fins in
    fin_cell rows), so the mesh topology does not depend on the depths; walls
    and roofs are subdivided to reach about target_tris triangles.  weld:
    merge coincident vertices (faces meeting at an edge share its vertices;
    otherwise every planar patch has its own).  sun_model, sun_triangles: as
    for city_block (with "emitter" the balconies and fins shade the walls
    and the ground in the forward solve and in the adjoint trace).

    meta: ``param_vertex``, ``param_index``, ``param_dir`` (d x / d depth,
    per element vertex), ``depths`` (initial), ``base_xyz`` (positions at
    depth 0), ``ground_vertex`` (mask), ``n_params``."""
    rng = np.random.default_rng(seed)
    nb_ = nx * ny
    heights = rng.uniform(hmin, hmax, nb_)
    px, py = width + street, depth + street
    ncx, ncy = int(round(nx * px)), int(round(ny * py))
    occupied = np.zeros((ncx, ncy), bool)
    bld = []
    for b in range(nb_):
        i, j = b % nx, b // nx
        x0, y0 = i * px + street / 2.0, j * py + street / 2.0
        occupied[int(round(x0)):int(round(x0 + width)), int(round(y0)):int(round(y0 + depth))] = True
        bld.append((x0, y0, float(heights[b])))
    xs, ts, off = [], [], 0
    ground, pv, pi_, pdir = [], [], [], []

    def add(x, t, is_ground=False, param=None, n=None, d0=None, wall_pt=None):
        nonlocal off
        xs.append(x)
        ts.append(t + off)
        ground.append(np.full(x.shape[0], is_ground))
        if param is not None:
            frac = np.clip(((x - wall_pt) @ n) / d0, 0.0, 1.0)
            pv.append(np.arange(off, off + x.shape[0]))
            pi_.append(np.full(x.shape[0], param))
            pdir.append(frac[:, None] * n[None, :])
        off += x.shape[0]

    # ground: maximal runs of free cells along x in each row of 1 m cells
    for cy in range(ncy):
        cx = 0
        while cx < ncx:
            if occupied[cx, cy]:
                cx += 1
                continue
            c0 = cx
            while cx < ncx and not occupied[cx, cy]:
                cx += 1
            add(*grid_patch([c0, cy, 0.0], [cx - c0, 0, 0], [0, 1.0, 0], cx - c0, 1), True)

    # elements first (their count fixes the wall subdivision); parameters:
    # 12 balcony groups per building, then one fin group per fin building
    n_bal = 12 * nb_
    depths = rng.uniform(*depth_range, n_bal + min(n_fin_buildings, nb_))
    ez = np.array([0.0, 0.0, 1.0])

    def extrusion(corner, a_vec, b_vec, n, d, na, nbb, param, bottom=True):
        # open box on the wall rectangle corner + [0,1] a_vec + [0,1] b_vec,
        # extruded by d along n
        corner = np.asarray(corner, np.float64)
        a_hat = a_vec / np.linalg.norm(a_vec)
        b_hat = b_vec / np.linalg.norm(b_vec)
        dn = d * n
        kw = dict(param=param, n=n, d0=d, wall_pt=corner)
        add(*_oriented_patch(corner + dn, a_vec, b_vec, na, nbb, n), **kw)              # front
        add(*_oriented_patch(corner + b_vec, a_vec, dn, na, depth_cells, b_hat), **kw)  # b = 1
        if bottom:
            add(*_oriented_patch(corner, a_vec, dn, na, depth_cells, -b_hat), **kw)     # b = 0
        add(*_oriented_patch(corner, b_vec, dn, nbb, depth_cells, -a_hat), **kw)        # a = 0
        add(*_oriented_patch(corner + a_vec, b_vec, dn, nbb, depth_cells, a_hat), **kw)  # a = 1

    walls = []
    for b, (x0, y0, h) in enumerate(bld):
        x1, y1 = x0 + width, y0 + depth
        # (start, tangent, length, outward normal) of the 4 facades; q = 0 is the -y facade
        fac = [(np.array([x0, y0, 0.0]), np.array([1.0, 0, 0]), width, np.array([0, -1.0, 0])),
               (np.array([x1, y0, 0.0]), np.array([0, 1.0, 0]), depth, np.array([1.0, 0, 0])),
               (np.array([x1, y1, 0.0]), np.array([-1.0, 0, 0]), width, np.array([0, 1.0, 0])),
               (np.array([x0, y1, 0.0]), np.array([0, -1.0, 0]), depth, np.array([-1.0, 0, 0]))]
        walls.append((b, h, fac))
        nf = int((h - 1.0) // floor)
        for q, (p0, t, L, n) in enumerate(fac):
            nbay = max(1, int(L // bay))
            bw = L / nbay
            for k in range(nbay):
                s0 = k * bw + 0.5
                for f in range(1, nf + 1):
                    band = min(2, 3 * (f - 1) // max(nf, 1))
                    param = (b * 4 + q) * 3 + band
                    corner = p0 + s0 * t + (f * floor - slab) * ez
                    extrusion(corner, (bw - 1.0) * t, slab * ez, n, depths[param],
                              balcony_cells, 1, param)
            if q == 0 and b < n_fin_buildings:
                param = n_bal + b
                nh = max(1, int(round(h / fin_cell)))
                for k in range(1, nbay):
                    corner = p0 + (k * bw - fin_width / 2.0) * t
                    extrusion(corner, fin_width * t, h * ez, n, depths[param], 1, nh, param,
                              bottom=False)
    elem_tris = sum(t.shape[0] for t in ts) - sum(
        t.shape[0] for t, g in zip(ts, ground) if g.all())
    ground_tris = sum(t.shape[0] for t, g in zip(ts, ground) if g.all())
    wall_area = sum(2 * (width + depth) * h for _, h, _ in walls)
    roof_area = nb_ * width * depth
    rest = max(target_tris - ground_tris - elem_tris, 1000)
    edge = np.sqrt(2.0 * (wall_area + roof_area) / rest)
    for b, h, fac in walls:
        nh = max(1, int(round(h / edge)))
        for p0, t, L, n in fac:
            add(*_oriented_patch(p0, L * t, h * ez, max(1, int(round(L / edge))), nh, n))
        x0, y0 = fac[0][0][0], fac[0][0][1]
        add(*grid_patch([x0, y0, h], [width, 0, 0], [0, depth, 0],
                        max(1, int(round(width / edge))), max(1, int(round(depth / edge)))))
    xyz = np.concatenate(xs)
    idx = np.concatenate(ts)
    gmask = np.concatenate(ground)
    v = np.concatenate(pv)
    dirs = np.concatenate(pdir)
    params = np.concatenate(pi_)
    if weld:
        # one vertex per position: faces of a box share their edge vertices
        # (coincident vertices of an element share depth fraction and group)
        uq, inv = np.unique(xyz.astype(np.float32), axis=0, return_inverse=True)
        inv = inv.ravel()
        xyz = uq.astype(np.float64)
        idx = inv[idx]
        g2 = np.zeros(uq.shape[0], bool)
        g2[inv[gmask]] = True
        gmask = g2
        v, first = np.unique(inv[v], return_index=True)
        dirs, params = dirs[first], params[first]
    nt = idx.shape[0]
    nq = nt // 2
    eps = np.repeat(rng.uniform(*eps_range, nq), 2).astype(np.float32)
    alpha = np.repeat(rng.uniform(*alpha_range, nq), 2)
    cap = np.repeat(rng.uniform(*capacity_range, nq), 2).astype(np.float32)
    el, az = np.radians(sun_elevation_deg), np.radians(sun_azimuth_deg)
    sun = np.array([np.cos(el) * np.cos(az), np.cos(el) * np.sin(az), np.sin(el)])
    sd = _uniform_scene("facade", xyz, idx, rho_d=1.0 - eps, heat_capacity=cap, kelvin=kelvin)
    if sun_model == "flux":
        cosi = np.clip(sd.normals() @ sun, 0.0, None)
        sd.heat_flux = (alpha * sun_flux * cosi).astype(np.float32)
    v = v.astype(np.uint32)
    params = params.astype(np.uint32)
    base = np.asarray(xyz, np.float64).copy()
    base[v] -= dirs * depths[params][:, None]
    areas = sd.triangle_areas()
    sd.meta.update(seed=seed, buildings=nb_, edge=float(edge), sun=sun.tolist(),
                   mean_triangle_area=float(areas.mean()), total_area=float(areas.sum()),
                   alpha=alpha.astype(np.float32), heights=heights,
                   param_vertex=v, param_index=params, param_dir=dirs.astype(np.float32),
                   depths=depths, base_xyz=base, ground_vertex=gmask,
                   n_params=depths.size, sun_model=sun_model)
    return _with_sun(sd, sun_model, sun, sun_flux, sun_triangles)


def facade_vertices(sd, depths):
    """Vertex positions of a facade scene at the given extrusion depths."""
    m = sd.meta
    x = m["base_xyz"].copy()
    d = np.asarray(depths, np.float64)
    x[m["param_vertex"]] += m["param_dir"].astype(np.float64) * d[m["param_index"]][:, None]
    return x.astype(np.float32)


def facade_parametrization(sd):
    """(vertex, param, dxdp, n_params) of the facade depths, for
    irtt_project_gradient (the reference's vertex-subset parametrizations,
    thermal_params.hpp:41-130)."""
    m = sd.meta
    return m["param_vertex"], m["param_index"], m["param_dir"], m["n_params"]


def sun_flux(sd, elevation_deg, azimuth_deg, flux=800.0):
    """Absorbed direct-sun flux per triangle (W/m^2) for a city block:
    alpha * flux * max(0, n.s), unshadowed."""
    el, az = np.radians(elevation_deg), np.radians(azimuth_deg)
    sun = np.array([np.cos(el) * np.cos(az), np.cos(el) * np.sin(az), np.sin(el)])
    cosi = np.clip(sd.normals() @ sun, 0.0, None)
    return (sd.meta["alpha"] * flux * cosi).astype(np.float32)


def bin_sources(sd, heat_flux, bin_mode=0):
    """External absorbed power per bin (W) for a per-triangle flux, assigned
    as the library assembles it: flux * area to a triangle bin, a third of it
    to each corner's vertex bin."""
    w = np.asarray(heat_flux, np.float32).astype(np.float64) * sd.triangle_areas()
    if bin_mode == 1:
        return w
    out = np.zeros(sd.nv)
    for c in range(3):
        np.add.at(out, sd.idx[:, c], w / 3.0)
    return out


def height_parametrization(sd, n_params=None):
    """Building heights as design parameters of a city block: raising
    building b by dh moves each of its vertices by z / h_b * dh along z (walls
    stretch, the roof lifts).  Returns (vertex, param, dxdp, n_params) for
    irtt_project_gradient; building b maps to parameter b % n_params.
    (The reference's vertex-subset parametrizations, thermal_params.hpp:41-130.)"""
    owner = sd.meta["vertex_building"]
    h = sd.meta["heights"]
    n = len(h) if n_params is None else n_params
    v = np.nonzero(owner >= 0)[0].astype(np.uint32)
    b = owner[v]
    d = np.zeros((v.size, 3), np.float32)
    d[:, 2] = sd.xyz[v, 2] / h[b]
    return v, (b % n).astype(np.uint32), d, n
