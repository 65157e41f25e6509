"""Pins the oracle's adjoint trace (thermal_tracing_unified.glsl:292-599) on
scenes where a visibility edge dominates the derivative, against central
finite differences with common random numbers (the same photons traced on
the displaced geometry) of phi(p) = lambda^T T(p) x4.  The trace returns
-d phi / d vertex, projected here onto the parameter.

Scenes (irtt.scenes): ``blocker_plates`` -- emitter plate, opaque strip,
receiver plate; the parameter moves the strip's edge at x = 0.6 (its width)
or the whole strip (its lateral position).  ``fin_plates`` -- a vertical fin
whose top edge the emitter's rays reach at grazing incidence; the parameter
is the fin's height.  x4 = 1 on the emitter only, lambda = 1 on the receiver
(the shadow side) or on the occluder.

What the cases establish (measured ratios adjoint / FD in the comments; the
bands allow for the trace's Monte-Carlo noise at 2.6 M photons):
* the reference's edge term carries the shadow side of a moving edge:
  0.90 on the blocker, with no other term contributing;
* it omits the occluder's own gain when its boundary triangles stretch:
  0.0 with lambda on the strip; the hit-triangle area Jacobian
  (ADJ_HIT_AREA) gives 0.99; the edge-gain extension (ADJ_EDGE_GAIN)
  0.88 for stretching but 1.49 for a rigid translation, which the continuous
  term already carries (0.99);
* its irradiance derivative is too small at grazing incidence: 0.18 on the
  fin, 1.02 with the projection factor (ADJ_EDGE_PROJECTED);
* every edge event's split (shadow, origin, edge vertices) sums to zero,
  so a rigid motion of the whole scene has no edge-term gradient;
* under a collimated sun (``balcony_wall``: a balcony's shadow on a wall),
  the reference's continuous term differentiates a point-to-point geometric
  term a parallel beam does not have: 0.26 (sun at 45 degrees) and 0.10 (30
  degrees, 60 degrees azimuth) of the balcony's FD gain where the true
  continuous part is zero; the beam transfer (ADJ_COLLIMATED) removes it,
  leaving the hit-area term at 0.99 and 0.98; the wall's loss is carried by
  the projected edge term (0.85 and 1.17).
"""
import numpy as np
import pytest

SIG = 5.670374419e-8
REF = 7            # area | continuous | edge: the reference's terms
GAIN = REF | 16    # + edge gain (extension)
HIT = REF | 32     # + hit-triangle area Jacobian (extension)
PROJ = REF | 64    # + projected edge factor (extension)
BEAM = 128         # collimated first segments: beam transfer (extension)


def _phi(oracle, sd, params, lam, x4):
    from oracle import physics
    sc = oracle.OracleScene.from_irtt(sd)
    acc, _, _ = oracle.transport(sc, params)
    _, _, Q = oracle.photon_weights(sc, params)
    _, ea, _, _ = physics.bins(sd, 0)
    return SIG * np.sum(lam * ea * x4) - Q * (lam @ (acc.astype(np.float64) @ x4))


def _fd(oracle, make, verts, axis, h, params, lam, x4):
    sp, sm = make(), make()
    sp.xyz[verts, axis] += h
    sm.xyz[verts, axis] -= h
    return (_phi(oracle, sp, params, lam, x4) - _phi(oracle, sm, params, lam, x4)) / (2 * h)


def _adj(oracle, sd, params, lam, x4, verts, axis, flags):
    g, _ = oracle.adjoint(oracle.OracleScene.from_irtt(sd), params, lam, x4, flags=flags)
    return -g[verts, axis].sum()


@pytest.fixture(scope="module")
def params(oracle):
    return oracle.make_params(256, 8, batch_seed=5, bin_mode=0)


def _weights(sd, lam_range):
    e0, e1 = sd.meta["emitter_verts"]
    x4 = np.zeros(sd.nv)
    x4[e0:e1] = 1.0
    lam = np.zeros(sd.nv)
    lam[lam_range[0]:lam_range[1]] = 1.0
    return lam, x4


def test_edge_term_carries_the_shadow_side(oracle, params):
    from irtt import scenes
    sd = scenes.blocker_plates()
    b1 = sd.meta["blocker_verts"][1]
    lam, x4 = _weights(sd, (b1, sd.nv))
    ev = sd.meta["edge_vertex"]
    fd = _fd(oracle, scenes.blocker_plates, ev, 0, 1e-2, params, lam, x4)
    ref = _adj(oracle, sd, params, lam, x4, ev, 0, REF)
    proj = _adj(oracle, sd, params, lam, x4, ev, 0, PROJ)
    rest = _adj(oracle, sd, params, lam, x4, ev, 0, REF & ~4)
    print(f"shadow side: FD {fd:.4e}, reference {ref / fd:.3f}, projected {proj / fd:.3f}")
    assert fd > 0          # a wider strip shades the receiver
    assert rest == 0.0     # only the edge term sees the receiver's loss
    assert 0.8 < ref / fd < 1.15
    assert 0.85 < proj / fd < 1.2


@pytest.mark.parametrize("receiver", [True, False])
def test_occluder_gain_needs_the_hit_area_term(oracle, params, receiver):
    """lambda on the strip: moving its edge out stretches its boundary
    triangles, which then absorb more (with or without a receiver behind)."""
    from irtt import scenes

    def make():
        return scenes.blocker_plates(receiver=receiver)
    sd = make()
    lam, x4 = _weights(sd, sd.meta["blocker_verts"])
    ev = sd.meta["edge_vertex"]
    fd = _fd(oracle, make, ev, 0, 1e-2, params, lam, x4)
    r = {k: _adj(oracle, sd, params, lam, x4, ev, 0, f) / fd
         for k, f in (("ref", REF), ("gain", GAIN), ("hit", HIT))}
    print(f"occluder side (receiver={receiver}): FD {fd:.4e}, ratios {r}")
    assert fd < 0
    assert abs(r["ref"]) < 0.05
    assert 0.9 < r["hit"] < 1.1
    assert 0.8 < r["gain"] < 1.1


def test_rigid_translation_needs_no_boundary_term(oracle, params):
    """Translating the strip moves its boundary with its material points: the
    continuous term alone carries the strip's change; the hit-area term adds
    nothing, the edge-gain term counts the boundary a second time."""
    from irtt import scenes
    sd = scenes.blocker_plates()
    b0, b1 = sd.meta["blocker_verts"]
    lam, x4 = _weights(sd, (b0, b1))
    bv = np.arange(b0, b1)
    fd = _fd(oracle, scenes.blocker_plates, bv, 0, 1e-2, params, lam, x4)
    r = {k: _adj(oracle, sd, params, lam, x4, bv, 0, f) / fd
         for k, f in (("ref", REF), ("gain", GAIN), ("hit", HIT))}
    print(f"rigid translation: FD {fd:.4e}, ratios {r}")
    assert 0.9 < r["ref"] < 1.1 and 0.9 < r["hit"] < 1.1
    assert r["gain"] > 1.3


def test_grazing_edge_needs_the_projection(oracle, params):
    from irtt import scenes
    sd = scenes.fin_plates()
    f1 = sd.meta["fin_verts"][1]
    lam, x4 = _weights(sd, (f1, sd.nv))
    ev = sd.meta["edge_vertex"]
    fd = _fd(oracle, scenes.fin_plates, ev, 2, 5e-3, params, lam, x4)
    ref = _adj(oracle, sd, params, lam, x4, ev, 2, REF)
    proj = _adj(oracle, sd, params, lam, x4, ev, 2, PROJ)
    print(f"grazing edge: FD {fd:.4e}, reference {ref / fd:.3f}, projected {proj / fd:.3f}")
    assert fd > 0
    assert 0.0 < ref / fd < 0.35
    assert 0.85 < proj / fd < 1.2


@pytest.mark.parametrize("flags", [4, 4 | 16, 4 | 64])
def test_edge_split_sums_to_zero(oracle, flags):
    """Each edge event adds -dE to the edge vertex and fractions of dE that
    sum to one to the shadow and origin vertices: the edge term's gradient
    summed over all vertices vanishes (translation invariance)."""
    from irtt import scenes
    sd = scenes.blocker_plates(n=8)
    rng = np.random.default_rng(4)
    lam = rng.random(sd.nv)
    x4 = rng.random(sd.nv)
    p = oracle.make_params(32, 8, batch_seed=1, bin_mode=0)
    g, st = oracle.adjoint(oracle.OracleScene.from_irtt(sd), p, lam, x4, flags=flags)
    assert st["occluded"] > 0
    # the FP32 fractions d1/(d1+d2) and d2/(d1+d2) sum to one within a few ulp
    assert np.all(np.abs(g.sum(0)) <= 1e-6 * np.abs(g).sum(0))


@pytest.mark.parametrize("sun", [(45.0, 0.0), (30.0, 60.0)])
def test_collimated_sun_needs_the_beam_transfer(oracle, sun):
    """A balcony's depth under a collimated sun: its top gains what the wall
    below loses.  With lambda on the balcony the derivative is the hit-area
    term (its triangles stretch; a parallel beam's transfer has no distance
    factor); with lambda on the wall it is the projected edge term."""
    from irtt import scenes
    el, az = sun

    def make(depth=1.0):
        return scenes.balcony_wall(depth=depth, sun_elevation_deg=el, sun_azimuth_deg=az)
    sd = make()
    own = sd.meta["owner"]
    ddir = sd.meta["depth_dir"].astype(np.float64)
    sv = sd.meta["suns"][0]["vertices"]
    x4 = np.zeros(sd.nv)
    x4[sv] = sd.kelvin[sv].astype(np.float64) ** 4
    p = oracle.make_params(64, 8, batch_seed=5, bin_mode=0)
    h = 2e-2
    out = {}
    for name, sel in (("balcony", own == 2), ("wall", own == 1)):
        lam = sel.astype(np.float64)
        fd = (_phi(oracle, make(1.0 + h), p, lam, x4) - _phi(oracle, make(1.0 - h), p, lam, x4)) \
            / (2 * h)
        r = {}
        for k, f in (("ref", REF), ("ext", REF | 32 | 64), ("beam", REF | 32 | 64 | BEAM),
                     ("cont", 2), ("cont_beam", 2 | BEAM)):
            g, _ = oracle.adjoint(oracle.OracleScene.from_irtt(sd), p, lam, x4, flags=f)
            r[k] = -(g * ddir).sum() / fd
        out[name] = r
        print(f"sun {sun}, lambda on the {name}: FD {fd:.4e}, ratios "
              + ", ".join(f"{k} {v:.3f}" for k, v in r.items()))
    b, w = out["balcony"], out["wall"]
    # translated and stretched planar faces keep their cosine: zero to rounding
    assert abs(b["cont_beam"]) < 1e-6 and abs(w["cont_beam"]) < 1e-6
    assert b["cont"] > 0.05             # the point-to-point term's spurious part
    assert 0.93 < b["beam"] < 1.07
    assert not 0.93 < b["ext"] < 1.07
    assert 0.8 < w["beam"] < 1.2 and w["beam"] == w["ext"]
