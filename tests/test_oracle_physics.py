"""Pins the physics-assembly and FP64 solver restatements
(thermal_data.cpp:56-81,216-313; thermal_solver.cpp:13-229)."""
import numpy as np

from oracle import physics, solver


def test_bin_areas_and_mass():
    from irtt import scenes
    sd = scenes.city_block(seed=1, nx=2, ny=2, target_tris=2000)
    raw, area = physics.triangle_areas(sd.xyz, sd.idx)
    for mode in (0, 1):
        a, ea, m, src = physics.bins(sd, mode)
        assert np.isclose(a.sum(), area.sum(), rtol=1e-12)
        # areal capacity (density=1, thickness=1) times area
        cap = sd.heat_capacity.astype(np.float64)
        assert np.isclose(m.sum(), (cap * raw).sum(), rtol=1e-12)
        assert np.all(ea <= a + 1e-12) and np.all(ea >= 0.84 * a - 1e-12)


def test_cotan_laplacian_properties():
    from irtt import scenes
    rng = np.random.default_rng(0)
    xyz, tri = scenes.grid_patch([0, 0, 0], [3, 0, 0], [0, 2, 0], 6, 4)
    xyz = xyz + np.concatenate([rng.normal(size=(xyz.shape[0], 2)) * 0.05,
                                np.zeros((xyz.shape[0], 1))], 1)
    nv = xyz.shape[0]
    L = physics.cotan_laplacian(xyz, tri, np.full(tri.shape[0], 2.5), nv)
    assert np.allclose(L, L.T, atol=1e-12)
    assert np.allclose(L.sum(1), 0.0, atol=1e-12)
    # a linear field is harmonic: (L f)_i = 0 at interior vertices
    f = 0.7 * xyz[:, 0] - 1.3 * xyz[:, 1]
    lf = L @ f.astype(np.float32).astype(np.float64)
    boundary = np.isin(np.arange(nv), np.unique(np.concatenate([
        np.arange(0, nv, 5), np.arange(4, nv, 5), np.arange(5), np.arange(nv - 5, nv)])))
    assert np.abs(lf[~boundary]).max() < 1e-5
    # positive semi-definite
    assert np.linalg.eigvalsh(L).min() > -1e-10


def test_newton_single_element_closed_form():
    """One emitting bin, nothing else visible: sigma eps A x^4 = q A + sigma eps A Ta^4
    (ambient through the column sum), x = (q / (sigma eps) + Ta^4)^(1/4)."""
    sig, eps, A, q, Ta = 5.670374419e-8, 0.9, 2.0, 500.0, 293.0
    T = np.array([[sig * eps * A]])
    b = solver.rhs([q * A], T.sum(0), Ta)
    x, it = solver.newton(T, b, np.array([False]), np.array([0.0]), np.array([0.0]))
    want = (q / (sig * eps) + Ta ** 4) ** 0.25
    assert abs(x[0] - want) < 1e-9
    assert abs(solver.steady_linear(T, b, np.array([False]), np.array([0.0]))[0] - want) < 1e-9


def test_newton_matches_linear_solve_in_y():
    """Steady state without conduction is linear in y = x^4; Newton (in x)
    must land on the same fixed point, with fixed bins respected."""
    rng = np.random.default_rng(1)
    n = 40
    K = rng.random((n, n)) * 1e-9
    np.fill_diagonal(K, 0)
    T = np.diag(K.sum(0) * 1.3) - K      # column-diagonally dominant, like the operator
    fixed = np.zeros(n, bool)
    fixed[:5] = True
    kfix = np.full(n, 280.0)
    b = rng.random(n) * 5.0
    x_lin = solver.steady_linear(T, b, fixed, kfix)
    x_new, it = solver.newton(T, b, fixed, kfix, np.zeros(n))
    assert np.abs(x_new - x_lin).max() < 1e-8
    assert np.all(x_new[:5] == 280.0)


def test_transient_step_limits():
    """Implicit Euler: dt -> 0 keeps the state, dt -> inf reaches steady state."""
    rng = np.random.default_rng(2)
    n = 20
    K = rng.random((n, n)) * 1e-9
    np.fill_diagonal(K, 0)
    T = np.diag(K.sum(0) * 1.5) - K
    b = rng.random(n) * 3.0
    fixed = np.zeros(n, bool)
    x0 = np.full(n, 290.0)
    mass = np.full(n, 1e4)
    xs = solver.steady_linear(T, b, fixed, x0)
    x_small, _ = solver.newton(T, b, fixed, x0, x0, mass=mass, xprev=x0, dt=1e-9, steady=False)
    x_big, _ = solver.newton(T, b, fixed, x0, x0, mass=mass, xprev=x0, dt=1e14, steady=False)
    assert np.abs(x_small - x0).max() < 1e-6
    assert np.abs(x_big - xs).max() < 1e-3
