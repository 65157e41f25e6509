"""GPU solver parity: the device Newton/BiCGSTAB solve against the FP64
oracle on the same operator (downloaded from the device).  Tolerance: max
|dT| <= 0.05 K (the stated bar); the measured error is printed."""
import numpy as np
import pytest

from oracle import physics, solver

pytestmark = pytest.mark.gpu

TOL_K = 0.05


def _oracle_rhs(s, sd, bin_mode, Ta):
    _, _, _, src = physics.bins(sd, bin_mode)
    return solver.rhs(src, s.column_sums(), Ta)


def test_bin_assembly_matches_restatement(irtt):
    from irtt import scenes
    sd = scenes.city_block(seed=1, nx=2, ny=2, target_tris=2000)
    for mode in (0, 1):
        with irtt.Scene.from_data(sd, mode) as s:
            a, ea, m = s.bin_data()
        oa, oea, om, _ = physics.bins(sd, mode)
        assert np.allclose(a, oa, rtol=1e-12) and np.allclose(ea, oea, rtol=1e-12)
        assert np.allclose(m, om, rtol=1e-12)


def test_single_plate_closed_form(irtt):
    """A lone plate sees nothing: x = (q / (sigma eps) + Ta^4)^(1/4) exactly."""
    from irtt import scenes
    sd = scenes.parallel_plates(n=4, emissivity=0.9)
    na = sd.meta["plate_a_tris"]
    nva = sd.meta["plate_a_verts"]
    xyz, idx = sd.xyz[:nva], sd.idx[:na]
    mats = sd.materials()[:na]
    with irtt.Scene(xyz, idx, mats, irtt.BIN_TRIANGLE) as s:
        s.build_transport(threads_per_tri=64, rays_per_thread=4)
        st = s.solve(irtt.solver_params(ambient_kelvin=293.0))
        x = s.temperatures()
    want = (500.0 / (5.670374419e-8 * 0.9) + 293.0 ** 4) ** 0.25
    print(f"closed form {want:.6f} K, max |dT| {np.abs(x - want).max():.2e} K, {st}")
    # the operator is stored in FP32 (relative rounding <= 6e-8 on sigma eps A,
    # i.e. ~1e-5 K at 360 K); the ambient term uses exact column sums
    assert st["converged"] and np.abs(x - want).max() < 1e-4


@pytest.mark.parametrize("jac", [0, 1])
def test_plates_steady_vs_fp64(irtt, jac):
    """Config 0 physics (eps 0.9, 500 W/m^2 on A, ambient 293 K, triangle bins,
    1024 x 1024), fewer photons: device solve vs FP64 solve of the same T."""
    from irtt import scenes
    sd = scenes.parallel_plates(emissivity=0.9)
    with irtt.Scene.from_data(sd, irtt.BIN_TRIANGLE) as s:
        s.build_transport(threads_per_tri=512, rays_per_thread=8, batch_seed=0)
        T = s.transport().astype(np.float64)
        b = _oracle_rhs(s, sd, 1, 293.0)
        st = s.solve(irtt.solver_params(jacobian=jac, ambient_kelvin=293.0, x_tol=1e-6,
                                        f_tol=1e-14))
        x = s.temperatures()
    fixed = np.zeros(T.shape[0], bool)
    x64 = solver.steady_linear(T, b, fixed, np.zeros(T.shape[0]))
    err = np.abs(x - x64).max()
    print(f"jacobian={jac} max|dT| vs FP64 = {err:.3e} K; plate A mean {x[:512].mean():.3f} K, "
          f"plate B mean {x[512:].mean():.3f} K; {st}")
    assert st["converged"] and err < TOL_K


def test_conduction_and_fixed_vs_fp64(irtt):
    """Vertex bins with the cotangent Laplacian and fixed bins: device solve vs
    FP64 Newton with the restated Laplacian."""
    from irtt import scenes
    sd = scenes.closed_box(4, emissivity=0.85)
    sd.conductivity[:] = 1.5
    sd.thickness[:] = 0.1
    sd.heat_flux[:32] = 300.0
    sd.kelvin[:] = 290.0
    sd.fixed[:25] = 1          # first face's vertices pinned
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        s.build_transport(threads_per_tri=256, rays_per_thread=4, batch_seed=5)
        T = s.transport().astype(np.float64)
        b = _oracle_rhs(s, sd, 0, -1.0)
        st = s.solve(irtt.solver_params(ambient_kelvin=-1.0, x_tol=1e-7, f_tol=1e-16))
        x = s.temperatures()
    L = physics.cotan_laplacian(sd.xyz, sd.idx, sd.conductivity * sd.thickness, sd.nv)
    fixed = sd.fixed.astype(bool)
    x64, _ = solver.newton(T, b, fixed, sd.kelvin.astype(np.float64), sd.kelvin.astype(np.float64),
                           lap=L, tol=1e-12)
    err = np.abs(x - x64).max()
    print(f"max|dT| vs FP64 = {err:.3e} K; {st}")
    assert np.all(x[:25] == 290.0)
    assert st["converged"] and err < TOL_K


def test_transient_step_vs_fp64(irtt):
    from irtt import scenes
    sd = scenes.parallel_plates(n=8, emissivity=0.9, kelvin=293.0)
    sd.heat_capacity[:] = 2e4  # areal capacity (density 1, thickness 1)
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        s.build_transport(threads_per_tri=256, rays_per_thread=8, batch_seed=3)
        T = s.transport().astype(np.float64)
        b = _oracle_rhs(s, sd, 0, 293.0)
        _, _, mass = s.bin_data()
        x0 = s.temperatures()
        st = s.solve(irtt.solver_params(steady=False, dt=300.0, ambient_kelvin=293.0,
                                        x_tol=1e-8, f_tol=1e-16))
        x1 = s.temperatures()
    fixed = np.zeros(len(x0), bool)
    x64, _ = solver.newton(T, b, fixed, x0, x0, mass=mass, xprev=x0, dt=300.0, steady=False,
                           tol=1e-12)
    err = np.abs(x1 - x64).max()
    print(f"transient step: max|dT| vs FP64 = {err:.3e} K, mean rise {np.mean(x1 - x0):.4f} K; {st}")
    assert st["converged"] and err < TOL_K


@pytest.mark.parametrize("restart", [30, 5])
def test_gmres_steady_conduction_and_adjoint_vs_fp64(irtt, restart):
    """IRTT_KRYLOV_GMRES: the Newton steps and the adjoint solve with
    restarted GMRES (same preconditioner and stop rule as BiCGSTAB), on the
    conduction + fixed-bin case against the FP64 Newton, and the transposed
    solve against a dense FP64 solve.  restart 5 exercises restarts."""
    from irtt import scenes
    sd = scenes.closed_box(4, emissivity=0.85)
    sd.conductivity[:] = 1.5
    sd.thickness[:] = 0.1
    sd.heat_flux[:32] = 300.0
    sd.kelvin[:] = 290.0
    sd.fixed[:25] = 1
    p = irtt.solver_params(ambient_kelvin=-1.0, x_tol=1e-9, f_tol=1e-18, lin_rtol=1e-12,
                           krylov=irtt.KRYLOV_GMRES, gmres_restart=restart, max_linear=400)
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        s.build_transport(threads_per_tri=256, rays_per_thread=4, batch_seed=5)
        T = s.transport().astype(np.float64)
        b = _oracle_rhs(s, sd, 0, -1.0)
        st = s.solve(p)
        x = s.temperatures()
        rng = np.random.default_rng(1)
        rhs = rng.normal(size=s.n_bins)
        lam, ast = s.adjoint_solve(rhs, params=p)
        # BiCGSTAB on the same system for comparison
        s.set_vertex_state(sd.kelvin, sd.fixed)
        s.solve(irtt.solver_params(ambient_kelvin=-1.0, x_tol=1e-9, f_tol=1e-18, lin_rtol=1e-12))
        xb = s.temperatures()
    L = physics.cotan_laplacian(sd.xyz, sd.idx, sd.conductivity * sd.thickness, sd.nv)
    fixed = sd.fixed.astype(bool)
    x64, _ = solver.newton(T, b, fixed, sd.kelvin.astype(np.float64),
                           sd.kelvin.astype(np.float64), lap=L, tol=1e-12)
    err = np.abs(x - x64).max()
    J = T * (4.0 * x ** 3)[None, :] + (L.toarray() if hasattr(L, "toarray") else L)
    f = ~fixed
    l64 = np.zeros_like(x)
    l64[f] = np.linalg.solve(J[np.ix_(f, f)].T, rhs[f])
    rel = np.abs(lam - l64).max() / np.abs(l64).max()
    print(f"GMRES({restart}): max |dT| vs FP64 {err:.2e} K, vs BiCGSTAB "
          f"{np.abs(x - xb).max():.2e} K, adjoint rel {rel:.2e}; {st}; adjoint {ast}")
    assert st["converged"] and err < 1e-6 and np.abs(x - xb).max() < 1e-6
    # the Krylov stop rule bounds the residual at 1e-12 relative; the solution
    # error is that times the conditioning of J (about 1e4 here)
    assert rel < 1e-7


def test_gmres_on_sparse_city(irtt):
    from irtt import scenes
    sd = scenes.city_block(seed=6, nx=2, ny=2, target_tris=2500)
    src = physics.bins(sd, 0)[3]
    out = []
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        s.build_transport(threads_per_tri=64, rays_per_thread=4, batch_seed=2,
                          layout=irtt.LAYOUT_SPARSE)
        T = s.transport().astype(np.float64)
        cs = s.column_sums()
        for kr in (irtt.KRYLOV_BICGSTAB, irtt.KRYLOV_GMRES):
            s.set_temperatures(np.zeros(s.n_bins))
            st = s.solve(irtt.solver_params(ambient_kelvin=293.0, x_tol=1e-9, f_tol=1e-18,
                                            lin_rtol=1e-12, krylov=kr))
            out.append((s.temperatures(), st))
    n = T.shape[0]
    x64 = solver.steady_linear(T, solver.rhs(src, cs, 293.0), np.zeros(n, bool), np.zeros(n))
    for x, st in out:
        print(f"{st}: max |dT| vs FP64 {np.abs(x - x64).max():.2e} K")
        assert st["converged"] and np.abs(x - x64).max() < 1e-7


def _bicgstab_np(A, b, iters):
    """Right-Jacobi-preconditioned BiCGSTAB (van der Vorst 1992) from x = 0,
    exactly `iters` iterations (no stop test): the iteration the device runs."""
    dinv = 1.0 / np.diag(A)
    x = np.zeros_like(b)
    r = b.copy()
    R = r.copy()
    rho = alpha = omega = 1.0
    p = np.zeros_like(b)
    v = np.zeros_like(b)
    for k in range(iters):
        rho_new = R @ r
        beta = 0.0 if k == 0 else (rho_new / rho) * (alpha / omega)
        p = r + beta * (p - omega * v)
        ph = dinv * p
        v = A @ ph
        alpha = rho_new / (R @ v)
        s = r - alpha * v
        sh = dinv * s
        t = A @ sh
        omega = (t @ s) / (t @ t)
        x = x + alpha * ph + omega * sh
        r = s - omega * t
        rho = rho_new
    return x


@pytest.mark.parametrize("max_linear", [3, 11])
def test_bicgstab_honours_max_linear(irtt, max_linear):
    """A linear solve stops after exactly max_linear iterations, also when
    max_linear is not a multiple of the iterations enqueued per host check:
    the adjoint lambda equals max_linear restated iterations on J^T."""
    from irtt import scenes
    sd = scenes.parallel_plates(n=8, emissivity=0.9)
    rng = np.random.default_rng(7)
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        s.build_transport(threads_per_tri=128, rays_per_thread=4, batch_seed=2)
        T = s.transport().astype(np.float64)
        x = 300.0 + 20.0 * rng.random(s.n_bins)
        rhs = rng.normal(size=s.n_bins)
        lam, st = s.adjoint_solve(rhs, x=x, params=irtt.solver_params(
            max_linear=max_linear, lin_rtol=1e-300))
    JT = (T * (4.0 * x ** 3)[None, :]).T
    want = _bicgstab_np(JT, rhs, max_linear)
    err = np.abs(lam - want).max() / np.abs(want).max()
    print(f"max_linear={max_linear}: iterations {st['linear_iters']}, rel. diff vs restated "
          f"{err:.2e}")
    assert st["linear_iters"] == max_linear
    assert err < 1e-9
