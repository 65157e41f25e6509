"""The design loop on the device (irtt/design.py DesignLoop with
irtt/params.py maps, constraints and optimisers).

Scene: scenes.blocker_plates -- emitter plate fixed at 500 K, an opaque
strip fixed at 293 K, a free receiver plate (ambient 293 K); the design
parameter translates the strip along x (AxisTranslate), the objective is the
receiver's area-weighted mean temperature.  Moving the strip moves its shadow
across the receiver, so the derivative is carried by the edge term.

* FD mode 0 (central differences of full forward evaluations, common random
  numbers) against the adjoint with the projected edge factor and the
  hit-area term: same sign, ratio in [0.6, 1.4] (at 2.9 M photons; the
  oracle measures 0.90-1.02 for these terms on this scene).
* FD mode 1 returns one FD gradient per step size, largest step first.
* FD mode 2 (transport-only FD with x and lambda held) agrees in sign with
  the projected adjoint trace of the same term.
* With an InPolygonXZ constraint that the strip violates, the loop's
  gradient is the library's design gradient plus the projected penalty
  gradient, and the objective includes the penalty.
* Five Adam steps lower the objective; the optimiser keeps the best state.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FLAGS = 1 | 2 | 4 | 32 | 64   # area, continuous, edge, hit area, projected edge factor


def _setup(irtt, constraints=None):
    from irtt import params as P, scenes
    from irtt.design import DesignLoop, LinearObjective
    sd = scenes.blocker_plates(n=16)
    e0, e1 = sd.meta["emitter_verts"]
    b0, b1 = sd.meta["blocker_verts"]
    sd.kelvin[:] = 293.0
    sd.kelvin[e0:e1] = 500.0
    sd.fixed[:b1] = 1
    par = P.Parametrization(sd.xyz, [(np.arange(b0, b1), P.AxisTranslate((1.0, 0.0, 0.0)), None)])
    tp = irtt.transport_params(256, 8, batch_seed=5, flags=irtt.TRANSPORT_KEEP_ABSORBERS)
    sp = irtt.solver_params(steady=True, ambient_kelvin=293.0, x_tol=1e-7, f_tol=1e-14,
                            lin_rtol=1e-10)
    loop = DesignLoop(irtt, sd, par, None, tp, sp, constraints=constraints, adjoint_flags=FLAGS)
    area = loop.s.bin_data()[0]
    w = np.zeros(sd.nv)
    w[b1:] = area[b1:]
    loop.objective = LinearObjective(w / w.sum())
    return loop, sd


def test_fd_modes_vs_adjoint(irtt):
    loop, sd = _setup(irtt)
    with loop:
        p0 = np.zeros(1)
        fd, adj = loop.fd_check(p0, 2e-2)
        r = adj[0] / fd[0]
        print(f"FD mode 0: fd {fd[0]:.5g} adjoint {adj[0]:.5g} ratio {r:.3f}")
        assert fd[0] != 0.0 and 0.6 < r < 1.4
        steps, fds, adj2 = loop.fd_sweep(p0, 2e-2, n_steps=3)
        assert np.allclose(steps, [2e-2 * np.sqrt(10), 2e-2, 2e-2 / np.sqrt(10)])
        assert fds.shape == (3, 1) and np.isclose(fds[1, 0], fd[0])
        assert np.array_equal(adj2, adj)   # the gradient is bit-reproducible
        tfd, tadj = loop.transport_fd(p0, 2e-2)
        print(f"FD mode 2: fd {tfd[0]:.5g} adjoint trace {tadj[0]:.5g}")
        assert np.sign(tfd[0]) == np.sign(tadj[0]) and 0.5 < tadj[0] / tfd[0] < 1.5


def test_constraint_enters_value_and_gradient(irtt):
    from irtt import params as P
    loop0, sd = _setup(irtt)
    b0, b1 = sd.meta["blocker_verts"]
    with loop0:
        L0 = loop0.forward(np.zeros(1))
        g0 = loop0.gradient()
    con = P.Constraints([P.InPolygonXZ(np.arange(b0, b1), 50.0,
                                       "-1,-1; 0.5,-1; 0.5,2; -1,2")])
    loop, _ = _setup(irtt, con)
    with loop:
        L = loop.forward(np.zeros(1))
        g = loop.gradient()
        gp = np.zeros((sd.nv, 3))
        pen = con.evaluate(loop.xyz, gp)
        trip = loop.par.triplets(np.zeros(1))
        assert pen > 0.0 and L == pytest.approx(L0 + pen, rel=1e-12)
        assert g[0] == pytest.approx(g0[0] + irtt.project_gradient(gp, *trip)[0], rel=1e-12)
        assert g[0] > g0[0]   # the penalty pushes the strip towards -x


def test_adam_lowers_objective(irtt):
    from irtt import params as P
    loop, _ = _setup(irtt)
    with loop:
        opt = P.Adam(step_size=0.05)
        hist = loop.run(np.zeros(1), opt, 5)
        L = hist["objective"]
        print("Adam objective:", np.round(L, 4), "p:", hist["params"])
        assert L.size == 5 and np.isfinite(L).all()
        assert L.min() < L[0]
        assert opt.best_value == L.min()
