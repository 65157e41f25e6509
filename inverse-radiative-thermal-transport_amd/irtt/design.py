"""Design loop over a parametrised scene: objective, gradient, finite-
difference checks and descent, each evaluation through the library.

The reference runs the same loop inside ThermalRenderer
(src/implementations/thermal_renderer/thermal_renderer.cpp:1135-1337:
forward simulation, adjoint, constraints, projectGradient, optimiser step,
geometry update; the FD-check modes at :1150-1256; thermal_params.cpp:101-156,
thermal_objective.cpp:19-78).  One evaluation here is:
irtt_scene_update_vertices -> LBVH -> photon transport (same seed: common
random numbers between steps and finite-difference evaluations) -> Newton
solve; the gradient is irtt_design_gradient (adjoint solve, adjoint trace,
source term, projection through the parametrization's Jacobian) plus the
constraint penalties projected the same way.
"""
import time

import numpy as np


def ground_weights(s, sd):
    """dL/dx of the mean ground temperature L = sum_i A_i x_i / sum_i A_i over
    the ground bins (a linear objective; the reference's objective is the
    weighted squared deviation, thermal_objective.cpp:47-57, params.Objective)."""
    area = s.bin_data()[0]
    g = np.asarray(sd.meta["ground_vertex"], bool)
    w = np.where(g, area, 0.0)
    return w / w.sum()


class LinearObjective:
    """L = w . x (a weighted mean temperature); eval -> (value, dL/dx)."""

    def __init__(self, w):
        self.w = np.asarray(w, np.float64)
        self.missing = 0

    def update_targets(self, x, step=0):
        return False

    def eval(self, x, step=0):
        return float(self.w @ np.asarray(x, np.float64)), self.w


class LinearParametrization:
    """x = base + sum_k dxdp_k p[param_k] at vertex_k: a vertex-subset map
    with a constant Jacobian (extrusion depths, building heights)."""

    def __init__(self, base_xyz, vertex, param, dxdp, n_params):
        self.base = np.asarray(base_xyz, np.float64).reshape(-1, 3)
        self.v = np.asarray(vertex, np.int64)
        self.k = np.asarray(param, np.int64)
        self.d = np.asarray(dxdp, np.float32).reshape(-1, 3)
        self.n_params = int(n_params)
        if self.v.size and (self.v.max() >= self.base.shape[0] or self.k.max() >= self.n_params):
            raise ValueError("parametrization entry out of range")

    def positions(self, p):
        p = np.asarray(p, np.float64)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        x = self.base.copy()
        np.add.at(x, self.v, self.d.astype(np.float64) * p[self.k][:, None])
        return x

    def triplets(self, p):
        return self.v.astype(np.uint32), self.k.astype(np.uint32), self.d, self.n_params


class DesignLoop:
    """Objective, gradient and optimisation of a scene under a
    parametrization.

    irtt: the binding module; sd: SceneData; par: an object with
    ``positions(p)`` (V x 3 world positions) and ``triplets(p)`` ((vertex,
    param, dxdp, n_params) for irtt_project_gradient), such as
    params.Parametrization or LinearParametrization; objective: an object
    with ``eval(x) -> (value, dL/dx)`` over the bins (params.Objective,
    LinearObjective); constraints: params.Constraints or None (penalties on
    vertex positions, added to the objective and, through the same
    projection, to the gradient)."""

    def __init__(self, irtt, sd, par, objective, transport, solver, constraints=None,
                 adjoint_flags=None, bin_mode=None, kelvin=None, edge_ratio=0.05, tiny_eps=1e-2):
        self.irtt = irtt
        self.sd = sd
        self.par = par
        self.objective = objective
        self.constraints = constraints
        self.tp = transport
        self.sp = solver
        self.flags = adjoint_flags if adjoint_flags is not None else (
            irtt.ADJ_AREA | irtt.ADJ_CONTINUOUS | irtt.ADJ_EDGE)
        self.edge_ratio, self.tiny_eps = edge_ratio, tiny_eps
        self.bin_mode = irtt.BIN_VERTEX if bin_mode is None else bin_mode
        self.s = irtt.Scene.from_data(sd, self.bin_mode)
        self.kelvin = np.asarray(sd.kelvin if kelvin is None else kelvin, np.float32)
        self.p = None
        self.timing = {}

    def close(self):
        self.s.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _set_state(self):
        if self.bin_mode == self.irtt.BIN_VERTEX:
            self.s.set_vertex_state(self.kelvin, self.sd.fixed)
        else:
            k = self.kelvin
            if k.size == self.sd.nv:
                k = k[self.sd.idx].mean(1).astype(np.float32)
            fx = self.sd.fixed
            if fx is not None and np.asarray(fx).size == self.sd.nv:
                fx = np.asarray(fx, bool)[self.sd.idx].all(1)
            self.s.set_bin_state(k, fx)

    def set_geometry(self, p):
        """Vertex positions at p, LBVH and photon transport."""
        t0 = time.perf_counter()
        self.xyz = self.par.positions(p)
        self.s.update_vertices(self.xyz.astype(np.float32))
        self.tst = self.s.build_transport(self.tp)
        self.timing["transport_ms"] = (time.perf_counter() - t0) * 1e3
        self.p = np.array(p, np.float64)

    def forward(self, p):
        """Objective plus penalties at p (geometry, transport, solve)."""
        self.set_geometry(p)
        t1 = time.perf_counter()
        self._set_state()
        sst = self.s.solve(self.sp)
        self.x = self.s.temperatures()
        self.timing["solve_ms"] = (time.perf_counter() - t1) * 1e3
        self.timing.update(nnz=self.tst["nnz"], rays=self.tst["rays"], converged=sst["converged"])
        if getattr(self.objective, "missing", 0):
            self.objective.update_targets(self.x)
        L, self.dLdx = self.objective.eval(self.x)
        self.gpen = np.zeros((self.xyz.shape[0], 3))
        self.penalty = self.constraints.evaluate(self.xyz, self.gpen) if self.constraints else 0.0
        self.value = float(L + self.penalty)
        return self.value

    def gradient(self, flags=None):
        """dL/dp at the state of the last forward()."""
        if self.p is None:
            raise RuntimeError("gradient() needs a forward() first")
        t0 = time.perf_counter()
        trip = self.par.triplets(self.p)
        dp, gx, lam, st = self.s.design_gradient(self.dLdx, params=trip, solver=self.sp,
                                                 edge_ratio=self.edge_ratio, tiny_eps=self.tiny_eps,
                                                 flags=self.flags if flags is None else flags)
        if self.constraints is not None and self.gpen.any():
            dp = dp + self.irtt.project_gradient(self.gpen, *trip)
        self.grad_x, self.lam = gx, lam
        self.timing["gradient_ms"] = (time.perf_counter() - t0) * 1e3
        return dp

    def run(self, p0, optimizer, steps, log=None, lo=None, hi=None):
        """`steps` optimiser steps from p0 (thermal_renderer.cpp:1259-1278),
        each update clipped to [lo, hi] when bounds are given.  Returns a dict
        with the objective and gradient-norm histories, the final parameters
        and the optimiser's best state."""
        p = np.array(p0, np.float64)
        hist, gnorm = [], []
        for k in range(steps):
            L = self.forward(p)
            g = self.gradient()
            hist.append(L)
            gnorm.append(float(np.linalg.norm(g)))
            if log:
                log(k, L, self.timing)
            p = optimizer.step(p, g, L)
            if lo is not None or hi is not None:
                p = np.clip(p, lo, hi)
        return {"objective": np.array(hist), "grad_norm": np.array(gnorm), "params": p,
                "best_value": optimizer.best_value, "best_params": optimizer.best_params}

    def fd_gradient(self, p, h, dims=None):
        """Central differences (L(p + h/2 e_k) - L(p - h/2 e_k)) / h of the
        full forward evaluation, same photon seed on both sides (the
        reference's finiteDifferenceApprox)."""
        p = np.array(p, np.float64)
        dims = range(p.size) if dims is None else dims
        out = np.full(p.size, np.nan)
        for k in dims:
            q = p.copy()
            q[k] = p[k] + 0.5 * h
            lp = self.forward(q)
            q[k] = p[k] - 0.5 * h
            lm = self.forward(q)
            out[k] = (lp - lm) / h
        return out

    def fd_check(self, p, h, dims=None):
        """FD mode 0 (thermal_renderer.cpp:1152-1164): central FD with step h
        and the adjoint gradient at p.  Returns (fd, adjoint)."""
        fd = self.fd_gradient(p, h, dims)
        self.forward(p)
        return fd, self.gradient()

    def fd_sweep(self, p, h, dims=None, n_steps=11, factor=np.sqrt(10.0)):
        """FD mode 1 (thermal_renderer.cpp:1165-1209): central FD over
        n_steps step sizes h * factor^(n/2 - i), largest first, and the
        adjoint gradient at p.  Returns (steps, fd (n_steps x P), adjoint)."""
        steps = h * factor ** (n_steps // 2 - np.arange(n_steps, dtype=np.float64))
        fd = np.stack([self.fd_gradient(p, hh, dims) for hh in steps])
        self.forward(p)
        return steps, fd, self.gradient()

    def transport_fd(self, p, h, dims=None):
        """FD mode 2 (thermal_renderer.cpp:1210-1256): with the solution x
        and the adjoint state lambda at p held fixed, central differences of
        the transport term only, -lambda^T (T(p + h/2) - T(p - h/2)) x^4 / h
        (T row = absorbing bin, column = emitting bin), next to the projected
        adjoint trace of the same term.  Returns (fd, adjoint)."""
        p = np.array(p, np.float64)
        self.forward(p)
        self.gradient()
        lam, x4 = self.lam.copy(), self.x.astype(np.float64) ** 4
        trip = self.par.triplets(p)
        g, _ = self.s.adjoint_trace(lam, x4, edge_ratio=self.edge_ratio, tiny_eps=self.tiny_eps,
                                    flags=self.flags)
        adj = self.irtt.project_gradient(g, *trip)
        dims = range(p.size) if dims is None else dims
        fd = np.full(p.size, np.nan)
        for k in dims:
            q = p.copy()
            q[k] = p[k] + 0.5 * h
            self.set_geometry(q)
            tp = lam @ self.s.apply_operator(x4)[0]
            q[k] = p[k] - 0.5 * h
            self.set_geometry(q)
            tm = lam @ self.s.apply_operator(x4)[0]
            fd[k] = -(tp - tm) / h
        self.set_geometry(p)
        return fd, adj


class FacadeDesign(DesignLoop):
    """A facade scene (scenes.facade) with its extrusion depths as the
    design parameters and the mean ground temperature as the objective."""

    def __init__(self, irtt, sd, transport, solver, x0=None, adjoint_flags=0, constraints=None):
        m = sd.meta
        par = LinearParametrization(m["base_xyz"], m["param_vertex"], m["param_index"],
                                    m["param_dir"], m["n_params"])
        super().__init__(irtt, sd, par, None, transport, solver, constraints=constraints,
                         adjoint_flags=adjoint_flags or None, kelvin=x0)
        self.objective = LinearObjective(ground_weights(self.s, sd))
        self.depths = np.asarray(m["depths"], np.float64).copy()

    def central_difference(self, depths, k, h):
        d = np.array(depths, np.float64)
        d[k] += h
        lp = self.forward(d)
        d[k] -= 2 * h
        lm = self.forward(d)
        return (lp - lm) / (2 * h)

    def descend(self, steps, max_step=0.05, lo=0.0, hi=1.5, log=None):
        """Normalised gradient descent on the depths (each step moves the
        largest-gradient depth by max_step), clipped to [lo, hi].  Returns the
        objective history and the best (objective, depths)."""
        d = self.depths.copy()
        hist, best = [], (np.inf, d.copy())
        for k in range(steps):
            L = self.forward(d)
            g = self.gradient()
            hist.append(L)
            if L < best[0]:
                best = (L, d.copy())
            if log:
                log(k, L, self.timing)
            gmax = np.abs(g).max()
            if not gmax > 0:
                break
            d = np.clip(d - max_step * g / gmax, lo, hi)
        return np.array(hist), best

