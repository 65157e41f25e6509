"""Design loop over a parametrised scene: objective, gradient, finite
differences and gradient descent, each step through the library.

The reference runs the same loop inside ThermalRenderer
(src/implementations/thermal_renderer/thermal_renderer.cpp:1257-1306:
forward simulation, adjoint, projectGradient, parameter update, geometry
update; thermal_params.cpp:101-156, thermal_objective.cpp:19-78).  Here one
step is: irtt_scene_update_vertices -> LBVH -> photon transport (same seed:
common random numbers between steps and finite-difference evaluations) ->
Newton solve -> irtt_design_gradient (adjoint solve, adjoint trace, source
term, projection).
"""
import time

import numpy as np

from . import scenes as _scenes


def ground_weights(s, sd):
    """dL/dx of the mean ground temperature L = sum_i A_i x_i / sum_i A_i over
    the ground bins (a linear objective; the reference's objective is the
    weighted squared deviation, thermal_objective.cpp:47-57)."""
    area = s.bin_data()[0]
    g = np.asarray(sd.meta["ground_vertex"], bool)
    w = np.where(g, area, 0.0)
    return w / w.sum()


class FacadeDesign:
    """A facade scene on the device with its extrusion depths as the design
    parameters (scenes.facade)."""

    def __init__(self, irtt, sd, transport, solver, x0=None, adjoint_flags=0):
        self.irtt = irtt
        self.sd = sd
        self.tp = transport
        self.sp = solver
        # 0: the reference's terms (area | continuous | edge)
        self.flags = adjoint_flags or (irtt.ADJ_AREA | irtt.ADJ_CONTINUOUS | irtt.ADJ_EDGE)
        self.s = irtt.Scene.from_data(sd, irtt.BIN_VERTEX)
        self.x0 = np.asarray(sd.kelvin if x0 is None else x0, np.float32)
        self.w = ground_weights(self.s, sd)
        self.par = _scenes.facade_parametrization(sd)
        self.depths = np.asarray(sd.meta["depths"], np.float64).copy()
        self.timing = {}

    def close(self):
        self.s.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def forward(self, depths):
        """Objective at the given depths (geometry, LBVH, transport, solve)."""
        t0 = time.perf_counter()
        self.s.update_vertices(_scenes.facade_vertices(self.sd, depths))
        tst = self.s.build_transport(self.tp)
        t1 = time.perf_counter()
        self.s.set_vertex_state(self.x0, self.sd.fixed)
        sst = self.s.solve(self.sp)
        x = self.s.temperatures()
        t2 = time.perf_counter()
        self.timing = {"transport_ms": (t1 - t0) * 1e3, "solve_ms": (t2 - t1) * 1e3,
                       "nnz": tst["nnz"], "rays": tst["rays"], "converged": sst["converged"]}
        return float(self.w @ x), x

    def gradient(self, flags=None):
        """dL/d(depths) at the state of the last forward()."""
        t0 = time.perf_counter()
        dp, gx, lam, st = self.s.design_gradient(self.w, params=self.par, solver=self.sp,
                                                 flags=flags or self.flags)
        self.timing["gradient_ms"] = (time.perf_counter() - t0) * 1e3
        return dp

    def central_difference(self, depths, k, h):
        d = np.array(depths, np.float64)
        d[k] += h
        lp, _ = self.forward(d)
        d[k] -= 2 * h
        lm, _ = self.forward(d)
        return (lp - lm) / (2 * h)

    def descend(self, steps, max_step=0.05, lo=0.0, hi=1.5, log=None):
        """Normalised gradient descent on the depths (each step moves the
        largest-gradient depth by max_step), clipped to [lo, hi].  Returns the
        objective history and the best depths."""
        d = self.depths.copy()
        hist, best = [], (np.inf, d.copy())
        for k in range(steps):
            L, _ = self.forward(d)
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
