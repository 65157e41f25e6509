"""FP64 restatement of the energy-balance solve (test infrastructure only).

Follows ``src/implementations/thermal_renderer/thermal_solver.cpp``:
residual (:13-37), Jacobian (:39-96), damped Newton (:119-229), with the
library's extensions (absorbed source ``b`` and ambient exchange through the
operator's column sums).  The linear solves are dense LU (numpy), so the
result is the FP64 fixed point of the same discrete system, independent of
any Krylov iteration.

Pinned by: the single-element closed form x = (q / (sigma eps A))^(1/4)
(+ ambient), and for steady state without conduction the exact linear solve
in y = x^4 (``steady_linear``), which the Newton result must reproduce.
"""
import numpy as np


def rhs(b_source, colsum, ambient_kelvin):
    b = np.asarray(b_source, np.float64).copy()
    if ambient_kelvin is not None and ambient_kelvin >= 0:
        b += ambient_kelvin ** 4 * np.asarray(colsum, np.float64)
    return b


def residual(T, x, b, fixed, kfix, lap=None, mass=None, xprev=None, dt=300.0, steady=True):
    x = np.where(fixed, kfix, x)
    rad = T @ (x ** 4) - b
    if lap is not None:
        rad = rad + lap @ x
    r = rad if steady else mass * (x - xprev) + dt * rad
    return np.where(fixed, 0.0, r)


def jacobian(T, x, fixed, lap=None, mass=None, dt=300.0, steady=True):
    """Exact Jacobian T diag(4x^3) + L (+ M / dt scaling), fixed rows -> identity."""
    J = T * (4.0 * x ** 3)[None, :]
    if lap is not None:
        J = J + (lap.toarray() if hasattr(lap, "toarray") else lap)
    if not steady:
        J = dt * J + np.diag(mass)
    J[fixed, :] = 0.0
    J[:, fixed] = 0.0
    J[fixed, fixed] = 1.0
    return J


def newton(T, b, fixed, kfix, x0, lap=None, mass=None, xprev=None, dt=300.0, steady=True,
           tol=1e-10, max_iter=100):
    """Damped Newton with exact dense LU steps, thermal_solver.cpp:119-229."""
    T = np.asarray(T, np.float64)
    fixed = np.asarray(fixed, bool)
    x = np.where(fixed, kfix, np.asarray(x0, np.float64))
    if steady:
        x = np.where((~fixed) & (x == 0.0), 1.0, x)
    r = residual(T, x, b, fixed, kfix, lap, mass, xprev, dt, steady)
    rn = r @ r
    for it in range(max_iter):
        J = jacobian(T, x, fixed, lap, mass, dt, steady)
        dx = -np.linalg.solve(J, r)
        a = 1.0
        while True:
            xn = x + a * dx
            rnew = residual(T, xn, b, fixed, kfix, lap, mass, xprev, dt, steady)
            rnn = rnew @ rnew
            if (rnn <= rn and np.isfinite(rnn)) or a < 1e-6:
                break
            a *= 0.5
        x, r, rn = xn, rnew, rnn
        if np.max(np.abs(a * dx)) < tol:
            break
    return np.where(fixed, kfix, x), it + 1


def steady_linear(T, b, fixed, kfix):
    """Steady state without conduction: T y = b is linear in y = x^4.
    Free rows: T_ff y_f = b_f - T_fx y_x."""
    T = np.asarray(T, np.float64)
    fixed = np.asarray(fixed, bool)
    f = ~fixed
    yx = np.asarray(kfix, np.float64)[fixed] ** 4
    rhs_ = np.asarray(b, np.float64)[f] - T[np.ix_(f, fixed)] @ yx
    yf = np.linalg.solve(T[np.ix_(f, f)], rhs_)
    x = np.asarray(kfix, np.float64).copy()
    x[f] = np.sign(yf) * np.abs(yf) ** 0.25
    return x
