// history.cpp -- forward and adjoint drivers over time steps.
//
// Replace ThermalRenderer::runForwardSim and runAdjointSim
// (src/implementations/thermal_renderer/thermal_renderer.cpp:958-1133): the
// forward driver keeps every step's temperatures, the adjoint driver walks
// them backwards with the transposed Jacobian of each step and hands the
// history to the adjoint photon trace.  The history is indexed with the bin
// count as stride throughout (the reference's shader reads it with the
// history length as stride, correct only for one step).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "irtt_internal.hpp"

namespace irtt {

void simulate_history(Scene &s, const irtt_solver_params &p, uint32_t steps, double *history,
                      irtt_solve_stats &st)
{
    if (steps == 0) throw Error(IRTT_ERR_ARG, "steps must be positive");
    if (p.steady && steps != 1) throw Error(IRTT_ERR_ARG, "a steady solve is one step");
    const uint32_t nb = s.nb;
    st = irtt_solve_stats{};
    for (uint32_t t = 0; t < steps; t++) {
        irtt_solve_stats one{};
        solve(s, p, one);
        std::memcpy(history + (size_t)t * nb, s.x.data(), nb * sizeof(double));
        st.newton_iters += one.newton_iters;
        st.linear_iters += one.linear_iters;
        st.line_search_halvings += one.line_search_halvings;
        st.residual_sq = std::max(st.residual_sq, one.residual_sq);
        st.max_dx = std::max(st.max_dx, one.max_dx);
        st.solve_ms += one.solve_ms;
        st.converged = (t == 0 ? 1u : st.converged) & one.converged;
    }
}

void adjoint_history(Scene &s, const irtt_solver_params &p, uint32_t steps, const double *hist,
                     const double *w, const double *targets, uint32_t target_stride,
                     double *lam_hist, double *objective, const irtt_adjoint_params *ap,
                     double *grad_x, irtt_solve_stats &st)
{
    if (steps == 0) throw Error(IRTT_ERR_ARG, "steps must be positive");
    if (p.steady && steps != 1) throw Error(IRTT_ERR_ARG, "a steady history has one step");
    const uint32_t nb = s.nb;
    const double dt = p.dt > 0.0 ? p.dt : 300.0;
    const double scale = p.steady ? 1.0 : 1.0 / (double)steps;
    std::vector<double> g(nb), src(nb), lam(nb), lam_next(nb, 0.0);
    double L = 0.0;
    st = irtt_solve_stats{};
    for (uint32_t t = steps; t-- > 0;) {
        const double *x = hist + (size_t)t * nb;
        const double *tg = targets + (size_t)t * target_stride;
        double obj = 0.0;
        for (uint32_t i = 0; i < nb; i++) {
            double d = x[i] - tg[i];
            g[i] = w[i] * d;
            obj += 0.5 * w[i] * d * d;
        }
        L += scale * obj;
        for (uint32_t i = 0; i < nb; i++) {
            double v = p.steady ? g[i] : dt * g[i] + s.bin_mass[i] * lam_next[i];
            src[i] = s.bin_fixed[i] ? 0.0 : v;
        }
        irtt_solve_stats one{};
        adjoint_solve(s, p, x, src.data(), lam.data(), one);
        st.linear_iters += one.linear_iters;
        st.solve_ms += one.solve_ms;
        st.converged = (t == steps - 1 ? 1u : st.converged) & one.converged;
        for (uint32_t i = 0; i < nb; i++) lam_hist[(size_t)t * nb + i] = scale * lam[i];
        lam_next = lam;
    }
    if (objective) *objective = L;
    if (!grad_x) return;
    // history rows for the trace: (lambda_t, x_t^4), and the ambient term
    // Ta^4 colsum(T) of every step as one row (-Ta^4, sum_t lambda_t)
    const bool amb = p.ambient_kelvin >= 0.0;
    const uint32_t rows = steps + (amb ? 1u : 0u);
    std::vector<double> lr((size_t)rows * nb, 0.0), xr((size_t)rows * nb, 0.0);
    for (uint32_t t = 0; t < steps; t++)
        for (uint32_t i = 0; i < nb; i++) {
            const double xv = hist[(size_t)t * nb + i], x2 = xv * xv;
            lr[(size_t)t * nb + i] = lam_hist[(size_t)t * nb + i];
            xr[(size_t)t * nb + i] = x2 * x2;
        }
    if (amb) {
        const double Ta = p.ambient_kelvin, Ta4 = Ta * Ta * Ta * Ta;
        for (uint32_t i = 0; i < nb; i++) {
            lr[(size_t)steps * nb + i] = -Ta4;
            double sum = 0.0;
            for (uint32_t t = 0; t < steps; t++) sum += lam_hist[(size_t)t * nb + i];
            xr[(size_t)steps * nb + i] = sum;
        }
    }
    irtt_adjoint_params a = ap ? *ap : irtt_adjoint_params{0.05f, 1e-2f, 0u, 0u};
    irtt_adjoint_stats ast{};
    adjoint_trace(s, a, rows, lr.data(), xr.data(), grad_x, ast);
    // the source term b (the same every step): sum_t lambda_t^T db/dv
    std::vector<double> lsum(nb, 0.0);
    for (uint32_t t = 0; t < steps; t++)
        for (uint32_t i = 0; i < nb; i++) lsum[i] += lam_hist[(size_t)t * nb + i];
    source_gradient(s, lsum.data(), grad_x);
}

}  // namespace irtt
