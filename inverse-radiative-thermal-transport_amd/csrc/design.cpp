// design.cpp -- the gradient of an objective with respect to design
// parameters at a solved state: adjoint solve, adjoint trace, the source
// term's geometry derivative, projection onto the parameters.
//
// The reference runs these steps inside ThermalRenderer::runAdjointSim and
// Parametrization::projectGradient
// (src/implementations/thermal_renderer/thermal_renderer.cpp:1028-1133,
//  thermal_params.cpp:130-156); its sun and sky are fixed-temperature
// emitters, so their geometry dependence is part of T.  Here an absorbed
// flux per triangle is a source term b = flux * area, whose derivative
// lambda^T db/dv is added on the host.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "irtt_internal.hpp"

namespace irtt {

// grad += lambda^T db/dv, b_bin = sum over the bin's triangles of
// flux_t A_t (triangle bins) or flux_t A_t / 3 (vertex bins).
// dA/da = (b - c) x n / 2, dA/db = (c - a) x n / 2, dA/dc = (a - b) x n / 2
// with n the unit normal of (b - a) x (c - a); a clamped (degenerate) area
// has no derivative.
void source_gradient(const Scene &s, const double *lam, double *grad)
{
    if (s.source_explicit) return;
    for (uint32_t t = 0; t < s.nt; t++) {
        const double flux = (double)s.mat[t].heat_flux;
        if (flux == 0.0) continue;
        const uint32_t iv[3] = {s.idx[3 * t], s.idx[3 * t + 1], s.idx[3 * t + 2]};
        const double w = s.bin_mode == IRTT_BIN_TRIANGLE
                             ? lam[t]
                             : (lam[iv[0]] + lam[iv[1]] + lam[iv[2]]) / 3.0;
        if (w == 0.0) continue;
        double P[3][3];
        for (int c = 0; c < 3; c++)
            for (int k = 0; k < 3; k++) P[c][k] = (double)s.xyz[3 * iv[c] + k];
        double e1[3], e2[3];
        for (int k = 0; k < 3; k++) {
            e1[k] = P[1][k] - P[0][k];
            e2[k] = P[2][k] - P[0][k];
        }
        double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                       e1[0] * e2[1] - e1[1] * e2[0]};
        const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (!(0.5 * len >= 10.0 * FLT_EPSILON)) continue;
        for (int k = 0; k < 3; k++) n[k] /= len;
        for (int c = 0; c < 3; c++) {
            const double *p = P[(c + 1) % 3], *q = P[(c + 2) % 3];
            const double d[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
            const double g[3] = {d[1] * n[2] - d[2] * n[1], d[2] * n[0] - d[0] * n[2],
                                 d[0] * n[1] - d[1] * n[0]};
            double *o = grad + 3 * (size_t)iv[c];
            for (int k = 0; k < 3; k++) o[k] += flux * w * 0.5 * g[k];
        }
    }
}

void design_gradient(Scene &s, const irtt_solver_params &p, const double *dLdx,
                     const irtt_adjoint_params &ap, uint32_t nnz, const uint32_t *vertex,
                     const uint32_t *param, const float *dxdp, uint32_t n_params, double *dp,
                     double *grad_out, double *lam_out, irtt_solve_stats &st)
{
    const uint32_t nb = s.nb;
    std::vector<double> x = s.x, lam(nb);
    adjoint_solve(s, p, x.data(), dLdx, lam.data(), st);
    gradient_from_adjoint(s, p, x.data(), lam.data(), ap, nnz, vertex, param, dxdp, n_params, dp,
                          grad_out);
    if (lam_out) std::copy(lam.begin(), lam.end(), lam_out);
}

void gradient_from_adjoint(Scene &s, const irtt_solver_params &p, const double *x,
                           const double *lam, const irtt_adjoint_params &ap, uint32_t nnz,
                           const uint32_t *vertex, const uint32_t *param, const float *dxdp,
                           uint32_t n_params, double *dp, double *grad_out)
{
    const uint32_t nb = s.nb, nv = s.nv;
    // history rows: (s lambda, x^4) for the operator term T x^4, and
    // (-s Ta^4, lambda) for the ambient term Ta^4 colsum(T)
    const double sc = p.steady ? 1.0 : (p.dt > 0.0 ? p.dt : 300.0);
    const bool amb = p.ambient_kelvin >= 0.0;
    const uint32_t rows = amb ? 2u : 1u;
    std::vector<double> lr((size_t)rows * nb), xr((size_t)rows * nb);
    const double Ta = p.ambient_kelvin, Ta4 = Ta * Ta * Ta * Ta;
    for (uint32_t i = 0; i < nb; i++) {
        const double x2 = x[i] * x[i];
        lr[i] = sc * lam[i];
        xr[i] = x2 * x2;
        if (amb) {
            lr[nb + i] = -sc * Ta4;
            xr[nb + i] = lam[i];
        }
    }
    std::vector<double> grad(3 * (size_t)nv);
    irtt_adjoint_stats ast{};
    adjoint_trace(s, ap, rows, lr.data(), xr.data(), grad.data(), ast);
    source_gradient(s, lr.data(), grad.data());
    if (dp) {
        for (uint32_t k = 0; k < n_params; k++) dp[k] = 0.0;
        int rc = irtt_project_gradient(grad.data(), nv, nnz, vertex, param, dxdp, n_params, dp);
        if (rc) throw Error(rc, "projection: " + std::string(irtt_last_error()));
    }
    if (grad_out) std::copy(grad.begin(), grad.end(), grad_out);
}

}  // namespace irtt
