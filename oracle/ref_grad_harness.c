/*
 * ref_grad_harness.c -- C entry points around the reference's generated
 * gradient headers (test infrastructure only: it builds a checker for the
 * oracle, nothing in the library links it).
 *
 * The reference evaluates two symbolic derivatives inline in its adjoint
 * shader by #including generated plain-arithmetic headers
 * (assets/shader/raytracing_thermal/thermal_tracing_unified.glsl:330-338,
 * 409-421):
 *   symGradArea3D.h     -- dadai, dadbi, dadci = d area(a, b, c) / d{a, b, c}
 *   symGradGeomTerm3D.h -- geomTerm = G and dgdai ... dgdcj = dG / d{six vertices}
 * Both headers are C-compatible expression lists over variables the
 * including scope declares.  This file declares those variables and includes
 * the headers where they lie in the reference tree (the include path is set
 * by the Makefile's `ref` target); nothing of them is copied here.
 * SCALAR is double so the harness evaluates the reference's formulas without
 * FP32 rounding.
 */
#include <math.h>

#define SCALAR double

void ref_area_gradient(const double *a, const double *b, const double *c, double *ga,
                       double *gb, double *gc)
{
    const SCALAR ai_1 = a[0], ai_2 = a[1], ai_3 = a[2];
    const SCALAR bi_1 = b[0], bi_2 = b[1], bi_3 = b[2];
    const SCALAR ci_1 = c[0], ci_2 = c[1], ci_3 = c[2];
    SCALAR dadai[3], dadbi[3], dadci[3];
#include "symGradArea3D.h"
    for (int k = 0; k < 3; k++) {
        ga[k] = dadai[k];
        gb[k] = dadbi[k];
        gc[k] = dadci[k];
    }
}

/* v: 18 doubles (ai bi ci aj bj cj), beta: barycentric weights of corners 1
 * and 2 on each triangle (bi1 bi2 bj1 bj2).  out: 18 doubles dG/dv.
 * Returns G. */
double ref_geom_term_grad(const double *v, const double *beta, double tinyEps, double *out)
{
    const SCALAR ai_1 = v[0], ai_2 = v[1], ai_3 = v[2];
    const SCALAR bi_1 = v[3], bi_2 = v[4], bi_3 = v[5];
    const SCALAR ci_1 = v[6], ci_2 = v[7], ci_3 = v[8];
    const SCALAR aj_1 = v[9], aj_2 = v[10], aj_3 = v[11];
    const SCALAR bj_1 = v[12], bj_2 = v[13], bj_3 = v[14];
    const SCALAR cj_1 = v[15], cj_2 = v[16], cj_3 = v[17];
    const SCALAR beta_i_1 = beta[0], beta_i_2 = beta[1];
    const SCALAR beta_j_1 = beta[2], beta_j_2 = beta[3];
    SCALAR geomTerm = 0.0;
    SCALAR dgdai[3], dgdbi[3], dgdci[3], dgdaj[3], dgdbj[3], dgdcj[3];
#include "symGradGeomTerm3D.h"
    for (int k = 0; k < 3; k++) {
        out[k] = dgdai[k];
        out[3 + k] = dgdbi[k];
        out[6 + k] = dgdci[k];
        out[9 + k] = dgdaj[k];
        out[12 + k] = dgdbj[k];
        out[15 + k] = dgdcj[k];
    }
    return geomTerm;
}
