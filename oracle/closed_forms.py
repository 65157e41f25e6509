"""Closed forms that pin the oracle (test infrastructure only).

* ``parallel_square_view_factor`` -- view factor between two coaxial parallel
  rectangles a x b at distance c (Hottel; Howell's catalogue of radiation
  configuration factors, configuration C-11).  0.19982 for unit squares at
  unit distance.
* ``pt_emission_fraction`` -- photon fraction bounds used by the tests.
"""
import math


def parallel_rect_view_factor(a, b, c):
    """F_12 between directly opposed parallel rectangles a x b separated by c."""
    X = a / c
    Y = b / c
    t1 = math.log(math.sqrt((1 + X * X) * (1 + Y * Y) / (1 + X * X + Y * Y)))
    t2 = X * math.sqrt(1 + Y * Y) * math.atan(X / math.sqrt(1 + Y * Y))
    t3 = Y * math.sqrt(1 + X * X) * math.atan(Y / math.sqrt(1 + X * X))
    t4 = X * math.atan(X) + Y * math.atan(Y)
    return 2.0 / (math.pi * X * Y) * (t1 + t2 + t3 - t4)


def two_plate_absorbed_fraction(F, eps_a, eps_b, bounces=64):
    """Fraction of the power emitted by plate A that plate B absorbs when both
    plates are gray diffuse (reflectance 1-eps) and radiosity is uniform over
    each plate (the approximation behind the Gebhart factor):
    G_AB = eps_b F sum_k ((1-eps_a)(1-eps_b) F^2)^k = eps_b F / (1 - rho_a rho_b F^2)."""
    ra, rb = 1.0 - eps_a, 1.0 - eps_b
    return eps_b * F / (1.0 - ra * rb * F * F)
