"""Pins the oracle's gradient formulas against the reference's own generated
headers: symGradArea3D.h (area gradient) and symGradGeomTerm3D.h (the
geometric term G and dG/dv), compiled where they lie by ``make -C oracle ref``
into ``oracle/_ref`` (``oracle/ref_grad_harness.c`` declares the variables
they read and write).  The oracle evaluates in FP32 and writes d ln|G|/dv;
the headers are evaluated in FP64, and dG/dv / G is compared.  Bounds: G to
1e-5 relative, each segment's gradient to 1e-5 of its largest component,
the area gradient to 1e-5 of its largest component (measured: below 1e-6)."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def ref(oracle):
    if oracle.ref_lib() is None:
        pytest.skip("reference headers not available (no oracle/_ref build, no reference tree)")
    return oracle


def _segment(rng, facing=True):
    tri_i = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float) + rng.normal(size=(3, 3)) * 0.2
    tri_j = np.array([[0, 0, 2], [0, 1, 2], [1, 0, 2]], float) + rng.normal(size=(3, 3)) * 0.2
    tri_j += rng.normal(size=3) * 0.5
    if not facing:
        tri_j = tri_j[[0, 2, 1]]
    v = np.concatenate([tri_i, tri_j]).astype(np.float32)
    bi = rng.dirichlet([1, 1, 1])[1:].astype(np.float32)
    bj = rng.dirichlet([1, 1, 1])[1:].astype(np.float32)
    return v, bi, bj


@pytest.mark.parametrize("eps", [1e-2, 0.0, 0.3])
def test_geom_term_matches_reference_header(ref, eps):
    rng = np.random.default_rng(17)
    worst_g = worst_d = 0.0
    for k in range(300):
        v, bi, bj = _segment(rng, facing=(k % 3 != 0))
        G, dln = ref.geom_term_grad(v, bi, bj, eps)
        Gr, dG = ref.ref_geom_term_grad(v.astype(np.float64), bi, bj, eps)
        assert Gr != 0.0
        eg = abs(G - Gr) / abs(Gr)
        want = dG / Gr
        ed = np.abs(dln - want).max() / np.abs(want).max()
        worst_g, worst_d = max(worst_g, eg), max(worst_d, ed)
        assert eg < 1e-5, (k, G, Gr)
        assert ed < 1e-5, (k, dln, want)
    print(f"eps={eps}: max rel. error G {worst_g:.2e}, d ln|G|/dv {worst_d:.2e} over 300 segments")


def test_area_gradient_matches_reference_header(ref):
    rng = np.random.default_rng(18)
    worst = 0.0
    for _ in range(300):
        t = rng.normal(size=(3, 3)).astype(np.float32)
        g = np.array(ref.area_gradient(*t), np.float64)
        gr = np.array(ref.ref_area_gradient(*t.astype(np.float64)))
        e = np.abs(g - gr).max() / np.abs(gr).max()
        worst = max(worst, e)
        assert e < 1e-5
    print(f"max rel. error of the area gradient {worst:.2e} over 300 triangles")


def test_reference_header_is_a_true_derivative(ref):
    """The harness itself: the header's dG/dv against FP64 central
    differences of its own G (guards against a wrong variable binding)."""
    rng = np.random.default_rng(19)
    for _ in range(20):
        v, bi, bj = _segment(rng)
        v = v.astype(np.float64)
        _, dG = ref.ref_geom_term_grad(v, bi, bj, 1e-2)
        for k in range(6):
            for c in range(3):
                e = np.zeros((6, 3))
                e[k, c] = 1e-6
                gp, _ = ref.ref_geom_term_grad(v + e, bi, bj, 1e-2)
                gm, _ = ref.ref_geom_term_grad(v - e, bi, bj, 1e-2)
                fd = (gp - gm) / 2e-6
                assert abs(fd - dG[k, c]) <= 1e-5 * (1 + np.abs(dG).max())
