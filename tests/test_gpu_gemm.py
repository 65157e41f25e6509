"""Products with many right-hand sides (irtt_apply_operator).

FP16 form (MFMA GEMM): unit vectors pick single operator entries, whose FP16
quantisation (row scale = FP32(max off-diagonal |T_ij| * 2^-14), entry =
FP16(FP32(T_ij / scale))) is restated in numpy, so every output element is
checked bit for bit at its position, forward and transposed; random vectors
stay within the FP16 rounding bound |err_i| <= 2^-9 (|T| |x|)_i (two FP16 roundings of 2^-11 each plus FP32 sums).  The FP32
form equals single FP64-accumulated products to rounding."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scene(irtt, n=16):
    from irtt import scenes
    sd = scenes.city_block(seed=4, nx=2, ny=2, target_tris=2500)
    s = irtt.Scene.from_data(sd, irtt.BIN_VERTEX)
    s.build_transport(threads_per_tri=64, rays_per_thread=4, batch_seed=3)
    return s


def _quantised(T):
    """Off-diagonal FP16 entries times their row scale, as the device holds them."""
    off = T.copy()
    np.fill_diagonal(off, 0.0)
    scale = (np.abs(off).max(axis=1) * np.float32(2.0 ** -14)).astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        q = np.where(scale[:, None] > 0, off / scale[:, None], np.float32(0)).astype(np.float32)
    h = q.astype(np.float16).astype(np.float64)
    return h * scale[:, None].astype(np.float64)


@pytest.mark.parametrize("transposed", [False, True])
def test_fp16_gemm_unit_vectors_bit_exact(irtt, transposed):
    with _scene(irtt) as s:
        T = s.transport()
        n = T.shape[0]
        s.select_operator(irtt.OP_DENSE_F16)
        rng = np.random.default_rng(0)
        cols = rng.choice(n, size=70, replace=False)  # two 64-vector chunks, ragged tiles
        X = np.zeros((cols.size, n))
        X[np.arange(cols.size), cols] = 1.0
        Y = s.apply_operator(X, transposed=transposed)
    Hq = _quantised(T)
    d = np.diag(T).astype(np.float64)
    for r, j in enumerate(cols):
        want = (Hq[j, :] if transposed else Hq[:, j]).copy()
        want[j] += d[j]
        np.testing.assert_array_equal(Y[r], want, err_msg=f"vector {r} (unit {j})")


@pytest.mark.parametrize("R", [1, 5, 32, 37, 64, 100])
def test_fp16_gemm_random_vectors_within_rounding(irtt, R):
    with _scene(irtt) as s:
        T = s.transport().astype(np.float64)
        s.select_operator(irtt.OP_DENSE_F16)
        rng = np.random.default_rng(R)
        X = (250.0 + 100.0 * rng.random((R, T.shape[0]))) ** 4
        X[:, ::7] *= -0.5
        Y = s.apply_operator(X)
        Yt = s.apply_operator(X, transposed=True)
    A = np.abs(T)
    for got, want, bound in ((Y, X @ T.T, np.abs(X) @ A.T), (Yt, X @ T, np.abs(X) @ A)):
        err = np.abs(got - want)
        ratio = (err / bound).max()
        print(f"R={R}: max |err| / (|T||x|) = {ratio:.2e}")
        assert ratio < 2.0 ** -9


@pytest.mark.parametrize("R", [3, 37])
def test_fp32_batch_is_an_fp64_product(irtt, R):
    """FP32 form: the FP64 MFMA GEMM (FP32 entries widened exactly, FP64
    products and sums) equals the FP64 product to 1e-13 of sum |T||x|."""
    with _scene(irtt) as s:
        T = s.transport().astype(np.float64)
        rng = np.random.default_rng(R)
        X = rng.normal(size=(R, T.shape[0]))
        Y = s.apply_operator(X)
        Yt = s.apply_operator(X, transposed=True)
        y1 = s.apply_operator(X[R - 1:])
    A = np.abs(T)
    for got, want, bound in ((Y, X @ T.T, np.abs(X) @ A.T), (Yt, X @ T, np.abs(X) @ A)):
        ratio = (np.abs(got - want) / bound).max()
        print(f"R={R}: max |err| / (|T||x|) = {ratio:.2e}")
        assert ratio < 1e-13
    np.testing.assert_array_equal(y1[0], Y[R - 1])  # a vector's product ignores its neighbours
