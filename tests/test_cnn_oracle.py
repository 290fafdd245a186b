"""Pins of the NumPy restatement of the proposal network (oracle/cnn_oracle.py):
bf16 rounding against its definition, convolution against a direct loop and
scipy's correlation, analytic gradients against central finite differences,
and the input preparation against closed forms."""
import numpy as np
import pytest

import cnn_oracle as co


def test_bf16_round_to_nearest_even():
    one = 1.0
    vals = np.array([one, one + 2.0 ** -8, one + 3 * 2.0 ** -8, one + 2.0 ** -7, -2.5, 0.0, 65504.0,
                     1.0 + 2.0 ** -8 + 2.0 ** -20], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -6, 1.0 + 2.0 ** -7, -2.5, 0.0, 65536.0, 1.0 + 2.0 ** -7])
    assert np.array_equal(co.bf16(vals), want)
    assert np.isnan(co.bf16(np.float32(np.nan)))


def _direct_conv(x, W, b, s):
    B, H, Wd, C = x.shape
    ho, wo = (H - 1) // s + 1, (Wd - 1) // s + 1
    z = np.zeros((B, ho, wo, W.shape[1]))
    for n in range(B):
        for oy in range(ho):
            for ox in range(wo):
                for o in range(W.shape[1]):
                    acc = b[o]
                    for ky in range(3):
                        for kx in range(3):
                            iy, ix = oy * s + ky - 1, ox * s + kx - 1
                            if 0 <= iy < H and 0 <= ix < Wd:
                                acc += W[3 * ky + kx, o, :] @ x[n, iy, ix, :]
                    z[n, oy, ox, o] = acc
    return z


@pytest.mark.parametrize("s", [1, 2])
def test_conv_matches_direct_loop(s):
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 7, 6, 3))
    W = rng.standard_normal((9, 5, 3))
    b = rng.standard_normal(5)
    np.testing.assert_allclose(co.conv(x, W, b, s), _direct_conv(x, W, b, s), rtol=1e-12, atol=1e-12)


def test_conv_matches_scipy_correlate():
    from scipy.signal import correlate2d
    rng = np.random.default_rng(4)
    x = rng.standard_normal((1, 9, 11, 2))
    W = rng.standard_normal((9, 3, 2))
    z = co.conv(x, W, np.zeros(3), 1)
    for o in range(3):
        ref = sum(correlate2d(x[0, :, :, c], W[:, o, c].reshape(3, 3), mode="same") for c in range(2))
        np.testing.assert_allclose(z[0, :, :, o], ref, rtol=1e-12, atol=1e-12)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(5)
    params = co.init_params(7).astype(np.float64)
    params += rng.uniform(-0.02, 0.02, params.size)          # non-zero biases
    x0 = rng.standard_normal((2, 64, 64, 2))
    tgt = np.array([[0.7, -1.9, 1.0, 1.0, 0.5], [-2.5, 0.3, 0.0, 0.8, 1.0]])
    loss, g, _ = co.loss_and_grads(params, x0, tgt, rounding=False)
    idx = np.concatenate([rng.choice(co.NPARAM - 99, 40, replace=False), np.arange(co.NPARAM - 99, co.NPARAM)])
    h = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        num = (co.loss_and_grads(p1, x0, tgt, rounding=False)[0] - co.loss_and_grads(p2, x0, tgt, rounding=False)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 + 1e-4 * abs(num), (i, num, g[i])


def test_pack_unpack_roundtrip():
    p = co.init_params(1)
    assert p.size == co.NPARAM == 28451
    assert np.array_equal(co.pack(co.unpack(p)), p.astype(np.float64))


def test_box4_and_crops_closed_forms():
    ramp = np.add.outer(np.arange(16.0), 10 * np.arange(24.0))
    d = co.box4(ramp)
    assert d.shape == (4, 6)
    np.testing.assert_allclose(d, np.add.outer(4 * np.arange(4) + 1.5, 10 * (4 * np.arange(6) + 1.5)))
    const = np.full((80, 90), 3.0)
    c = co.crops(const, const, [(0, 0, 0, 0)])
    assert np.all(c == 0.0)                                   # zero variance: all zero
    rng = np.random.default_rng(2)
    img = rng.uniform(1, 2, (100, 120))
    c = co.crops(img, img, [(10, 20, -5, 70)])
    np.testing.assert_allclose(c[0, :, :, 0].mean(), 0.0, atol=1e-12)
    np.testing.assert_allclose(c[0, :, :, 0].std(), 1.0, rtol=1e-12)
    assert np.all(c[0, :, :5, 1] == c[0, 0, 0, 1])            # outside the image: one (normalised) value
