"""CPU: the per-sample field generator of the dense workload (synth.field_pair_torch)
and the truth it defines (synth.window_truth).

* A constant field reproduces the exact Fourier shift of the master (the sign
  convention aligned(x) = master(x - D)) to within the windowed-sinc kernel's
  stated error power.
* On a smooth varying field the FP64 oracle's offsets match the template-mean
  displacement window_truth reports.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


def test_constant_field_matches_fourier_shift(synth):
    f = lambda x, y: (0 * x + 2.37, 0 * x - 1.61)
    c = lambda x, y: 0 * x + 1.0
    m, s = synth.field_pair_torch(384, 384, f, c, device="cpu", rows=128)
    mc = m[..., 0].double().numpy() + 1j * m[..., 1].double().numpy()
    sc = s[..., 0].double().numpy() + 1j * s[..., 1].double().numpy()
    ref = synth.fourier_shift(mc, 2.37, -1.61)
    inner = (slice(80, 304), slice(80, 304))
    rel = np.mean(np.abs(sc[inner] - ref[inner]) ** 2) / np.mean(np.abs(ref[inner]) ** 2)
    assert rel < 0.045
    # the opposite shift is far off: the sign convention is pinned
    bad = synth.fourier_shift(mc, -2.37, 1.61)
    assert np.mean(np.abs(sc[inner] - bad[inner]) ** 2) / np.mean(np.abs(bad[inner]) ** 2) > 1.0


def test_window_truth_matches_oracle_offsets(xg, orc, synth):
    a, lam = 4.0, 900.0
    mod = lambda v: np if isinstance(v, np.ndarray) else torch
    f = lambda x, y: (a * mod(x).sin(2 * np.pi * x / lam), -0.5 * a * mod(y).cos(2 * np.pi * y / lam))
    c = lambda x, y: 0 * x + 0.97
    nx = ny = 768
    m, s = synth.field_pair_torch(nx, ny, f, c, device="cpu", rows=128)
    mh, sh = m.numpy(), s.numpy()
    xs = 16
    x, y = xg.get_locations(nx, ny, 9, 9, xs, xs)
    w = xg.make_windows(x, y, 4 * xs, 4 * xs)
    locs = np.stack([w["mx0"], w["my0"], w["sx0"], w["sy0"]], 1).astype(np.int32)
    res = orc.run(orc.make_params(xs, xs, interp_factor=16), mh, sh, locs, precision=64)
    dx, dy = xg.offsets(res, 2)
    tx, ty, sx, sy = synth.window_truth(f, x.astype(np.float64), y.astype(np.float64), xs)
    ok = res["corr"] > 50
    assert ok.sum() > 0.8 * len(res)
    ex, ey = dx - tx, dy - ty
    print("median |error| dx %.3f dy %.3f px over %d windows (field spread up to %.2f px)"
          % (np.median(np.abs(ex[ok])), np.median(np.abs(ey[ok])), ok.sum(), max(sx.max(), sy.max())))
    assert np.median(np.abs(ex[ok])) < 0.1 and np.median(np.abs(ey[ok])) < 0.15
    # the window centre's value differs from the truth by the field's variation
    assert np.ptp(tx) > 4 and np.ptp(ty) > 2
