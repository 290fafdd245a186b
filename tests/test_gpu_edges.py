"""GPU: edge cases of the C ABI, each against the FP32 CPU restatement.

* An empty call (n = 0) succeeds and writes nothing, on every path and
  pointer mode.
* One location per batch (max_batch 1) on the strip and per-patch paths.
* The largest window the parameters allow (xsearch = ysearch = 256: 1024 x
  1024 windows, 512 x 512 templates), the largest range-interpolation
  factors (ri 8 and 16) and the smallest high-res factor (1: an 8 x 8 grid,
  the paired column sums with rows 0..4).
* All-zero images: every window is zero, the surface has no positive value,
  and the result is the defined one (peak 0, corr 0), as in the restatement.
"""
import numpy as np
import pytest

from conftest import gpu_present

pytestmark = pytest.mark.gpu

FIELDS = ("xoff", "yoff", "corr", "xfrac", "yfrac", "cmax", "ip", "jp", "hip", "hjp")


@pytest.fixture(scope="module")
def pair(synth):
    if not gpu_present():
        pytest.skip("no GPU")
    return synth.constant_shift_pair(1536, 2048, -2.6, 4.1, 0.85, seed=51, noise_seed=52)


def _locs(w):
    return np.stack([w["mx0"], w["my0"], w["sx0"], w["sy0"]], 1).astype(np.int32)


def _oracle_params(orc, p):
    return orc.make_params(p.xsearch, p.ysearch, ri=p.ri, interp_flag=p.interp_flag,
                           interp_factor=p.interp_factor, corr_flag=p.corr_flag)


def _same(got, ref, label):
    for f in FIELDS:
        a, b = got[f], ref[f]
        eq = (a.view(np.uint32) == b.view(np.uint32)) if a.dtype.kind == "f" else (a == b)
        bad = np.nonzero(~eq)[0]
        assert bad.size == 0, "%s: %s differs at %s: %s vs %s" % (label, f, bad[:4], a[bad[:4]], b[bad[:4]])
    np.testing.assert_allclose(got["cave"], ref["cave"], rtol=1e-12, err_msg=label)


def _windows(xg, n, npx, npy, seed, m_nx=1536, m_ny=2048):
    rng = np.random.default_rng(seed)
    w = np.zeros(n, xg.WINDOW_DTYPE)
    w["mx0"] = rng.integers(0, m_nx - npx + 1, n)
    w["my0"] = rng.integers(0, m_ny - npy + 1, n)
    w["sx0"], w["sy0"] = w["mx0"], w["my0"]
    return w


@pytest.mark.parametrize("xs,corr", [(64, 2), (16, 2), (32, 0), (64, 0), (32, 1), (64, 1), (16, 1)],
                         ids=["strips", "per_patch", "time", "time256", "time4", "time4_256", "time4_per_patch"])
def test_empty_call(xg, pair, xs, corr):
    import torch
    m, s = pair
    p = xg.make_params(xs, xs, corr_flag=corr)
    w = np.zeros(0, xg.WINDOW_DTYPE)
    with xg.Engine(p, max_batch=8) as eng:
        eng.set_images(m, s)
        assert len(eng.correlate(w)) == 0
        sentinel = torch.full((4 * xg.RESULT_DTYPE.itemsize,), 0x5a, dtype=torch.uint8, device="cuda")
        eng.correlate(w, out=sentinel, flags=xg.NO_SYNC)
        dw = torch.zeros((0, 4), dtype=torch.int32, device="cuda")
        eng.correlate(dw, out=sentinel)
        eng.sync()
        assert bool((sentinel == 0x5a).all())


@pytest.mark.parametrize("xs", [64, 16])
def test_one_location_per_batch(xg, orc, pair, xs):
    m, s = pair
    p = xg.make_params(xs, xs)
    w = _windows(xg, 5, p.npx, p.npy, seed=xs)
    with xg.Engine(p, max_batch=1) as eng:
        eng.set_images(m, s)
        got = eng.correlate(w)
    _same(got, orc.run(_oracle_params(orc, p), m, s, _locs(w), precision=32), "max_batch 1, xs %d" % xs)


def test_largest_window(xg, orc, pair):
    """xsearch = ysearch = 256 (the limit of xcorr_hip_params): 1024 x 1024
    windows through the generic LDS kernels; one window inside the images and
    one partly outside."""
    m, s = pair
    p = xg.make_params(256, 256, interp_factor=8)
    w = np.zeros(2, xg.WINDOW_DTYPE)
    w["mx0"], w["my0"] = [200, -300], [400, 1500]
    w["sx0"], w["sy0"] = w["mx0"] + [-3, 2], w["my0"] + [4, -1]
    with xg.Engine(p) as eng:
        eng.set_images(m, s)
        got = eng.correlate(w)
    _same(got, orc.run(_oracle_params(orc, p), m, s, _locs(w), precision=32), "1024 x 1024 windows")
    assert got["corr"][0] > 50


@pytest.mark.parametrize("ri", [8, 16])
def test_largest_range_interpolation(xg, orc, pair, ri):
    m, s = pair
    p = xg.make_params(32, 32, ri=ri)
    w = _windows(xg, 6, p.npx, p.npy, seed=ri)
    with xg.Engine(p, max_batch=4) as eng:
        eng.set_images(m, s)
        got = eng.correlate(w)
    _same(got, orc.run(_oracle_params(orc, p), m, s, _locs(w), precision=32), "ri %d" % ri)


@pytest.mark.parametrize("xs,corr", [(64, 2), (16, 2), (32, 0), (64, 0), (32, 1), (64, 1)])
def test_high_res_factor_one(xg, orc, pair, xs, corr):
    m, s = pair
    p = xg.make_params(xs, xs, interp_factor=1, corr_flag=corr)
    w = _windows(xg, 12, p.npx, p.npy, seed=70 + xs)
    with xg.Engine(p, max_batch=8) as eng:
        eng.set_images(m, s)
        got = eng.correlate(w)
    _same(got, orc.run(_oracle_params(orc, p), m, s, _locs(w), precision=32), "interp 1, xs %d" % xs)


@pytest.mark.parametrize("xs,corr", [(64, 2), (32, 2), (16, 2), (32, 0), (64, 0), (32, 1), (64, 1)])
def test_zero_images(xg, orc, xs, corr):
    if not gpu_present():
        pytest.skip("no GPU")
    z = np.zeros((768, 768, 2), np.int16)
    p = xg.make_params(xs, xs, corr_flag=corr)
    w = _windows(xg, 4, p.npx, p.npy, seed=3, m_nx=768, m_ny=768)
    with xg.Engine(p) as eng:
        eng.set_images(z, z)
        got = eng.correlate(w)
    ref = orc.run(_oracle_params(orc, p), z, z, _locs(w), precision=32)
    _same(got, ref, "zero images")
    assert np.all(got["corr"] == 0) and np.all(got["ip"] == ref["ip"])


@pytest.mark.parametrize("xs,corr", [(64, 0), (64, 1), (32, 1)], ids=["time256", "time4_256", "time4"])
def test_time_windows_across_the_edges(xg, orc, pair, xs, corr):
    """The matrix-core time searches (-time on 256 x 256 windows, -time4 on 256 x
    256 and 128 x 128) on windows that hang over every side of the images (the
    samples outside read as zero) and one entirely outside."""
    m, s = pair
    p = xg.make_params(xs, xs, corr_flag=corr)
    w = np.zeros(6, xg.WINDOW_DTYPE)
    corners = [(-100, 40), (1536 - 150, 300), (500, -90), (700, 2048 - 120), (-60, -70), (-400, 900)]
    for i, (x0, y0) in enumerate(corners):
        w[i] = (x0, y0, x0 + 3, y0 - 2)
    with xg.Engine(p, max_batch=4) as eng:
        eng.set_images(m, s)
        got = eng.correlate(w)
    _same(got, orc.run(_oracle_params(orc, p), m, s, _locs(w), precision=32), "time edges, xs %d corr %d" % (xs, corr))
