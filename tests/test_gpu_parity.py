"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle.

* Bit-exact against the FP32 sequential restatement (oracle/xcorr_seq32.c):
  integer peaks, NCC values, sub-pixel offsets, stage dumps.
* Within tolerance of the FP64 textbook restatement (oracle/xcorr_ref64.c):
  |corr| < 0.05 (percent units), |xfrac|,|yfrac| <= 1/ifc, peaks equal.
* Full-size properties: the injected constant shift is recovered.
"""
import numpy as np
import pytest

from conftest import gpu_present

pytestmark = pytest.mark.gpu

FIELDS_EXACT = ("xoff", "yoff", "corr", "xfrac", "yfrac", "cmax", "ip", "jp", "hip", "hjp")


@pytest.fixture(scope="module")
def pair(synth):
    if not gpu_present():
        pytest.skip("no GPU")
    return synth.constant_shift_pair(1024, 2048, 3.37, -12.81, 0.8)


def _windows(xg, p, nxl, nyl, m_nx=1024, m_ny=2048, rshift=0, ashift=0):
    x, y = xg.get_locations(m_nx, m_ny, nxl, nyl, p.xsearch, p.ysearch)
    return xg.make_windows(x, y, p.npx, p.npy, rshift, ashift), x, y


def _oracle_params(orc, p):
    return orc.make_params(p.xsearch, p.ysearch, ri=p.ri, interp_flag=p.interp_flag,
                           interp_factor=p.interp_factor, corr_flag=p.corr_flag)


def _locs(w):
    return np.stack([w["mx0"], w["my0"], w["sx0"], w["sy0"]], 1).astype(np.int32)


def _assert_bit_exact(got, ref, label):
    for f in FIELDS_EXACT:
        a, b = got[f], ref[f]
        if a.dtype.kind == "f":
            same = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))
        else:
            same = a == b
        bad = np.nonzero(~same)[0]
        assert bad.size == 0, "%s: field %s differs at %s: gpu %s oracle %s" % (
            label, f, bad[:5], a[bad[:5]], b[bad[:5]])
    np.testing.assert_allclose(got["cave"], ref["cave"], rtol=1e-12, err_msg=label)


CASES = [
    # xsearch, ysearch, ri, interp_factor, interp_flag
    (16, 16, 2, 16, 1),
    (32, 32, 2, 16, 1),
    (64, 64, 2, 16, 1),
    (32, 64, 2, 8, 1),
    (64, 32, 1, 16, 1),
    (16, 32, 4, 4, 1),
    (32, 32, 2, 16, 0),
    (128, 128, 2, 8, 1),
    (64, 8, 2, 16, 1),
    (64, 4, 2, 8, 1),
    (8, 64, 2, 16, 1),
    (64, 64, 4, 16, 1),
    (64, 64, 1, 16, 0),
    (32, 32, 2, 2, 1),          # high-res grid of 16 columns (rows not uniform over a wave)
    (64, 64, 2, 32, 1),         # high-res grid of 256 columns (one column per thread)
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "xs%d_ys%d_ri%d_ifc%d_i%d" % c)
def test_bit_exact_vs_fp32_oracle(xg, orc, pair, case):
    xs, ys, ri, ifc, interp = case
    m, s = pair
    p = xg.make_params(xs, ys, ri=ri, interp_factor=ifc, interp_flag=interp)
    w, _, _ = _windows(xg, p, 5, 4)
    with xg.Engine(p, max_batch=8) as eng:          # 20 locations -> 3 batches
        eng.set_images(m, s)
        got = eng.correlate(w)
    ref = orc.run(_oracle_params(orc, p), m, s, _locs(w), precision=32)
    _assert_bit_exact(got, ref, "case %s" % (case,))


def test_stage_dumps_bit_exact(xg, orc, pair):
    m, s = pair
    p = xg.make_params(64, 64)
    w, _, _ = _windows(xg, p, 3, 2)
    with xg.Engine(p) as eng:
        eng.set_images(m, s)
        got = eng.correlate(w, flags=xg.KEEP_SURFACE)
        for slot in range(len(w)):
            amp1, amp2 = eng.dump(0, slot), eng.dump(1, slot)
            means, surf = eng.dump(3, slot), eng.dump(2, slot)
            mw = (m[w["my0"][slot]:w["my0"][slot] + p.npy, w["mx0"][slot]:w["mx0"][slot] + p.npx]
                  .astype(np.float32))
            sw = (s[w["sy0"][slot]:w["sy0"][slot] + p.npy, w["sx0"][slot]:w["sx0"][slot] + p.npx]
                  .astype(np.float32))
            rec, d = orc.patch(_oracle_params(orc, p), mw[..., 0] + 1j * mw[..., 1],
                               sw[..., 0] + 1j * sw[..., 1], precision=32, dump=True)
            assert np.array_equal(amp1.view(np.uint32), d["amp1"].view(np.uint32)), "amp master, slot %d" % slot
            # the slave's amplitudes are kept where the NCC reads them: the template
            t = (slice(p.ysearch, p.ysearch + p.ny_corr), slice(p.xsearch, p.xsearch + p.nx_corr))
            assert np.array_equal(amp2[t].view(np.uint32), d["amp2"][t].view(np.uint32)), "amp slave, slot %d" % slot
            assert np.array_equal(means, d["mean"]), "means, slot %d" % slot
            assert np.array_equal(surf.view(np.uint32), d["surf"].view(np.uint32)), "surface, slot %d" % slot
            assert got["ip"][slot] == rec["ip"] and got["jp"][slot] == rec["jp"]


@pytest.mark.parametrize("xs,nxl,nyl", [(32, 10, 8), (64, 8, 10)], ids=["xs32", "xs64_headline"])
def test_tolerance_vs_fp64_oracle(xg, orc, pair, xs, nxl, nyl):
    """GPU (FP32 specification) against the FP64 textbook path; xs 64 is the
    bench's shape (256 x 256 windows, 16x high-res), where the spectral demean,
    Hermitian completion and c1 + i c2 packing all take part."""
    m, s = pair
    p = xg.make_params(xs, xs)
    w, _, _ = _windows(xg, p, nxl, nyl)
    with xg.Engine(p) as eng:
        eng.set_images(m, s)
        got = eng.correlate(w)
    ref = orc.run(_oracle_params(orc, p), m, s, _locs(w), precision=64)
    # integer peaks: exact (these inputs have well separated peaks)
    assert np.array_equal(got["ip"], ref["ip"]) and np.array_equal(got["jp"], ref["jp"])
    # stated tolerances: corr within 0.05 percent points, sub-pixel within one high-res step
    assert np.max(np.abs(got["corr"] - ref["corr"])) < 0.05
    step = 1.0 / p.interp_factor
    assert np.max(np.abs(got["xfrac"] - ref["xfrac"])) <= step + 1e-6
    assert np.max(np.abs(got["yfrac"] - ref["yfrac"])) <= step + 1e-6
    np.testing.assert_allclose(got["cmax"], ref["cmax"], rtol=1e-5)


def test_small_workload_recovers_shift(xg, orc, pair):
    """The small workload: 1024 x 2048, -nx 20 -ny 50 -xsearch 64 -ysearch 64 (1000 patches)."""
    m, s = pair
    p = xg.make_params(64, 64)
    w, x, y = _windows(xg, p, 20, 50)
    assert len(w) == 1000
    with xg.Engine(p) as eng:
        eng.set_images(m, s)
        got = eng.correlate(w)
    dx, dy = xg.offsets(got, p.ri)
    # resolution 1/(ri*ifc) = 1/32 px in range, 1/16 px in azimuth; xcorr's corr^(1/4)
    # interpolation is biased by up to a few steps
    assert abs(np.median(dx) - 3.37) < 0.1
    assert abs(np.median(dy) + 12.81) < 0.2
    assert np.median(got["corr"]) > 50.0
    # bit-exact against the FP32 restatement on a sample of 40 patches
    idx = np.arange(0, 1000, 25)
    ref = orc.run(_oracle_params(orc, p), m, s, _locs(w[idx]), precision=32)
    _assert_bit_exact(got[idx], ref, "small workload sample")


def test_windows_outside_image_read_zero(xg, orc, pair):
    m, s = pair
    p = xg.make_params(32, 32)
    w = np.zeros(4, xg.WINDOW_DTYPE)
    w["mx0"] = [-40, 1000, 500, -200]
    w["my0"] = [-10, 2000, -100, 50]
    w["sx0"] = w["mx0"] + [3, -5, 0, 7]
    w["sy0"] = w["my0"] + [-2, 4, 1, 0]
    with xg.Engine(p) as eng:
        eng.set_images(m, s)
        got = eng.correlate(w)
    ref = orc.run(_oracle_params(orc, p), m, s, _locs(w), precision=32)
    _assert_bit_exact(got, ref, "edge windows")


@pytest.mark.parametrize("xs,corr", [(32, 2), (64, 2), (128, 2), (32, 0), (16, 2)],
                         ids=["strips128", "strips256", "strips512", "time_strips128", "per_patch64"])
def test_real_float_format(xg, orc, pair, xs, corr):
    m, s = pair
    mf = np.ascontiguousarray(m[..., 0].astype(np.float32))
    sf = np.ascontiguousarray(s[..., 0].astype(np.float32))
    p = xg.make_params(xs, xs, format=xg.FLOAT32, corr_flag=corr)
    w, _, _ = _windows(xg, p, 4, 4)
    with xg.Engine(p) as eng:
        eng.set_images(mf, sf)
        got = eng.correlate(w)
    ref = orc.run(_oracle_params(orc, p), mf, sf, _locs(w), precision=32, fmt=1)
    _assert_bit_exact(got, ref, "real float")


def test_device_pointer_path(xg, orc, pair):
    import torch
    m, s = pair
    p = xg.make_params(32, 32)
    w, _, _ = _windows(xg, p, 6, 5)
    dm, ds = torch.from_numpy(m).cuda(), torch.from_numpy(s).cuda()
    dw = torch.from_numpy(w.view(np.int32).reshape(-1, 4).copy()).cuda()
    dout = torch.zeros(len(w) * xg.RESULT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    with xg.Engine(p) as eng:
        eng.set_images(dm, ds)
        eng.correlate(dw, out=dout)
        eng.sync()
        got = dout.cpu().numpy().view(xg.RESULT_DTYPE)
        host = eng.correlate(w)
    _assert_bit_exact(got, host, "device pointers vs host pointers")


@pytest.mark.parametrize("xs", [32, 64, 128])
def test_peak_on_surface_border_bit_exact(xg, orc, synth, xs):
    """Offsets that put the integer peak on or next to the border of the lag
    window, so the 8 x 8 high-res neighbourhood wraps across surface rows or
    leaves the surface (d_cppow_strip's flat index, conv_ampl_demean.cu:75-95).
    The default kernels keep only that neighbourhood (recomputing its rows,
    or, for 128 x 128 windows, from the values the lanes still hold);
    with XCORR_HIP_KEEP_SURFACE they write the whole surface: both equal the
    restatement bit for bit, and a surface dump needs the flag."""
    k = xs / 64.0
    p = xg.make_params(xs, xs)
    n = 4 * xs
    # lags cover D in (-31.5, 32] px in range (ri 2) and (-63, 64] px in azimuth (scaled by xs / 64)
    shifts = [(31.7, 0.3), (-31.3, 1.2), (0.4, 63.6), (-0.6, -62.7), (31.6, 63.4), (-31.2, -62.6), (31.9, -62.8)]
    for dx, dy in shifts:
        m, s = synth.constant_shift_pair(2 * n + 64, 2 * n + 64, dx * k, dy * k, 0.9, seed=31, noise_seed=32)
        w = np.zeros(6, xg.WINDOW_DTYPE)
        w["mx0"] = [0, 32, 64, 16, 48, 60]
        w["my0"] = [0, 64, 32, 48, 16, 60]
        w["sx0"], w["sy0"] = w["mx0"], w["my0"]
        with xg.Engine(p, max_batch=8) as eng:
            eng.set_images(m, s)
            nbr = eng.correlate(w)
            full = eng.correlate(w, flags=xg.KEEP_SURFACE)
            eng.correlate(w[:1], flags=xg.KEEP_SURFACE)        # dumps reach the last call's last batch
            assert eng.dump(2, 0).max() == full["cmax"][0]
            eng.correlate(w[:1])
            with pytest.raises(xg.XcorrError, match="KEEP_SURFACE"):
                eng.dump(2, 0)
        ref = orc.run(_oracle_params(orc, p), m, s, _locs(w), precision=32)
        label = "xs %d shift (%.1f, %.1f)" % (xs, dx * k, dy * k)
        _assert_bit_exact(nbr, full, label + " neighbourhood vs whole surface")
        _assert_bit_exact(nbr, ref, label)
        edge = (np.minimum(nbr["ip"], 2 * xs - 1 - nbr["ip"]) < 4) | (np.minimum(nbr["jp"], 2 * xs - 1 - nbr["jp"]) < 4)
        assert edge.sum() >= 3, label
