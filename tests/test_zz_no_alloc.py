"""GPU: xcorr_hip_correlate_noalloc against xcorr_hip_correlate.

The no-alloc entry point runs the per-patch kernels on buffers carved from
one static device array per GPU.  Its results must equal the existing entry
point's bit for bit, on two calls in a row -- the second with fewer (and
different) windows, so nothing the first call left in the shared array can
leak into the second -- for every kind of kernel chain: the 256 x 256
register path (strip-shaped contexts), 512 x 512, the generic LDS kernels,
and -time and -time4 on the int8 matrix cores (128 x 128 and 256 x 256
windows).  Windows include some partly outside the images.
"""
import numpy as np
import pytest

from conftest import gpu_present

pytestmark = pytest.mark.gpu

FIELDS = ("xoff", "yoff", "corr", "xfrac", "yfrac", "cmax", "ip", "jp", "hip", "hjp")

CASES = [
    # xsearch, ysearch, corr_flag (0 time, 1 time4, 2 freq), interp factor, windows of call 1, of call 2
    (64, 64, 2, 16, 100, 20),       # 48 windows per internal batch: 3 batches, then 1
    (128, 128, 2, 8, 30, 7),        # 12 per batch
    (32, 32, 2, 8, 400, 41),        # 128 x 128 windows through the register-resident per-patch kernels
    (16, 32, 2, 16, 500, 50),       # 388 per batch
    (32, 32, 0, 8, 400, 37),        # 339 per batch
    (64, 64, 0, 16, 100, 9),        # 256 x 256 -time on the int8 matrix cores: 85 per batch
    (32, 32, 1, 16, 450, 9),        # -time4, 339 per batch
    (64, 64, 1, 16, 100, 9),        # 256 x 256 -time4: 85 per batch
]


@pytest.fixture(scope="module")
def pair(synth):
    if not gpu_present():
        pytest.skip("no GPU")
    import torch
    m, s = synth.constant_shift_pair(1536, 2048, 3.37, -12.81, 0.8, seed=21, noise_seed=22)
    return torch, m, s, torch.from_numpy(m).cuda(), torch.from_numpy(s).cuda()


def _windows(xg, p, n, seed, m_nx=1536, m_ny=2048):
    rng = np.random.default_rng(seed)
    w = np.zeros(n, xg.WINDOW_DTYPE)
    w["mx0"] = rng.integers(-40, m_nx - p.npx + 40, n)
    w["my0"] = rng.integers(-40, m_ny - p.npy + 40, n)
    w["sx0"] = w["mx0"] + rng.integers(-6, 7, n)
    w["sy0"] = w["my0"] + rng.integers(-6, 7, n)
    return w


def _same(got, ref, label):
    """Every peak, offset and correlation field bit for bit; the surface mean
    cave (an FP64 sum whose order differs between the strip and per-patch
    kernels) to 1e-12 relative, the bar every parity test holds it to."""
    for f in FIELDS:
        a, b = got[f], ref[f]
        eq = (a.view(np.uint32) == b.view(np.uint32)) if a.dtype.kind == "f" else (a == b)
        bad = np.nonzero(~eq)[0]
        assert bad.size == 0, "%s: %s differs at %s: %s vs %s" % (label, f, bad[:4], a[bad[:4]], b[bad[:4]])
    np.testing.assert_allclose(got["cave"], ref["cave"], rtol=1e-12, err_msg=label)


@pytest.mark.parametrize("case", CASES, ids=lambda c: "xs%d_ys%d_corr%d_ifc%d" % c[:4])
def test_noalloc_matches_correlate_on_two_calls(xg, pair, case):
    torch, m, s, mt, st = pair
    xs, ys, cf, ifc, n1, n2 = case
    p = xg.make_params(xs, ys, interp_factor=ifc, corr_flag=cf)
    print("no-alloc case xs %d ys %d corr_flag %d ifc %d: %d then %d windows" % case, flush=True)
    with xg.Engine(p, max_batch=64) as eng:
        eng.set_images(mt, st)
        for call, (n, seed) in enumerate(((n1, 100 + xs), (n2, 200 + xs))):
            w = _windows(xg, p, n, seed)
            ref = eng.correlate(w)                                   # the existing entry point
            dw = torch.from_numpy(w.view(np.int32).reshape(-1, 4).copy()).cuda()
            out = torch.full((n * xg.RESULT_DTYPE.itemsize,), 0xff, dtype=torch.uint8, device="cuda")
            eng.correlate_noalloc(dw, out)
            got = out.cpu().numpy().view(xg.RESULT_DTYPE)
            _same(got, ref, "call %d (%d windows)" % (call + 1, n))
        # queued without a wait; the engine stream orders it
        out.fill_(0xff)
        eng.correlate_noalloc(dw, out, flags=xg.NO_SYNC)
        eng.sync()
        _same(out.cpu().numpy().view(xg.RESULT_DTYPE), ref, "NO_SYNC call")


def test_noalloc_rejects_bad_arguments(xg, pair):
    torch, m, s, mt, st = pair
    p = xg.make_params(32, 32)
    with xg.Engine(p, max_batch=8) as eng:
        dw = torch.zeros((4, 4), dtype=torch.int32, device="cuda")
        out = torch.zeros(4 * xg.RESULT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
        with pytest.raises(xg.XcorrError, match="no images"):
            eng.correlate_noalloc(dw, out)
        eng.set_images(mt, st)
        with pytest.raises(xg.XcorrError, match="only accepted flag"):
            eng.correlate_noalloc(dw, out, flags=xg.DEVICE_PTRS)
        eng.correlate_noalloc(dw[:0], out)                           # n = 0 is a no-op
