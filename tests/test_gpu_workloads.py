"""GPU: the measured workloads themselves, at the sizes the bench runs them.

* Strip workload: a 10800 x 27648 complex-int16 pair (offset field
  dx = 2.1 + 1e-4 x, dy = -7.6 + 3e-4 y, coherence per 512 x 512 block in
  [0.3, 0.95]), -nx 100 -ny 400 -xsearch 64 -ysearch 64, interp 16: 40,000
  windows of 256 x 256 in five 8192-window batches that alternate between two
  contexts.  Host results, device results and serial batches agree bit for
  bit; the injected field is recovered; windows sampled evenly over the strip
  order (so every batch contributes) equal the FP32 CPU restatement bit for
  bit.  Reference loop: do_correlation xcorr.cu:299-378.
* Learned proposal on the same strip at the default confidence threshold:
  every window's result is a plain xcorr run (the plain engine for the
  fallback, the shrunken engine on the displaced window otherwise), a sample
  of the shrunken searches equals the restatement, and the composed offsets
  stay within 0.05 px RMSE of plain xcorr.  The throughput ratio against
  plain xcorr is printed.
* Dense tracking, reduced to a 2048 x 4096 pair: 128 x 128 windows on a
  31-px grid, the exact time-domain search (-time, on the int8 matrix cores)
  on one checkerboard half and the frequency path with 8x high-res on the
  other, both classes in ONE graph launch (autotuned, radix plans included),
  replayed twice, bit-exact against the restatement with each class's plans.
"""
import importlib
import os
import time

import numpy as np
import pytest

from conftest import ROOT, gpu_present

pytestmark = pytest.mark.gpu

FIELDS = ("xoff", "yoff", "corr", "xfrac", "yfrac", "cmax", "ip", "jp", "hip", "hjp")
WEIGHTS = os.path.join(ROOT, "gmtsar-gpu_amd", "weights", "proposal_cnn.npy")
NTHREADS = min(16, len(os.sched_getaffinity(0)))


def _same(got, ref, label, fields=FIELDS):
    for f in fields:
        a, b = got[f], ref[f]
        eq = (a.view(np.uint32) == b.view(np.uint32)) if a.dtype.kind == "f" else (a == b)
        bad = np.nonzero(~eq)[0]
        assert bad.size == 0, "%s: %s differs at %s: %s vs %s" % (label, f, bad[:4], a[bad[:4]], b[bad[:4]])


def _locs(w):
    return np.stack([w["mx0"], w["my0"], w["sx0"], w["sy0"]], 1).astype(np.int32)


def _oracle_params(orc, p):
    return orc.make_params(p.xsearch, p.ysearch, ri=p.ri, interp_flag=p.interp_flag,
                           interp_factor=p.interp_factor, corr_flag=p.corr_flag)


def _oracle_sample(orc, p, m_h, s_h, w, plans=None):
    """FP32 restatement on windows w, reading only the lines they need."""
    lo = max(int(min(w["my0"].min(), w["sy0"].min())), 0)
    hi = int(max(w["my0"].max(), w["sy0"].max())) + p.npy
    loc = _locs(w)
    loc[:, 1] -= lo
    loc[:, 3] -= lo
    with orc.fft_plans(plans or {}):
        return orc.run(_oracle_params(orc, p), m_h[lo:hi], s_h[lo:hi], loc, precision=32, nthreads=NTHREADS)


@pytest.fixture(scope="module")
def strip(xg, synth):
    if not gpu_present():
        pytest.skip("no GPU")
    import torch
    print("strip workload: generating the 10800 x 27648 pair", flush=True)
    m, s = synth.strip_pair_torch(10800, 27648, lambda x: 2.1 + 1e-4 * x, lambda y: -7.6 + 3e-4 * y, device="cuda")
    torch.cuda.synchronize()
    p = xg.make_params(64, 64)
    x, y = xg.get_locations(10800, 27648, 100, 400, 64, 64)
    w = xg.make_windows(x, y, p.npx, p.npy)
    yield torch, m, s, p, w, x, y
    del m, s
    torch.cuda.empty_cache()


@pytest.fixture(scope="module")
def strip_host(strip):
    torch, m, s = strip[:3]
    return m.cpu().numpy(), s.cpu().numpy()


@pytest.fixture(scope="module")
def strip_plain(xg, strip):
    """Plain xcorr of the strip workload as the bench runs it (8192-window
    batches, two contexts), with host results."""
    torch, m, s, p, w = strip[:5]
    with xg.Engine(p, max_batch=8192) as eng:
        eng.set_images(m, s)
        host = eng.correlate(w)
        dout = torch.zeros(len(w) * xg.RESULT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
        eng.correlate(w, out=dout, flags=xg.NO_SYNC)
        eng.sync()
        dev = dout.cpu().numpy().view(xg.RESULT_DTYPE).copy()
        serial = eng.correlate(w, flags=xg.SERIAL)
    return host, dev, serial


def test_strip_workload_paths_agree(xg, strip, strip_plain):
    host, dev, serial = strip_plain
    assert len(host) == 40000
    _same(dev, host, "device results vs host results")
    _same(serial, host, "serial batches vs overlapped batches")
    np.testing.assert_array_equal(dev["cave"], host["cave"])


def test_strip_workload_recovers_injected_field(xg, strip, strip_plain):
    torch, m, s, p, w, x, y = strip
    got = strip_plain[0]
    dx, dy = xg.offsets(got, p.ri)
    ex = dx - (2.1 + 1e-4 * x)
    ey = dy - (-7.6 + 3e-4 * y)
    good = got["corr"] > 50
    print("corr > 50: %d of %d; median error dx %.4f dy %.4f px; median |error| dx %.4f dy %.4f px; "
          "90th pct |error| dx %.4f dy %.4f px" % (good.sum(), len(got), np.median(ex[good]), np.median(ey[good]),
                                                    np.median(np.abs(ex[good])), np.median(np.abs(ey[good])),
                                                    np.percentile(np.abs(ex[good]), 90),
                                                    np.percentile(np.abs(ey[good]), 90)))
    # amplitude correlation runs at about the square of the coherence (U[0.3, 0.95]
    # per block), so about a quarter of the windows are above 50
    assert good.sum() > 8000
    # the generator shifts each 512 x 512 block by the field at its centre (up to
    # 0.026 px / 0.077 px from the window's own value); xcorr's corr^(1/4)
    # high-res step is biased toward whole lags by up to a few 1/16 steps
    assert abs(np.median(ex[good])) < 0.1 and abs(np.median(ey[good])) < 0.2
    assert np.median(np.abs(ex[good])) < 0.25 and np.median(np.abs(ey[good])) < 0.35


def test_strip_workload_sample_bit_exact(xg, orc, strip, strip_host, strip_plain):
    torch, m, s, p, w = strip[:5]
    order = np.lexsort((w["my0"], w["mx0"]))          # the strip planner's order for this pair
    idx = np.sort(order[::78])                         # 513 windows, every batch of 8192 contributes
    assert len(idx) >= 512
    ref = _oracle_sample(orc, p, *strip_host, w[idx])
    _same(strip_plain[0][idx], ref, "strip workload sample")
    np.testing.assert_allclose(strip_plain[0]["cave"][idx], ref["cave"], rtol=1e-12)


def test_proposal_on_strip_workload(xg, orc, strip, strip_host, strip_plain):
    torch, m, s, p, w = strip[:5]
    prop = importlib.import_module("gmtsar_gpu_amd.proposal")
    params = np.load(WEIGHTS, allow_pickle=False)
    pr = prop.Proposer(params, max_batch=8192)
    try:
        pr.set_images(m, s)
        torch.cuda.synchronize()
        pred = pr.predict(w)
        dx, dy, res, path = pr.run(w, pred=pred)
        t0 = time.perf_counter()
        pr.run(w)
        t_prop = time.perf_counter() - t0
        t0 = time.perf_counter()
        pr.full.correlate(w)
        t_plain = time.perf_counter() - t0
        plain = strip_plain[0]
        pdx, pdy = prop.compose(plain, w, 2)
        k0 = path == 0
        _same(res[k0], plain[k0], "fallback windows vs plain xcorr")
        assert np.array_equal(dx[k0], pdx[k0]) and np.array_equal(dy[k0], pdy[k0])
        rx, ry = np.rint(pred[0]).astype(np.int64), np.rint(pred[1]).astype(np.int64)
        for k, eng in ((1, pr.small), (2, pr.small_ni)):
            sel = np.nonzero(path == k)[0]
            if sel.size == 0:
                continue
            ww = prop.shrunk_windows(w[sel], rx[sel], ry[sel], pr.xs_small)
            ref = eng.correlate(ww)
            _same(res[sel], ref, "path %d vs the shrunken engine" % k)
            cdx, cdy = prop.compose(ref, ww, 2)
            assert np.array_equal(dx[sel], cdx) and np.array_equal(dy[sel], cdy)
            pick = sel[:: max(1, sel.size // 128)][:128]
            wo = prop.shrunk_windows(w[pick], rx[pick], ry[pick], pr.xs_small)
            _same(res[pick], _oracle_sample(orc, eng.params, *strip_host, wo), "path %d sample vs oracle" % k)
        e = np.concatenate([dx - pdx, dy - pdy]).astype(np.float64)
        rmse = float(np.sqrt(np.mean(e ** 2)))
        print("proposal at tau %.2f: shrunken %.1f %% of %d windows, RMSE vs plain %.4f px, "
              "%.0f patches/s vs plain %.0f patches/s (x%.2f, host results, one call each)"
              % (pr.tau, 100 * (path > 0).mean(), len(w), rmse, len(w) / t_prop, len(w) / t_plain, t_plain / t_prop))
        assert (path > 0).any()
        assert rmse <= 0.05
    finally:
        pr.close()


@pytest.fixture(scope="module")
def dense(xg, synth):
    if not gpu_present():
        pytest.skip("no GPU")
    import torch
    rng = np.random.default_rng(5)
    field = synth.sinusoid_field(rng)
    coh = synth.smooth_coherence(rng)
    nx, ny = 2048, 4096
    m, s = synth.field_pair_torch(nx, ny, field, coh, device="cuda")
    torch.cuda.synchronize()
    # x_inc = (2048 - 192) / (57 + 3) = 30, y_inc = (4096 - 192) / (125 + 1) = 30
    x, y = xg.get_locations(nx, ny, 57, 125, 32, 32)
    nxl, nyl = 57, 125
    i, j = np.meshgrid(np.arange(nyl), np.arange(nxl), indexing="ij")
    half = ((i + j) & 1).ravel() == 1
    work = []
    for cf, sel in ((xg.TIME, half), (xg.FREQ, ~half)):
        p = xg.make_params(32, 32, interp_factor=8, corr_flag=cf)
        w = xg.make_windows(x[sel], y[sel], p.npx, p.npy)
        out = torch.zeros(len(w) * xg.RESULT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
        work.append(dict(params=p, master=m, slave=s, windows=w, out=out))
    return torch, m, s, work


def test_dense_workload_one_graph_bit_exact(xg, orc, dense):
    torch, m, s, work = dense
    assert sum(len(q["windows"]) for q in work) == 57 * 125
    with xg.Multi(work, tune=True, tune_radix=True) as mu:
        classes, pc = mu.classes()
        assert len(classes) == 2 and sorted(c[0].corr_flag for c in classes) == [xg.TIME, xg.FREQ]
        got = []
        for rep in range(2):
            for q in work:
                q["out"].fill_(0xff)
            torch.cuda.synchronize()
            mu.run()
            torch.cuda.synchronize()
            now = [q["out"].cpu().numpy().view(xg.RESULT_DTYPE).copy() for q in work]
            if got:
                for a, b in zip(now, got):
                    _same(a, b, "graph replay %d" % rep)
            got = now
    m_h, s_h = m.cpu().numpy(), s.cpu().numpy()
    for i, q in enumerate(work):
        p = q["params"]
        _, t = classes[pc[i]]
        plans = {p.npx: t["radix_x"]}
        plans[p.npy] = t["radix_y"]
        w = q["windows"]
        idx = np.arange(0, len(w), max(1, len(w) // 192))
        ref = _oracle_sample(orc, p, m_h, s_h, w[idx], plans)
        print("dense class corr_flag %d: %d windows, tuning %s, %d sampled" % (p.corr_flag, len(w), t, len(idx)))
        _same(got[i][idx], ref, "dense class %d" % p.corr_flag)
        np.testing.assert_allclose(got[i]["cave"][idx], ref["cave"], rtol=1e-12)
        assert np.median(got[i]["corr"]) > 10
