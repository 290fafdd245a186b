"""GPU parity of the strip path (square windows of 128, 256 and 512 samples:
-xsearch = -ysearch = 32, 64, 128; ri 2, frequency mode).

Windows that share the master column, the slave column and the slave line
offset are computed from shared strip lines.  These cases exercise the
planner: overlapping, touching, duplicate and isolated windows in one column,
mixed slave offsets, unsorted input, windows across the image edges, batches
that split a strip, and the three pointer modes.  Every result must equal the
FP32 CPU restatement bit for bit, in input order.  A caller-set radix plan on
a strip-shaped context runs the per-patch kernels, also bit-exact.
"""
import numpy as np
import pytest

from conftest import gpu_present
from test_gpu_parity import _assert_bit_exact, _locs, _oracle_params

pytestmark = pytest.mark.gpu

SIZES = [32, 64, 128]        # xsearch = ysearch; windows 4 xsearch square


@pytest.fixture(scope="module")
def pairs(synth):
    if not gpu_present():
        pytest.skip("no GPU")
    out = {}
    for xs in SIZES:
        k = xs / 64.0
        out[xs] = synth.constant_shift_pair(int(768 * k), int(1536 * k), -2.63 * k, 5.41 * k, 0.8)
    return out


def _strip_windows(xg, xs):
    """The 256-window layout scaled to windows of 4 xs samples."""
    k = xs / 64.0
    rng = np.random.default_rng(7)
    rows = []
    # column A: dense overlapping run, a duplicate, a touching window, a gap
    for my0 in (100, 137, 137, 150, 300, 356, 612, 868, 1100):
        rows.append((200, my0, 203, my0 - 5))
    # column A with a second slave offset (separate strips, same master lines)
    for my0 in (120, 180, 240):
        rows.append((200, my0, 203, my0 + 9))
    # column B: random overlapping lines
    for my0 in sorted(rng.integers(0, 1200, 10)):
        rows.append((420, int(my0), 418, int(my0) + 2))
    # edges: left/right/top/bottom overhang, fully outside
    rows += [(-60, 20, -57, 18), (600, 1400, 603, 1395), (-100, -200, -98, -190), (700, -30, 705, -25),
             (-300, 500, -297, 495), (900, 900, 905, 900)]
    w = np.rint(np.array(rows, np.float64) * k).astype(np.int32)
    perm = rng.permutation(len(w))                 # unsorted input order
    return np.ascontiguousarray(w[perm]).view(xg.WINDOW_DTYPE).reshape(-1)


@pytest.mark.parametrize("xs", SIZES)
def test_strip_planner_bit_exact(xg, orc, pairs, xs):
    m, s = pairs[xs]
    p = xg.make_params(xs, xs)
    w = _strip_windows(xg, xs)
    ref = orc.run(_oracle_params(orc, p), m, s, _locs(w), precision=32)
    for mb in (7, 64):                              # 7: strips split across batches
        with xg.Engine(p, max_batch=mb) as eng:
            eng.set_images(m, s)
            got = eng.correlate(w)
        _assert_bit_exact(got, ref, "xs %d max_batch %d" % (xs, mb))


@pytest.mark.parametrize("xs", SIZES)
def test_strip_shape_with_radix_plan_runs_per_patch(xg, orc, pairs, xs):
    m, s = pairs[xs]
    p = xg.make_params(xs, xs)
    w = _strip_windows(xg, xs)
    plan = {32: [8, 16], 64: [4, 4, 16], 128: [16, 16, 2]}[xs]
    with xg.Engine(p, max_batch=16) as eng:
        eng.set_images(m, s)
        eng.set_tuning(radix_x=plan, radix_y=plan)
        got = eng.correlate(w)
        default = {32: [16, 8], 64: [16, 16], 128: [8, 8, 8]}[xs]
        eng.set_tuning(radix_x=default, radix_y=default)   # back to the default plans: strips again
        again = eng.correlate(w)
    with orc.fft_plans({4 * xs: plan}):
        ref = orc.run(_oracle_params(orc, p), m, s, _locs(w), precision=32)
    _assert_bit_exact(got, ref, "xs %d radix %s" % (xs, plan))
    _assert_bit_exact(again, orc.run(_oracle_params(orc, p), m, s, _locs(w), precision=32), "xs %d default" % xs)


@pytest.mark.parametrize("xs", SIZES)
def test_strip_pointer_modes(xg, pairs, xs):
    import torch
    m, s = pairs[xs]
    p = xg.make_params(xs, xs)
    w = _strip_windows(xg, xs)
    dm, ds = torch.from_numpy(m).cuda(), torch.from_numpy(s).cuda()
    dw = torch.from_numpy(w.view(np.int32).reshape(-1, 4).copy()).cuda()
    nbytes = len(w) * xg.RESULT_DTYPE.itemsize
    with xg.Engine(p, max_batch=9) as eng:
        eng.set_images(dm, ds)
        host = eng.correlate(w)
        d1 = torch.zeros(nbytes, dtype=torch.uint8, device="cuda")
        eng.correlate(dw, out=d1)                   # device windows and results
        d2 = torch.zeros(nbytes, dtype=torch.uint8, device="cuda")
        eng.correlate(w, out=d2, flags=xg.NO_SYNC)  # host windows, device results
        eng.sync()
    for d, label in ((d1, "device windows"), (d2, "host windows, device results")):
        got = d.cpu().numpy().view(xg.RESULT_DTYPE)
        _assert_bit_exact(got, host, label)


@pytest.mark.parametrize("xs", SIZES)
def test_strip_dumps_in_shared_strip(xg, orc, pairs, xs):
    m, s = pairs[xs]
    n = 4 * xs
    k = xs / 64.0
    p = xg.make_params(xs, xs)
    w = np.rint(np.array([(300, my0, 297, my0 + 3) for my0 in (400, 450, 520, 560)]) * k).astype(np.int32).view(
        xg.WINDOW_DTYPE).reshape(-1)
    with xg.Engine(p) as eng:
        eng.set_images(m, s)
        eng.correlate(w, flags=xg.KEEP_SURFACE)
        for slot in (1, 2):
            amp1, amp2 = eng.dump(0, slot), eng.dump(1, slot)
            means, surf = eng.dump(3, slot), eng.dump(2, slot)
            mw = m[w["my0"][slot]:w["my0"][slot] + n, w["mx0"][slot]:w["mx0"][slot] + n].astype(np.float32)
            sw = s[w["sy0"][slot]:w["sy0"][slot] + n, w["sx0"][slot]:w["sx0"][slot] + n].astype(np.float32)
            _, d = orc.patch(_oracle_params(orc, p), mw[..., 0] + 1j * mw[..., 1], sw[..., 0] + 1j * sw[..., 1],
                             precision=32, dump=True)
            t = (slice(n // 4, 3 * n // 4), slice(n // 4, 3 * n // 4))
            assert np.array_equal(amp1.view(np.uint32), d["amp1"].view(np.uint32)), "amp master, slot %d" % slot
            assert np.array_equal(amp2[t].view(np.uint32), d["amp2"][t].view(np.uint32)), "amp slave, slot %d" % slot
            assert np.array_equal(means, d["mean"]), "means, slot %d" % slot
            assert np.array_equal(surf.view(np.uint32), d["surf"].view(np.uint32)), "surface, slot %d" % slot


def test_time_strips_bit_exact(xg, orc, pairs):
    """-time on 128 x 128 windows (xs = ys = 32) also runs on strip lines: the
    planner cases above, against the restatement's full integer search."""
    m, s = pairs[32]
    p = xg.make_params(32, 32, corr_flag=xg.TIME, interp_factor=8)
    w = _strip_windows(xg, 32)
    ref = orc.run(_oracle_params(orc, p), m, s, _locs(w), precision=32)
    for mb in (7, 64):
        with xg.Engine(p, max_batch=mb) as eng:
            eng.set_images(m, s)
            got = eng.correlate(w)
        _assert_bit_exact(got, ref, "time max_batch %d" % mb)


@pytest.mark.parametrize("corr", ["freq", "time"])
def test_strip_batches_on_two_contexts(xg, orc, pairs, corr):
    """With more than one batch, strip batches alternate between the engine's
    context and a second one (own buffers and stream).  Host results, device
    results and XCORR_HIP_SERIAL results (one stream) equal the restatement,
    and dumps after the call read the buffers of whichever context ran the
    last batch: they equal the dumps of the same location correlated alone."""
    import torch
    xs = 32
    m, s = pairs[xs]
    p = xg.make_params(xs, xs, corr_flag=xg.TIME if corr == "time" else xg.FREQ, interp_factor=8)
    w = _strip_windows(xg, xs)
    ref = orc.run(_oracle_params(orc, p), m, s, _locs(w), precision=32)
    nbytes = len(w) * xg.RESULT_DTYPE.itemsize
    checked = 0
    for mb in (4, 5, 6, 9):                         # even and odd batch counts
        with xg.Engine(p, max_batch=mb) as eng:
            eng.set_images(m, s)
            got = eng.correlate(w)
            _assert_bit_exact(got, ref, "%s max_batch %d, host results" % (corr, mb))
            _assert_bit_exact(eng.correlate(w, flags=xg.SERIAL), ref, "%s max_batch %d, serial" % (corr, mb))
            d = torch.zeros(nbytes, dtype=torch.uint8, device="cuda")
            eng.correlate(w, out=d, flags=xg.NO_SYNC | xg.KEEP_SURFACE)
            eng.sync()
            _assert_bit_exact(d.cpu().numpy().view(xg.RESULT_DTYPE), ref, "%s max_batch %d, device" % (corr, mb))
            dumps = {}
            for loc in range(len(w)):
                try:
                    dumps[loc] = [eng.dump(k, loc) for k in range(4)]
                except xg.XcorrError:
                    continue                        # not in the last batch
            assert dumps, "no location of the last batch found"
            t = (slice(xs, 3 * xs), slice(xs, 3 * xs))   # the slave amplitudes are kept on the template only
            for loc, got_d in dumps.items():
                eng.correlate(w[loc:loc + 1], flags=xg.KEEP_SURFACE)   # one batch, on the engine's own context
                for k in range(4):
                    one = eng.dump(k, 0)
                    a, b = (got_d[k][t], one[t]) if k == 1 else (got_d[k], one)
                    assert np.array_equal(np.ascontiguousarray(a).view(np.uint8),
                                          np.ascontiguousarray(b).view(np.uint8)), (corr, mb, loc, k)
                checked += 1
    assert checked >= 4
