"""GPU parity of the time-domain search (-time, -time4): exact integer
correlation over every lag (do_time_int_xcorr.cu:146-189), the scan-order peak,
then the high-res refinement of the correlation surface.  Bit-exact against the
FP32 CPU restatement; the full-search surface against an int64 brute force."""
import numpy as np
import pytest

from conftest import gpu_present

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pair(synth):
    if not gpu_present():
        pytest.skip("no GPU")
    return synth.constant_shift_pair(512, 768, 2.4, -1.7, 0.7, seed=11, noise_seed=12)


def _run(xg, orc, pair, xs, ys, corr_flag, ri=2, ifc=16, interp=1, nxl=4, nyl=3):
    m, s = pair
    p = xg.make_params(xs, ys, ri=ri, interp_factor=ifc, interp_flag=interp, corr_flag=corr_flag)
    x, y = xg.get_locations(512, 768, nxl, nyl, xs, ys)
    w = xg.make_windows(x, y, p.npx, p.npy)
    with xg.Engine(p, max_batch=5) as eng:
        eng.set_images(m, s)
        got = eng.correlate(w)
        first = (len(w) - 1) // 5 * 5
        eng.correlate(w[first:first + 1], flags=xg.KEEP_SURFACE)   # dumps reach the last call's last batch
        surf0 = eng.dump(2, 0)
        amp = (eng.dump(0, 0), eng.dump(1, 0), eng.dump(3, 0))
    op = orc.make_params(xs, ys, ri=ri, interp_factor=ifc, interp_flag=interp, corr_flag=corr_flag)
    locs = np.stack([w["mx0"], w["my0"], w["sx0"], w["sy0"]], 1).astype(np.int32)
    ref = orc.run(op, m, s, locs, precision=32)
    return p, got, ref, surf0, amp


def _exact(got, ref):
    for f in ("xoff", "yoff", "corr", "xfrac", "yfrac", "cmax", "ip", "jp", "hip", "hjp"):
        a, b = got[f], ref[f]
        same = a.view(np.uint32) == b.view(np.uint32) if a.dtype.kind == "f" else a == b
        assert np.all(same), (f, np.nonzero(~same)[0][:5], a[~same][:5], b[~same][:5])
    np.testing.assert_allclose(got["cave"], ref["cave"], rtol=1e-12)


def _template_ints(a2, m2f, xs, ys=None):
    """The template's integers (int)(a - m2) (C truncation) in a window-sized
    array, zero elsewhere: strip lines and the time path leave the slave
    amplitudes outside the template unset."""
    ys = xs if ys is None else ys
    i2 = np.zeros(a2.shape, np.int64)
    i2[ys:3 * ys, xs:3 * xs] = (a2[ys:3 * ys, xs:3 * xs] - m2f).astype(np.int64)
    return i2


@pytest.mark.parametrize("xs,ys", [(8, 8), (16, 8), (16, 32), (32, 32)])
@pytest.mark.parametrize("corr_flag", [0, 1], ids=["time", "time4"])
def test_time_bit_exact(xg, orc, pair, xs, ys, corr_flag):
    p, got, ref, _, _ = _run(xg, orc, pair, xs, ys, corr_flag)
    _exact(got, ref)
    # -time4 reports 0 where the 4th-power ratio is <= 0.5 (do_time_int_xcorr.cu:135-139)
    assert np.all(got["corr"] > 0) if corr_flag == 0 else np.all(got["corr"] >= 0)


def test_time_surface_matches_int64_bruteforce(xg, orc, pair):
    p, got, ref, surf, (a1, a2, means) = _run(xg, orc, pair, 8, 8, 0, nxl=2, nyl=1)
    m1f, m2f = np.float32(means[0]), np.float32(means[1])
    i1 = (a1 - m1f).astype(np.int64)            # C truncation toward zero
    i2 = _template_ints(a2, m2f, 8)
    xs = ys = 8
    b = i2[ys:ys + 2 * ys, xs:xs + 2 * xs]
    brute = np.zeros((2 * ys, 2 * xs))
    for ic in range(2 * ys):
        for jc in range(2 * xs):
            a = i1[ic:ic + 2 * ys, jc:jc + 2 * xs]
            num, d1, d2 = int((a * b).sum()), int((a * a).sum()), int((b * b).sum())
            den = np.sqrt(float(d1) * float(d2))
            brute[ic, jc] = 0.0 if den == 0 else 100.0 * abs(num / den)
    assert np.array_equal(surf, brute)


def test_time_nointerp_and_norange(xg, orc, pair):
    _, got, ref, _, _ = _run(xg, orc, pair, 16, 16, 0, ri=1, interp=0)
    _exact(got, ref)
    assert np.all(got["xfrac"] == 0) and np.all(got["hip"] == -1)


def test_time_surface_xs32_matches_int64_bruteforce(xg, orc, pair):
    """-xsearch 32 -ysearch 32 -time runs on the int8 matrix cores (limb
    decomposition): its surface must equal the exact int64 correlation."""
    from numpy.lib.stride_tricks import sliding_window_view
    p, got, ref, surf, (a1, a2, means) = _run(xg, orc, pair, 32, 32, 0, nxl=2, nyl=1)
    _exact(got, ref)
    m1f, m2f = np.float32(means[0]), np.float32(means[1])
    i1 = (a1 - m1f).astype(np.int64)
    i2 = _template_ints(a2, m2f, 32)
    b = i2[32:96, 32:96]
    win = sliding_window_view(i1[:127, :127], (64, 64))          # (64, 64, 64, 64): lag (dy, dx), then (y, x)
    num = np.einsum("ijkl,kl->ij", win, b)
    d1 = np.einsum("ijkl,ijkl->ij", win, win)
    d2 = int((b * b).sum())
    den = np.sqrt(d1.astype(np.float64) * float(d2))
    brute = np.where(den == 0, 0.0, 100.0 * np.abs(num / np.where(den == 0, 1.0, den)))
    assert np.array_equal(surf, brute)


def _scaled(pair, fm, fs, left_only=False):
    m, s = pair
    out = []
    for img, f in ((m, fm), (s, fs)):
        v = img.astype(np.float64)
        if left_only:
            v[:, : v.shape[1] // 2] *= f
        else:
            v *= f
        out.append(np.clip(np.rint(v), -32768, 32767).astype(np.int16))
    return tuple(out)


@pytest.mark.parametrize("fm,fs,left", [(1, 1, False), (5, 1, False), (1, 5, False), (5, 5, False), (5, 5, True)],
                         ids=["small", "master_big", "template_big", "both_big", "mixed"])
def test_time_xs32_limb_counts(xg, orc, pair, fm, fs, left):
    """Windows whose integer values leave [-8256, 8127] need the third 7-bit
    limb; the others skip its products.  Every combination (and a batch that
    mixes them) must stay bit-exact, and the surface must equal the int64
    correlation."""
    from numpy.lib.stride_tricks import sliding_window_view
    sp = _scaled(pair, fm, fs, left)
    p, got, ref, surf, (a1, a2, means) = _run(xg, orc, sp, 32, 32, 0, nxl=4, nyl=2)
    _exact(got, ref)
    m1f, m2f = np.float32(means[0]), np.float32(means[1])
    i1 = (a1 - m1f).astype(np.int64)
    i2 = _template_ints(a2, m2f, 32)
    print("max |v| master %d template %d" % (np.abs(i1).max(), np.abs(i2[32:96, 32:96]).max()))
    b = i2[32:96, 32:96]
    win = sliding_window_view(i1[:127, :127], (64, 64))
    num = np.einsum("ijkl,kl->ij", win, b)
    d1 = np.einsum("ijkl,ijkl->ij", win, win)
    d2 = int((b * b).sum())
    den = np.sqrt(d1.astype(np.float64) * float(d2))
    brute = np.where(den == 0, 0.0, 100.0 * np.abs(num / np.where(den == 0, 1.0, den)))
    assert np.array_equal(surf, brute)


def _brute_surface(i1, i2, xs):
    """The -time surface of one window from int64 sums, one lag row at a time
    (the sums are exact; the quotient in float64 as the NCC forms it)."""
    from numpy.lib.stride_tricks import sliding_window_view
    n = 2 * xs
    b = i2[xs:xs + n, xs:xs + n]
    d2 = int((b * b).sum())
    out = np.zeros((n, n))
    for dy in range(n):
        win = sliding_window_view(i1[dy:dy + n, :2 * n - 1], (n, n))[0]      # (dx, y, x)
        num = np.einsum("jkl,kl->j", win, b)
        d1 = np.einsum("jkl,jkl->j", win, win)
        den = np.sqrt(d1.astype(np.float64) * float(d2))
        out[dy] = np.where(den == 0, 0.0, 100.0 * np.abs(num / np.where(den == 0, 1.0, den)))
    return out


@pytest.mark.parametrize("fm,fs,left", [(1, 1, False), (5, 1, False), (1, 5, False), (5, 5, True)],
                         ids=["small", "master_big", "template_big", "mixed"])
def test_time_xs64_matches_int64_bruteforce(xg, orc, pair, fm, fs, left):
    """-xsearch 64 -ysearch 64 -time (the reference's default search size, 256 x
    256 windows) runs on the int8 matrix cores as four lag quadrants of four
    sub-searches each: bit-exact against the restatement whatever limbs the
    values need, and its surface equal to the int64 correlation."""
    sp = _scaled(pair, fm, fs, left)
    p, got, ref, surf, (a1, a2, means) = _run(xg, orc, sp, 64, 64, 0, nxl=2, nyl=2)
    _exact(got, ref)
    m1f, m2f = np.float32(means[0]), np.float32(means[1])
    i1 = (a1 - m1f).astype(np.int64)
    i2 = np.zeros_like(i1)                       # the template only: strip lines hold the rest of no window
    i2[64:192, 64:192] = (a2[64:192, 64:192] - m2f).astype(np.int64)
    print("max |v| master %d template %d" % (np.abs(i1).max(), np.abs(i2).max()))
    assert np.array_equal(surf, _brute_surface(i1, i2, 64))


def test_time4_xs64_bit_exact(xg, orc, pair):
    """-time4 at the default search size (the int8 matrix-core search)."""
    _, got, ref, _, _ = _run(xg, orc, pair, 64, 64, 1, nxl=2, nyl=1)
    _exact(got, ref)


def _gatelli_surface(i1, i2, xs):
    """The -time4 surface of one window (do_time_int_xcorr.cu:103-144) from
    exact integer sums: with A = a^2, B = b^2 split as A = Ah 2^14 + Al, every
    partial sum of Ah B, Al B, Ah A and Al A stays below 2^63 in int64 and the
    totals are Python integers; float() rounds them to nearest as the 128-bit
    conversion does, then the ratio and 100 sqrt(2 g - 1) in float64."""
    import math
    from numpy.lib.stride_tricks import sliding_window_view
    n = 2 * xs
    A = i1 * i1
    B = i2[xs:xs + n, xs:xs + n] ** 2
    assert A.max() < 2 ** 31 and B.max() < 2 ** 31          # Ah B summed over 2^14 terms < 2^63
    b4 = (int(((B >> 14) * B).sum()) << 14) + int(((B & 16383) * B).sum())
    out = np.zeros((n, n))
    for dy in range(n):
        win = sliding_window_view(A[dy:dy + n, :2 * n - 1], (n, n))[0]      # (dx, y, x)
        wh, wl = win >> 14, win & 16383
        nh, nl = np.einsum("jkl,kl->j", wh, B), np.einsum("jkl,kl->j", wl, B)
        ah, al = np.einsum("jkl,jkl->j", wh, win), np.einsum("jkl,jkl->j", wl, win)
        for dx in range(n):
            num = (int(nh[dx]) << 14) + int(nl[dx])
            a4 = (int(ah[dx]) << 14) + int(al[dx])
            den = math.sqrt(float(a4) * float(b4))
            g = 0.0 if den == 0.0 else abs(float(num) / den)
            out[dy, dx] = 0.0 if g <= 0.5 else 100.0 * math.sqrt(g * 2.0 - 1.0)
    return out


@pytest.mark.parametrize("xs", [32, 64])
@pytest.mark.parametrize("fm,fs,left", [(1, 1, False), (0.25, 0.25, False), (0.25, 1, False), (1, 0.25, False),
                                        (0.25, 1, True), (8, 8, True), (0.25, 8, False)],
                         ids=["limb3_both", "limb3_none", "limb3_template", "limb3_master", "mixed",
                              "six_limbs_mixed", "six_limbs_template"])
def test_time4_matches_exact_sums(xg, orc, pair, xs, fm, fs, left):
    """-time4 on 128 x 128 and 256 x 256 windows runs on the int8 matrix cores
    (the squares as up to six 7-bit limbs: limb 3 only where a value leaves
    +-1448, limbs 4 and 5 only where one leaves +-16383): bit-exact against the
    restatement, and the surface equal to the one from exact integer sums."""
    sp = _scaled(pair, fm, fs, left)
    p, got, ref, surf, (a1, a2, means) = _run(xg, orc, sp, xs, xs, 1, nxl=2, nyl=2)
    _exact(got, ref)
    m1f, m2f = np.float32(means[0]), np.float32(means[1])
    i1 = (a1 - m1f).astype(np.int64)
    i2 = np.zeros_like(i1)                       # the template only: the time path leaves the rest unset
    i2[xs:3 * xs, xs:3 * xs] = (a2[xs:3 * xs, xs:3 * xs] - m2f).astype(np.int64)
    print("max |v| master %d template %d" % (np.abs(i1).max(), np.abs(i2).max()))
    assert np.array_equal(surf, _gatelli_surface(i1, i2, xs))


@pytest.mark.parametrize("xs", [32, 64])
@pytest.mark.parametrize("corr_flag", [0, 1], ids=["time", "time4"])
def test_time_values_beyond_limbs(xg, orc, pair, xs, corr_flag):
    """Float (-real) input whose window integers reach past 2^20, beyond the
    limbs (three of the values for -time, six of the squares for -time4): the
    matrix-core searches hand such a patch to their exact loop (int64 /
    128-bit), and the result stays bit-exact."""
    m, s = pair
    mf = np.ascontiguousarray(m[..., 0].astype(np.float32) * np.float32(1500.0))
    sf = np.ascontiguousarray(s[..., 0].astype(np.float32) * np.float32(1500.0))
    assert np.abs(mf).max() > 2 * (1 << 20) and np.abs(sf).max() > 2 * (1 << 20)
    p = xg.make_params(xs, xs, format=xg.FLOAT32, corr_flag=xg.TIME4 if corr_flag else xg.TIME)
    x, y = xg.get_locations(512, 768, 2, 2, xs, xs)
    w = xg.make_windows(x, y, p.npx, p.npy)
    with xg.Engine(p) as eng:
        eng.set_images(mf, sf)
        got = eng.correlate(w)
    op = orc.make_params(xs, xs, corr_flag=corr_flag)
    locs = np.stack([w["mx0"], w["my0"], w["sx0"], w["sy0"]], 1).astype(np.int32)
    ref = orc.run(op, mf, sf, locs, precision=32, fmt=1)
    _exact(got, ref)
