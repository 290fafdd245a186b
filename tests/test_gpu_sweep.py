"""GPU parity over a seeded random sweep of the parameter space.

Each case draws xsearch / ysearch, the range interpolation factor, the
high-res factor (or -nointerp), the correlation kind, the input format, the
batch size and a window list mixing a regular grid, random windows, windows
across the image edges and per-window slave offsets, then compares every
result with the FP32 CPU restatement bit for bit.  The draws cover the strip
paths (square 128 / 256 / 512 windows, -time on 128), the register per-patch
kernels and the generic LDS kernels in one run, in combinations the
hand-written cases do not list.
"""
import numpy as np
import pytest

from conftest import gpu_present
from test_gpu_parity import _assert_bit_exact, _locs, _oracle_params

pytestmark = pytest.mark.gpu

NX, NY = 1024, 1536


@pytest.fixture(scope="module")
def pair(synth):
    if not gpu_present():
        pytest.skip("no GPU")
    return synth.constant_shift_pair(NX, NY, 4.1, -6.3, 0.75, seed=31, noise_seed=32)


def _case(seed):
    rng = np.random.default_rng(seed)
    xs = int(rng.choice([8, 16, 32, 64, 128], p=[0.15, 0.25, 0.3, 0.2, 0.1]))
    ys = xs if rng.random() < 0.6 else int(rng.choice([8, 16, 32, 64]))
    corr = int(rng.choice([2, 2, 2, 0, 1]))
    if corr != 2 and max(xs, ys) > 32:        # the full integer search is O(lags x template)
        xs, ys = min(xs, 32), min(ys, 32)
    ri = int(rng.choice([2, 2, 2, 1, 4]))
    interp = int(rng.random() < 0.85)
    ifc = int(rng.choice([2, 4, 8, 16, 32]))
    fmt = int(rng.random() < 0.2)
    batch = int(rng.choice([3, 16, 64]))
    return xs, ys, corr, ri, interp, ifc, fmt, batch, rng


def _windows(xg, p, rng):
    npx, npy = p.npx, p.npy
    x, y = xg.get_locations(NX, NY, 5, 6, p.xsearch, p.ysearch)
    grid = xg.make_windows(x, y, npx, npy, int(rng.integers(-5, 6)), int(rng.integers(-5, 6)))
    g = grid.view(np.int32).reshape(-1, 4)
    extra = []
    for _ in range(8):                        # random windows, some across the image edges
        mx0 = int(rng.integers(-npx // 2, NX - npx // 2))
        my0 = int(rng.integers(-npy // 2, NY - npy // 2))
        extra.append((mx0, my0, mx0 + int(rng.integers(-9, 10)), my0 + int(rng.integers(-9, 10))))
    w = np.concatenate([g, np.array(extra, np.int32)])
    return np.ascontiguousarray(w[rng.permutation(len(w))]).view(xg.WINDOW_DTYPE).reshape(-1)


@pytest.mark.parametrize("seed", range(20))
def test_random_configuration_bit_exact(xg, orc, pair, seed):
    xs, ys, corr, ri, interp, ifc, fmt, batch, rng = _case(seed)
    m, s = pair
    if fmt:
        m = np.ascontiguousarray(m[..., 0].astype(np.float32))
        s = np.ascontiguousarray(s[..., 1].astype(np.float32))
    p = xg.make_params(xs, ys, ri=ri, interp_flag=interp, interp_factor=ifc, corr_flag=corr,
                       format=xg.FLOAT32 if fmt else xg.CINT16)
    w = _windows(xg, p, rng)
    label = "seed %d: xs %d ys %d corr %d ri %d interp %d ifc %d fmt %d batch %d" % (
        seed, xs, ys, corr, ri, interp, ifc, fmt, batch)
    print(label, len(w), "windows")
    with xg.Engine(p, max_batch=batch) as eng:
        eng.set_images(m, s)
        got = eng.correlate(w)
    ref = orc.run(_oracle_params(orc, p), m, s, _locs(w), precision=32, fmt=fmt)
    _assert_bit_exact(got, ref, label)
