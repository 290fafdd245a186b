"""GPU parity of streamed images (xcorr_hip_correlate_stream): the SLC lines
reach the device band by band through a reader callback, two band buffers in
flight.  Results must equal the in-memory path bit for bit, in input order, for
every path (strips, per-patch frequency, time search), with many small bands,
windows across the image edges and slave line offsets."""
import numpy as np
import pytest

from conftest import gpu_present
from test_gpu_parity import _assert_bit_exact

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pair(synth):
    if not gpu_present():
        pytest.skip("no GPU")
    return synth.constant_shift_pair(768, 1536, 1.73, -4.2, 0.85, seed=21, noise_seed=22)


def _grid_windows(xg, p, nxl, nyl, ashift=0, rshift=0):
    x, y = xg.get_locations(768, 1536, nxl, nyl, p.xsearch, p.ysearch)
    return xg.make_windows(x, y, p.npx, p.npy, rshift, ashift)


def _shuffled_with_edges(xg, w, p):
    rng = np.random.default_rng(5)
    extra = np.array([(-50, -30, -47, -20), (700, 1400, 703, 1390), (100, 1500, 100, 1510),
                      (300, -300, 300, -290)], np.int32)
    allw = np.concatenate([w.view(np.int32).reshape(-1, 4), extra])
    return np.ascontiguousarray(allw[rng.permutation(len(allw))]).view(xg.WINDOW_DTYPE).reshape(-1)


@pytest.mark.parametrize("xs,ys,corr,band", [(64, 64, 2, 300), (64, 64, 2, 700), (32, 32, 2, 150), (16, 16, 0, 80),
                                             (128, 128, 2, 600), (32, 32, 0, 150)],
                         ids=["strips_b300", "strips_b700", "strips32_b150", "time16_b80", "strips128_b600",
                              "time_strips32_b150"])
def test_stream_equals_resident(xg, pair, xs, ys, corr, band):
    m, s = pair
    p = xg.make_params(xs, ys, corr_flag=corr)
    w = _shuffled_with_edges(xg, _grid_windows(xg, p, 6, 12, ashift=7, rshift=-3), p)
    with xg.Engine(p, max_batch=37) as eng:
        eng.set_images(m, s)
        ref = eng.correlate(w)
        got = eng.correlate_stream(m, s, w, band_lines=band)
        again = eng.correlate(w)                     # the resident images are kept
    _assert_bit_exact(got, ref, "streamed, band %d" % band)
    _assert_bit_exact(again, ref, "resident after streaming")


def test_stream_default_band_and_memmap(xg, pair, tmp_path):
    m, s = pair
    mf, sf = tmp_path / "m.SLC", tmp_path / "s.SLC"
    m.tofile(mf)
    s.tofile(sf)
    mm = np.memmap(mf, dtype=np.int16, mode="r", shape=m.shape)
    sm = np.memmap(sf, dtype=np.int16, mode="r", shape=s.shape)
    p = xg.make_params(64, 64)
    w = _grid_windows(xg, p, 8, 20)
    with xg.Engine(p) as eng:
        eng.set_images(m, s)
        ref = eng.correlate(w)
        got = eng.correlate_stream(mm, sm, w)
    _assert_bit_exact(got, ref, "memmap, default bands")


def test_stream_rejects_window_taller_than_band(xg, pair):
    m, s = pair
    p = xg.make_params(64, 64)
    w = np.array([(100, 100, 100, 200)], np.int32).view(xg.WINDOW_DTYPE).reshape(-1)
    with xg.Engine(p) as eng:
        with pytest.raises(xg.XcorrError, match="spans more lines"):
            eng.correlate_stream(m, s, w, band_lines=300)     # 256 + 100 lines > 300
