"""GPU: launch tuning and multi-pair graph runs (through the C ABI).

* Tuning knobs that must not change a result bit (batch, LDS tiles, threads
  per block) give results identical to the default launch.
* A non-default FFT radix plan is bit-exact against the FP32 restatement run
  with the same plan (oracle.fft_plans).
* The autotuner returns valid settings and, without radix tuning, results
  identical to the untuned engine.
* A multi-pair workload (mixed window sizes, interpolation factors, a time-
  domain class, the 256 x 256 strip class) captured as one HIP graph gives,
  replay after replay, exactly the results of one engine per pair.
* The matrix-core time searches (-time and -time4 on 128 x 128 and 256 x 256
  windows, all on strip lines) as classes of one graph: every replay equals
  one engine per pair and the FP32 restatement.
"""
import numpy as np
import pytest

from conftest import gpu_present

pytestmark = pytest.mark.gpu

FIELDS = ("xoff", "yoff", "corr", "xfrac", "yfrac", "cmax", "ip", "jp", "hip", "hjp", "cave")


@pytest.fixture(scope="module")
def torch():
    if not gpu_present():
        pytest.skip("no GPU")
    import torch as t
    return t


@pytest.fixture(scope="module")
def pairs(synth):
    if not gpu_present():
        pytest.skip("no GPU")
    return {
        "a": synth.constant_shift_pair(640, 896, 1.75, -2.5, 0.85, seed=11, noise_seed=12),
        "b": synth.constant_shift_pair(1024, 1280, -3.25, 4.5, 0.7, seed=13, noise_seed=14),
        "c": synth.constant_shift_pair(1536, 1792, 5.5, -6.75, 0.9, seed=15, noise_seed=16),
    }


def _windows(xg, p, m_nx, m_ny, nxl, nyl):
    x, y = xg.get_locations(m_nx, m_ny, nxl, nyl, p.xsearch, p.ysearch)
    return xg.make_windows(x, y, p.npx, p.npy)


def _locs(w):
    return np.stack([w["mx0"], w["my0"], w["sx0"], w["sy0"]], 1).astype(np.int32)


def _oracle_params(orc, p):
    return orc.make_params(p.xsearch, p.ysearch, ri=p.ri, interp_flag=p.interp_flag,
                           interp_factor=p.interp_factor, corr_flag=p.corr_flag)


def _same(got, ref, label, fields=FIELDS):
    for f in fields:
        a, b = got[f], ref[f]
        if a.dtype.kind == "f":
            eq = (a.view(np.uint32 if a.itemsize == 4 else np.uint64) ==
                  b.view(np.uint32 if b.itemsize == 4 else np.uint64)) | (np.isnan(a) & np.isnan(b))
        else:
            eq = a == b
        bad = np.nonzero(~eq)[0]
        assert bad.size == 0, "%s: %s differs at %s: %s vs %s" % (label, f, bad[:4], a[bad[:4]], b[bad[:4]])


NEUTRAL = [
    # xsearch, ysearch, settings
    (32, 32, dict(threads=128)),
    (32, 32, dict(threads=512, rows_per_block=4, cols_per_block=16)),
    (32, 32, dict(rows_per_block=16, cols_per_block=2, batch=7)),
    (16, 32, dict(threads=64, rows_per_block=1, cols_per_block=16)),
    (128, 128, dict(threads=512, rows_per_block=1, cols_per_block=4, batch=5)),
]


@pytest.mark.parametrize("case", NEUTRAL, ids=lambda c: "xs%d_ys%d_%s" % (c[0], c[1], "_".join(
    "%s%d" % (k[:4], v) for k, v in sorted(c[2].items()))))
def test_neutral_tuning_is_bit_identical(xg, pairs, case):
    xs, ys, kw = case
    m, s = pairs["c"]
    p = xg.make_params(xs, ys, interp_factor=8)
    w = _windows(xg, p, 1536, 1792, 5, 4)
    with xg.Engine(p, max_batch=16) as eng:
        eng.set_images(m, s)
        base = eng.correlate(w)
        t = eng.set_tuning(**kw)
        for k, v in kw.items():
            assert getattr(t, k) == v
        got = eng.correlate(w)
    print("tuned", kw, "->", t.as_dict())
    _same(got, base, "tuning %s" % kw)


RADIX = [
    # xsearch, ysearch, radix_x, radix_y
    (32, 32, [8, 16], [8, 16]),
    (32, 32, [2, 8, 8], [2, 8, 8]),
    (32, 32, [4, 4, 8], [4, 4, 8]),
    (16, 32, [4, 16], [4, 4, 8]),
    (128, 128, [16, 16, 2], [16, 16, 2]),
    (64, 32, [4, 4, 16], [16, 8]),
]


@pytest.mark.parametrize("case", RADIX, ids=lambda c: "xs%d_ys%d_x%s_y%s" % (
    c[0], c[1], "x".join(map(str, c[2])), "x".join(map(str, c[3]))))
def test_radix_plan_bit_exact_vs_oracle_with_same_plan(xg, orc, pairs, case):
    xs, ys, rx, ry = case
    m, s = pairs["c"]
    p = xg.make_params(xs, ys, interp_factor=16)
    w = _windows(xg, p, 1536, 1792, 4, 3)
    with xg.Engine(p, max_batch=16) as eng:
        eng.set_images(m, s)
        eng.set_tuning(radix_x=rx, radix_y=ry)
        got = eng.correlate(w)
    plans = {p.npx: rx}
    plans[p.npy] = ry
    with orc.fft_plans(plans):
        ref = orc.run(_oracle_params(orc, p), m, s, _locs(w), precision=32)
    _same(got, ref, "radix %s/%s" % (rx, ry), fields=FIELDS[:-1])
    np.testing.assert_allclose(got["cave"], ref["cave"], rtol=1e-12)
    # a different rounding, same peaks on these well-correlated pairs
    refd = orc.run(_oracle_params(orc, p), m, s, _locs(w), precision=32)
    assert np.array_equal(got["ip"], refd["ip"]) and np.array_equal(got["jp"], refd["jp"])


def test_set_tuning_rejects_invalid(xg, pairs):
    m, s = pairs["a"]
    with xg.Engine(xg.make_params(32, 32), max_batch=8) as eng:
        for kw in (dict(threads=100), dict(threads=1024), dict(batch=9), dict(batch=0),
                   dict(rows_per_block=3), dict(rows_per_block=128), dict(radix_x=[16, 16], radix_y=[16, 16]),
                   dict(radix_x=[8, 16], radix_y=[16, 8]), dict(radix_x=[32, 4], radix_y=[32, 4])):
            with pytest.raises(xg.XcorrError):
                eng.set_tuning(**kw)
    with xg.Engine(xg.make_params(64, 64), max_batch=8) as eng:     # strip shape: any valid plan, any batch
        eng.set_tuning(radix_x=[4, 4, 16], radix_y=[4, 4, 16])
        eng.set_tuning(batch=4)
        with pytest.raises(xg.XcorrError):
            eng.set_tuning(batch=9)


def test_autotune_keeps_results(xg, orc, pairs):
    m, s = pairs["b"]
    p = xg.make_params(16, 32, interp_factor=8)      # not a strip shape: every launch setting is tuned
    w = _windows(xg, p, 1024, 1280, 10, 8)
    with xg.Engine(p, max_batch=64) as eng:
        eng.set_images(m, s)
        base = eng.correlate(w)
        t = eng.autotune(w)
        d = t.as_dict()
        print("autotune:", d)
        assert d["candidates"] >= 5 and d["ms_per_batch"] > 0
        assert 1 <= d["batch"] <= 64 and d["threads"] % 64 == 0
        assert d["radix_x"] == [8, 8] and d["radix_y"] == [16, 8]
        got = eng.correlate(w)
    _same(got, base, "autotuned")
    p = xg.make_params(32, 32, interp_factor=8)      # strip shape: radix plans (per-patch kernels) vs strips, batch
    w = _windows(xg, p, 1024, 1280, 10, 8)
    with xg.Engine(p, max_batch=64) as eng:
        eng.set_images(m, s)
        base = eng.correlate(w)
        d = eng.autotune(w, radix=True).as_dict()
        print("autotune (strips, radix plans tried):", d)
        # radix plans were timed against the strips (each on the per-patch kernels)
        assert d["candidates"] >= 4
        got = eng.correlate(w)
    if d["radix_x"] == [16, 8]:                     # the strips won: bit-identical to the untuned run
        _same(got, base, "autotuned strips")
    else:                                           # a plan won: bit-exact against the restatement with it
        with orc.fft_plans({p.npx: d["radix_x"]}):
            ref = orc.run(_oracle_params(orc, p), m, s, _locs(w), precision=32)
        _same(got, ref, "autotuned plan %s" % d["radix_x"], fields=FIELDS[:-1])


def _multi_workload(xg, torch, pairs):
    """Six pairs over three image pairs: strips (xs 64), xs 32 with ifc 8 and 16,
    xs 16 x 32, a -time pair and a 512 x 512 (xs 128) pair."""
    dev = torch.device("cuda:0")
    imgs = {k: (torch.from_numpy(v[0]).to(dev), torch.from_numpy(v[1]).to(dev), v) for k, v in pairs.items()}
    spec = [("c", 64, 64, 8, xg.FREQ, 6, 5), ("a", 32, 32, 8, xg.FREQ, 7, 9), ("b", 32, 32, 16, xg.FREQ, 9, 6),
            ("b", 16, 32, 16, xg.FREQ, 8, 7), ("a", 16, 16, 8, xg.TIME, 5, 6), ("c", 128, 128, 16, xg.FREQ, 3, 2),
            ("a", 32, 32, 8, xg.FREQ, 4, 5)]
    out = []
    for key, xs, ys, ifc, cf, nxl, nyl in spec:
        m_t, s_t, (m_h, s_h) = imgs[key]
        p = xg.make_params(xs, ys, interp_factor=ifc, corr_flag=cf)
        w = _windows(xg, p, m_h.shape[1], m_h.shape[0], nxl, nyl)
        res = torch.zeros(len(w) * xg.RESULT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
        out.append(dict(params=p, master=m_t, slave=s_t, windows=w, out=res, host=(m_h, s_h)))
    return out


def test_multi_graph_matches_single_engines(xg, torch, pairs):
    work = _multi_workload(xg, torch, pairs)
    expect = []
    for q in work:
        with xg.Engine(q["params"], max_batch=16) as eng:
            eng.set_images(*q["host"])
            expect.append(eng.correlate(q["windows"]))
    with xg.Multi(work) as mu:
        classes, pc = mu.classes()
        assert len(classes) == 6 and pc[1] == pc[6] and len(set(pc.tolist())) == 6
        for rep in range(3):
            for q in work:
                q["out"].fill_(0xff)
            torch.cuda.synchronize()
            mu.run()
            torch.cuda.synchronize()
            for i, q in enumerate(work):
                got = q["out"].cpu().numpy().view(xg.RESULT_DTYPE)
                _same(got, expect[i], "replay %d pair %d" % (rep, i))
            # eager work between replays
            torch.cuda.synchronize()


def test_multi_tuned_matches_oracle(xg, orc, torch, pairs):
    work = _multi_workload(xg, torch, pairs)
    with xg.Multi(work, tune=True, tune_radix=True) as mu:
        mu.run()
        torch.cuda.synchronize()
        classes, pc = mu.classes()
        for k, (p, t) in enumerate(classes):
            print("class xs%d ys%d ifc%d corr%d:" % (p.xsearch, p.ysearch, p.interp_factor, p.corr_flag), t)
        for i, q in enumerate(work):
            p = q["params"]
            _, t = classes[pc[i]]
            got = q["out"].cpu().numpy().view(xg.RESULT_DTYPE)
            plans = {p.npx: t["radix_x"]}
            plans[p.npy] = t["radix_y"]
            m_h, s_h = q["host"]
            with orc.fft_plans(plans):
                ref = orc.run(_oracle_params(orc, p), m_h, s_h, _locs(q["windows"]), precision=32)
            _same(got, ref, "tuned pair %d %s" % (i, t), fields=FIELDS[:-1])
            np.testing.assert_allclose(got["cave"], ref["cave"], rtol=1e-12)


def test_multi_time_classes_match_engines_and_oracle(xg, orc, torch, pairs):
    dev = torch.device("cuda:0")
    spec = [("c", 64, 64, 16, xg.TIME, 3, 3), ("b", 32, 32, 8, xg.TIME4, 5, 4), ("c", 64, 64, 16, xg.TIME4, 2, 3),
            ("a", 32, 32, 8, xg.TIME, 4, 5)]
    work = []
    for key, xs, ys, ifc, cf, nxl, nyl in spec:
        m_h, s_h = pairs[key]
        p = xg.make_params(xs, ys, interp_factor=ifc, corr_flag=cf)
        w = _windows(xg, p, m_h.shape[1], m_h.shape[0], nxl, nyl)
        res = torch.zeros(len(w) * xg.RESULT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
        work.append(dict(params=p, master=torch.from_numpy(m_h).to(dev), slave=torch.from_numpy(s_h).to(dev),
                         windows=w, out=res, host=(m_h, s_h)))
    expect = []
    for q in work:
        with xg.Engine(q["params"], max_batch=4) as eng:
            eng.set_images(*q["host"])
            expect.append(eng.correlate(q["windows"]))
        ref = orc.run(_oracle_params(orc, q["params"]), *q["host"], _locs(q["windows"]), precision=32)
        _same(expect[-1], ref, "engine vs restatement, class xs%d corr%d" % (q["params"].xsearch,
                                                                          q["params"].corr_flag), fields=FIELDS[:-1])
    with xg.Multi(work) as mu:
        classes, pc = mu.classes()
        assert len(classes) == 4
        for rep in range(3):
            for q in work:
                q["out"].fill_(0xff)
            torch.cuda.synchronize()
            mu.run()
            torch.cuda.synchronize()
            for i, q in enumerate(work):
                got = q["out"].cpu().numpy().view(xg.RESULT_DTYPE)
                _same(got, expect[i], "replay %d pair %d" % (rep, i))
