"""GPU: the proposal network's BF16 matrix-core kernels against the NumPy
restatement (oracle/cnn_oracle.py).

* Forward outputs of the layer-wise training kernels and of the fused
  inference kernel agree with the restatement (same bf16 rounding points,
  FP32 vs exact accumulation), and with each other bit for bit.
* Loss and gradients of every parameter group agree with the restatement's
  backward pass (relative error of each group's gradient < 2 %, cosine >
  0.999), so a wrong tap, stride, transpose or lane layout fails.
* Box-downsampling and normalised crops agree with their closed forms.
* Adam steps on the GPU drive the loss of a fixed batch down, and two steps
  move every parameter as a NumPy Adam (beta 0.9 / 0.999, eps 1e-8, bias
  correction) computes from the GPU's own gradients.
"""
import numpy as np
import pytest

from conftest import gpu_present

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env(xg):
    if not gpu_present():
        pytest.skip("no GPU")
    import importlib

    import torch
    import cnn_oracle
    prop = importlib.import_module("gmtsar_gpu_amd.proposal")
    return torch, prop, cnn_oracle


def _to_bf16_bits(x):
    import cnn_oracle
    v = cnn_oracle.bf16(x).astype(np.float32)
    return (v.view(np.uint32) >> 16).astype(np.uint16).view(np.int16)


def _params(co, seed):
    rng = np.random.default_rng(seed)
    p = co.unpack(co.init_params(seed).astype(np.float64))
    for l in range(1, 5):
        p["b%d" % l] = rng.uniform(0.0, 0.05, 32)
    p["Wh"] = rng.uniform(-0.3, 0.3, (3, 32))
    p["bh"] = rng.uniform(-0.1, 0.1, 3)
    return co.pack(p).astype(np.float32)


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, 64, 64, 2)) * 1.3


def test_forward_matches_oracle_and_fused_equals_layerwise(env):
    torch, prop, co = env
    n = 37
    params = _params(co, 1)
    x = _inputs(n, 2)
    xb = _to_bf16_bits(x)
    xv = co.bf16(x)
    with prop.ProposalNet(max_batch=16) as net:        # 37 samples -> 3 layer-wise batches
        net.set_params(params)
        x0 = torch.from_numpy(np.ascontiguousarray(xb)).cuda()
        o_lw = net.forward(x0).cpu().numpy()
        o_fu = net.infer(x0).cpu().numpy()
    _, _, o_ref = co.forward(params.astype(np.float64), xv)
    assert np.array_equal(o_lw.view(np.uint32), o_fu.view(np.uint32))
    err = np.abs(o_lw - o_ref)
    print("forward max abs err", err.max(), "max |o|", np.abs(o_ref).max())
    assert err.max() <= 2e-3 * (1.0 + np.abs(o_ref).max())


def test_gradients_match_oracle(env):
    torch, prop, co = env
    n = 16
    params = _params(co, 3)
    x = _inputs(n, 4)
    rng = np.random.default_rng(5)
    tgt = np.stack([rng.uniform(-3, 3, n), rng.uniform(-3, 3, n), rng.integers(0, 2, n),
                    rng.uniform(0.5, 1.0, n), rng.uniform(0.2, 1.0, n)], 1).astype(np.float32)
    with prop.ProposalNet(max_batch=n) as net:
        net.set_params(params)
        x0 = torch.from_numpy(np.ascontiguousarray(_to_bf16_bits(x))).cuda()
        loss, g = net.gradient(x0, torch.from_numpy(tgt).cuda())
    loss_ref, g_ref, _ = co.loss_and_grads(params.astype(np.float64), co.bf16(x), tgt.astype(np.float64))
    assert abs(loss - loss_ref) <= 1e-3 * abs(loss_ref)
    P, R = co.unpack(g.astype(np.float64)), co.unpack(g_ref)
    for k in P:
        a, b = P[k].ravel(), R[k].ravel()
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        cos = a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30)
        print("grad %s: rel %.2e cos %.6f |g| %.3e" % (k, rel, cos, np.linalg.norm(b)))
        assert rel < 2e-2 and cos > 0.999, k


def test_box4_and_crops_match_oracle(env, synth):
    torch, prop, co = env
    m, s = synth.constant_shift_pair(600, 900, 5.0, -7.0, 0.8, seed=21, noise_seed=22)
    mt, st = torch.from_numpy(m).cuda(), torch.from_numpy(s).cuda()
    corners = np.array([[10, 20, 12, 18], [-5, -8, -3, -10], [140, 200, 150, 190], [0, 0, 0, 0],
                        [100, 150, 101, 151]], np.int32)
    with prop.ProposalNet(max_batch=8) as net:
        net.set_params(_params(co, 6))
        dm, ds = net.box4(mt), net.box4(st)
        cr = net.crops(dm, ds, torch.from_numpy(corners).cuda())
        o1 = net.propose(dm, ds, torch.from_numpy(corners).cuda())
        o2 = net.infer(cr)
        dmh, crh = dm.cpu().numpy(), cr.cpu().numpy()
    amp = np.hypot(m[..., 0].astype(np.float64), m[..., 1].astype(np.float64))
    ref = co.box4(amp)
    np.testing.assert_allclose(dmh, ref, rtol=1e-5)
    refc = co.crops(dmh.astype(np.float64), ds.cpu().numpy().astype(np.float64), corners)
    got = (crh.view(np.uint16).astype(np.uint32) << 16).view(np.float32)
    np.testing.assert_allclose(got, co.bf16(refc), rtol=2 ** -7, atol=1e-3)
    o1h = o1.cpu().numpy()
    assert np.array_equal(o1h.view(np.uint32), o2.cpu().numpy().view(np.uint32))
    assert np.all(o1h != 0.0)


def test_adam_steps_reduce_loss(env):
    torch, prop, co = env
    n = 64
    x = _inputs(n, 8)
    rng = np.random.default_rng(9)
    tgt = np.stack([rng.uniform(-2, 2, n), rng.uniform(-2, 2, n), rng.integers(0, 2, n), np.ones(n),
                    np.ones(n)], 1).astype(np.float32)
    with prop.ProposalNet(max_batch=n) as net:
        net.set_params(co.init_params(11))
        x0 = torch.from_numpy(np.ascontiguousarray(_to_bf16_bits(x))).cuda()
        t = torch.from_numpy(tgt).cuda()
        first = net.train_step(x0, t, 2e-3, want_loss=True)
        for _ in range(300):
            net.train_step(x0, t, 2e-3)
        last = net.train_step(x0, t, 2e-3, want_loss=True)
        p = net.get_params()
    print("loss %.4f -> %.4f" % (first, last))
    assert np.all(np.isfinite(p))
    assert last < 0.5 * first


def test_adam_update_matches_numpy(env):
    torch, prop, co = env
    n = 32
    x = _inputs(n, 12)
    rng = np.random.default_rng(13)
    tgt = np.stack([rng.uniform(-2, 2, n), rng.uniform(-2, 2, n), rng.integers(0, 2, n), np.ones(n),
                    np.ones(n)], 1).astype(np.float32)
    lr = 1e-3
    with prop.ProposalNet(max_batch=n) as net:
        p0 = co.init_params(17)
        net.set_params(p0)
        x0 = torch.from_numpy(np.ascontiguousarray(_to_bf16_bits(x))).cuda()
        t = torch.from_numpy(tgt).cuda()
        _, g1 = net.gradient(x0, t)
        net.train_step(x0, t, lr)
        p1 = net.get_params()
        _, g2 = net.gradient(x0, t)
        net.train_step(x0, t, lr)
        p2 = net.get_params()
    f = np.float64
    m = v = np.zeros(p0.size)
    want = [p0.astype(f)]
    for step, g in ((1, g1.astype(f)), (2, g2.astype(f))):
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        want.append(want[-1] - lr * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8))
    # the weight gradients accumulate with FP32 atomics, so the gradient a step
    # uses may differ from the one fetched before it in the last bits: compare
    # the updates where the gradient is well away from zero
    for k, (got, prev) in enumerate(((p1, p0), (p2, p1))):
        upd, ref = got.astype(f) - prev.astype(f), want[k + 1] - want[k]
        sel = (np.abs(g1) > 1e-4 * np.abs(g1).max()) & (np.abs(g2) > 1e-4 * np.abs(g2).max())
        assert sel.sum() > 0.8 * p0.size
        # plus one float32 ulp of the parameter: the stored parameter is rounded
        tol = 2e-3 * np.abs(ref) + np.spacing(np.abs(got)).astype(f)
        bad = np.nonzero(sel & (np.abs(upd - ref) > tol))[0]
        assert bad.size == 0, "step %d: %s gpu %s numpy %s" % (k + 1, bad[:4], upd[bad[:4]], ref[bad[:4]])
