"""TEST INFRASTRUCTURE ONLY.  NumPy restatement of the learned offset proposal
network (the CNN the HIP kernels in csrc/cnn.hip train and run), used as the
checker of those kernels' forward pass and gradients.  Nothing shipped imports
it; only tests/ and bench.py's checks do.

The network has no counterpart in the reference (diego-romano/gmtsar-gpu has
no learned component); its definition is the one the engine documents:

  input   x0[64][64][2]: 4 x 4 box-downsampled amplitudes of the master and
          slave windows, each channel normalised to zero mean and unit
          variance, rounded to bfloat16
  conv l  z_l[oy][ox][co] = b_l[co] + sum_{tap, ci} bf16(W_l[tap][co][ci])
                             * x_{l-1}[oy*s_l + ky - 1][ox*s_l + kx - 1][ci]
          (tap = 3 ky + kx, zero padding), x_l = bf16(relu(z_l));
          l = 1..4: 2->32 stride 2, 32->32 stride 2, 32->32 stride 2, 32->32 stride 1
  head    f[c] = mean over the 8 x 8 pixels of x_4[.][.][c];
          o[j] = bh[j] + sum_c Wh[j][c] f[c]   (FP32 head weights)
  outputs o[0], o[1]: the slave-minus-master offset of the two crops in
          downsampled pixels (x, y); o[2]: confidence logit
  loss    mean over the batch of wr * (huber(o0 - tx) + huber(o1 - ty))
          + wc * bce(sigmoid(o2), tc)   (huber delta 1)

Backward keeps the kernels' rounding points: the gradient reaching each
conv's pre-activation is rounded to bfloat16 (it is an MFMA operand), every
sum is exact (float64 here, float32 accumulation on the GPU).

Pins (tests/test_cnn_oracle.py): with rounding off, the analytic gradients
match central finite differences; the convolution matches a direct loop and
scipy's 2-D correlation; bf16 rounding matches the round-to-nearest-even
definition on special values.
"""
import numpy as np

LAYERS = ((2, 2), (32, 2), (32, 2), (32, 1))   # (input channels, stride) per conv; 32 output channels each
COUT = 32
NPARAM = sum(9 * COUT * cin + COUT for cin, _ in LAYERS) + 3 * COUT + 3


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = np.isnan(a)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    r = (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    return np.where(nan, np.nan, r)


def unpack(params):
    """Split the flat FP32 parameter vector (the engine's layout): per conv
    W[9][32][cin] then b[32]; then Wh[3][32], bh[3]."""
    p = np.asarray(params, np.float64)
    assert p.size == NPARAM, p.size
    out, o = {}, 0
    for l, (cin, _) in enumerate(LAYERS):
        n = 9 * COUT * cin
        out["W%d" % (l + 1)] = p[o:o + n].reshape(9, COUT, cin)
        o += n
        out["b%d" % (l + 1)] = p[o:o + COUT]
        o += COUT
    out["Wh"] = p[o:o + 3 * COUT].reshape(3, COUT)
    o += 3 * COUT
    out["bh"] = p[o:o + 3]
    return out


def pack(d):
    parts = []
    for l in range(len(LAYERS)):
        parts += [d["W%d" % (l + 1)].ravel(), d["b%d" % (l + 1)].ravel()]
    parts += [d["Wh"].ravel(), d["bh"].ravel()]
    return np.concatenate(parts)


def init_params(seed=0):
    """He-uniform conv weights, zero biases, small head."""
    rng = np.random.default_rng(seed)
    d = {}
    for l, (cin, _) in enumerate(LAYERS):
        lim = np.sqrt(6.0 / (9 * cin))
        d["W%d" % (l + 1)] = rng.uniform(-lim, lim, (9, COUT, cin))
        d["b%d" % (l + 1)] = np.zeros(COUT)
    d["Wh"] = rng.uniform(-0.1, 0.1, (3, COUT))
    d["bh"] = np.zeros(3)
    return pack(d).astype(np.float32)


def _taps(xp, s, ho, wo):
    for ky in range(3):
        for kx in range(3):
            yield 3 * ky + kx, xp[:, ky:ky + s * ho:s, kx:kx + s * wo:s, :]


def conv(x, W, b, s):
    """z[b][oy][ox][co] for NHWC x, W[9][co][ci]."""
    B, H, Wd, C = x.shape
    ho, wo = (H - 1) // s + 1, (Wd - 1) // s + 1
    xp = np.zeros((B, H + 2, Wd + 2, C))
    xp[:, 1:-1, 1:-1, :] = x
    z = np.broadcast_to(b, (B, ho, wo, W.shape[1])).copy()
    for t, patch in _taps(xp, s, ho, wo):
        z += patch @ W[t].T
    return z


def forward(params, x0, rounding=True):
    """Activations and outputs.  x0: (B, 64, 64, 2) input values."""
    P = unpack(params)
    rnd = bf16 if rounding else (lambda v: np.asarray(v, np.float64))
    xs = [np.asarray(x0, np.float64)]
    for l, (cin, s) in enumerate(LAYERS):
        z = conv(xs[-1], rnd(P["W%d" % (l + 1)]), P["b%d" % (l + 1)], s)
        xs.append(rnd(np.maximum(z, 0.0)))
    f = xs[-1].mean(axis=(1, 2))
    o = f @ P["Wh"].T + P["bh"]
    return xs, f, o


def huber(r):
    a = np.abs(r)
    return np.where(a <= 1.0, 0.5 * r * r, a - 0.5)


def loss_terms(o, tgt):
    """tgt: (B, 5) = tx, ty, tc, wr, wc."""
    tx, ty, tc, wr, wc = (tgt[:, i] for i in range(5))
    q = o[:, 2]
    bce = np.maximum(q, 0) - q * tc + np.log1p(np.exp(-np.abs(q)))
    per = wr * (huber(o[:, 0] - tx) + huber(o[:, 1] - ty)) + wc * bce
    return per


def loss_and_grads(params, x0, tgt, rounding=True):
    """Mean loss and its gradient with respect to the flat parameters."""
    P = unpack(params)
    rnd = bf16 if rounding else (lambda v: np.asarray(v, np.float64))
    xs, f, o = forward(params, x0, rounding)
    B = o.shape[0]
    loss = float(loss_terms(o, tgt).mean())
    tx, ty, tc, wr, wc = (tgt[:, i] for i in range(5))
    do = np.zeros_like(o)
    do[:, 0] = wr * np.clip(o[:, 0] - tx, -1, 1) / B
    do[:, 1] = wr * np.clip(o[:, 1] - ty, -1, 1) / B
    do[:, 2] = wc * (1.0 / (1.0 + np.exp(-o[:, 2])) - tc) / B
    G = {"Wh": do.T @ f, "bh": do.sum(0)}
    df = do @ P["Wh"]
    hw = xs[-1].shape[1] * xs[-1].shape[2]
    dx = np.broadcast_to(df[:, None, None, :] / hw, xs[-1].shape).copy()
    for l in range(len(LAYERS) - 1, -1, -1):
        cin, s = LAYERS[l]
        xin, xout = xs[l], xs[l + 1]
        dz = rnd(dx * (xout > 0))
        B_, H, Wd, C = xin.shape
        ho, wo = xout.shape[1], xout.shape[2]
        xp = np.zeros((B_, H + 2, Wd + 2, C))
        xp[:, 1:-1, 1:-1, :] = xin
        Wl = rnd(P["W%d" % (l + 1)])
        dW = np.zeros_like(Wl)
        dxp = np.zeros_like(xp)
        for t, patch in _taps(xp, s, ho, wo):
            dW[t] = np.einsum("bhwo,bhwc->oc", dz, patch)
            ky, kx = divmod(t, 3)
            dxp[:, ky:ky + s * ho:s, kx:kx + s * wo:s, :] += dz @ Wl[t]
        G["W%d" % (l + 1)] = dW
        G["b%d" % (l + 1)] = dz.sum(axis=(0, 1, 2))
        dx = dxp[:, 1:-1, 1:-1, :]
    return loss, pack(G), o


def box4(amp):
    """4 x 4 box mean of an amplitude image (lines, samples); trailing partial
    blocks are dropped."""
    a = np.asarray(amp, np.float64)
    ny, nx = a.shape[0] // 4, a.shape[1] // 4
    return a[:4 * ny, :4 * nx].reshape(ny, 4, nx, 4).mean(axis=(1, 3))


def crops(ds_m, ds_s, corners):
    """Normalised 64 x 64 crops (B, 64, 64, 2) at downsampled corners
    (mx, my, sx, sy); samples outside an image read as zero."""
    out = np.zeros((len(corners), 64, 64, 2))
    for i, (mx, my, sx, sy) in enumerate(corners):
        for ch, (img, x0, y0) in enumerate(((ds_m, mx, my), (ds_s, sx, sy))):
            c = np.zeros((64, 64))
            ya, yb = max(y0, 0), min(y0 + 64, img.shape[0])
            xa, xb = max(x0, 0), min(x0 + 64, img.shape[1])
            if ya < yb and xa < xb:
                c[ya - y0:yb - y0, xa - x0:xb - x0] = img[ya:yb, xa:xb]
            mu = c.mean()
            sd = np.sqrt(max(((c - mu) ** 2).mean(), 0.0))
            out[i, :, :, ch] = (c - mu) / max(sd, 1e-6)
    return out
