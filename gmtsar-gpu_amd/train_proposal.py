#!/usr/bin/env python3
"""Train the proposal network on the GPU from plain-xcorr labels.

Data: synthetic complex-int16 pairs (speckle model of synth.py) with offset
fields |d| <= 32 px and coherence per 512 x 512 block uniform in [0.2, 0.95];
256 x 256 windows (-xsearch 64 -ysearch 64) on a 64-pixel grid; plain xcorr
on the GPU labels every window (dx, dy, corr).  Training runs entirely on the
project's HIP kernels (xcorr_cnn_train_step: BF16 matrix-core forward and
backward, Adam):

  stage 1  offset regression (huber on the crop offset in downsampled pixels),
           windows weighted by plain xcorr's correlation;
  stage 2  confidence: every window is re-correlated with the shrunken search
           (xsearch 32) centred on the network's rounded prediction; the label
           is 1 where that gives plain xcorr's dx and dy within `agree` px;
           training continues on both terms.

Writes the FP32 parameters (numpy .npy, 28451 values) and a JSON report with
the loss curve and held-out accuracy.

    python3 train_proposal.py --out weights/proposal_cnn.npy --report weights/train_report.json
"""
import argparse
import importlib.util
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _package():
    if "gmtsar_gpu_amd" in sys.modules:
        return sys.modules["gmtsar_gpu_amd"]
    spec = importlib.util.spec_from_file_location("gmtsar_gpu_amd", os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gmtsar_gpu_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def init_params(seed):
    """He-uniform conv weights, zero biases, small head (the layout of
    include/xcorr_cnn.h)."""
    rng = np.random.default_rng(seed)
    parts = []
    for cin in (2, 32, 32, 32):
        lim = math.sqrt(6.0 / (9 * cin))
        parts += [rng.uniform(-lim, lim, 9 * 32 * cin), np.zeros(32)]
    parts += [rng.uniform(-0.1, 0.1, 96), np.zeros(3)]
    return np.concatenate(parts).astype(np.float32)


def make_dataset(xg, prop, synth, torch, net, n_pairs, rng, dev, nx=4096, ny=8192, step=64, seed0=1000):
    """Windows, plain-xcorr results, crops (bf16 bits) and crop offsets of
    n_pairs synthetic pairs."""
    p = xg.make_params(64, 64)
    eng = xg.Engine(p, max_batch=8192)
    xs, labels, offs, wins, pairs = [], [], [], [], []
    for k in range(n_pairs):
        # offset field: linear in x (range) and y (azimuth), |d| <= 30 over the image
        ax, bx = rng.uniform(-30, 30), rng.uniform(-30, 30)
        ay, by = rng.uniform(-30, 30), rng.uniform(-30, 30)
        m, s = synth.strip_pair_torch(nx, ny, lambda x, a=ax, b=bx: a + (b - a) * x / nx,
                                      lambda y, a=ay, b=by: a + (b - a) * y / ny, coh_range=(0.2, 0.95),
                                      seed=seed0 + k, device=dev)
        torch.cuda.synchronize()                  # the engine's stream reads what torch's stream wrote
        gx = np.arange(0, nx - 256 + 1, step)
        gy = np.arange(0, ny - 256 + 1, step)
        yy, xx = np.meshgrid(gy, gx, indexing="ij")
        w = np.zeros(xx.size, xg.WINDOW_DTYPE)
        w["mx0"], w["my0"] = xx.ravel(), yy.ravel()
        w["sx0"], w["sy0"] = w["mx0"], w["my0"]
        eng.set_images(m, s)
        res = eng.correlate(w)
        dx, dy = prop.compose(res, w, 2)
        dsm, dss = net.box4(m), net.box4(s)
        corners, ox, oy = prop.corners_for(w)
        x0 = net.crops(dsm, dss, torch.from_numpy(corners).to(dev))
        xs.append(x0)
        labels.append(np.stack([dx, dy, res["corr"]], 1).astype(np.float64))
        offs.append(np.stack([ox, oy], 1))
        wins.append(w)
        pairs.append((m, s, dsm, dss))
        torch.cuda.synchronize()
    eng.close()
    return xs, labels, offs, wins, pairs


def targets(labels, offs, conf=None, wc=0.0):
    dx, dy, corr = labels[:, 0], labels[:, 1], labels[:, 2]
    t = np.zeros((len(dx), 5), np.float32)
    t[:, 0] = (dx - offs[:, 0]) / 4.0
    t[:, 1] = (dy - offs[:, 1]) / 4.0
    t[:, 2] = 0.0 if conf is None else conf
    t[:, 3] = np.clip(corr / 30.0, 0.05, 1.0)          # trust plain xcorr's labels by its correlation
    t[:, 4] = wc
    return t


def predict(net, x0, offs, torch):
    o = net.infer(x0).cpu().numpy().astype(np.float64)
    return 4.0 * o[:, 0] + offs[:, 0], 4.0 * o[:, 1] + offs[:, 1], 1.0 / (1.0 + np.exp(-o[:, 2]))


def agreement(xg, prop, torch, pairs, wins, preds, labels, agree, xs_small=16):
    """1 where the shrunken search at the rounded prediction reproduces plain
    xcorr's dx and dy within `agree` px."""
    eng = xg.Engine(xg.make_params(xs_small, xs_small), max_batch=8192)
    out = []
    for (m, s, _, _), w, (pdx, pdy, _), lab in zip(pairs, wins, preds, labels):
        ww = prop.shrunk_windows(w, np.rint(pdx), np.rint(pdy), xs_small)
        eng.set_images(m, s)
        r = eng.correlate(ww)
        dx, dy = prop.compose(r, ww, 2)
        out.append(((np.abs(dx - lab[:, 0]) <= agree) & (np.abs(dy - lab[:, 1]) <= agree)).astype(np.float32))
    eng.close()
    return out


def train(net, x0, tgt, steps, lr0, batch, rng, torch, dev, curve, tag):
    n = x0.shape[0]
    t_dev = torch.from_numpy(tgt).to(dev)
    xb = torch.empty((batch,) + tuple(x0.shape[1:]), dtype=x0.dtype, device=dev)
    tb = torch.empty(batch, 5, dtype=torch.float32, device=dev)
    acc = []
    for it in range(steps):
        idx = torch.from_numpy(rng.integers(0, n, batch)).to(dev)
        torch.index_select(x0, 0, idx, out=xb)
        torch.index_select(t_dev, 0, idx, out=tb)
        lr = lr0 * 0.5 * (1.0 + math.cos(math.pi * it / steps))
        want = it % 25 == 0 or it == steps - 1
        loss = net.train_step(xb, tb, max(lr, 1e-6), want_loss=want)
        if want:
            acc.append(loss)
        if it % 500 == 0 or it == steps - 1:
            p = net.get_params()
            curve.append([tag, it, round(float(np.mean(acc)), 5)])
            log("%s step %d lr %.2e loss(avg) %.4f |W| %.2f" % (tag, it, lr, np.mean(acc), np.linalg.norm(p)))
            acc = []


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "weights", "proposal_cnn.npy"))
    ap.add_argument("--report", default="")
    ap.add_argument("--pairs", type=int, default=30, help="training pairs of 4096 x 8192 (6720 windows each)")
    ap.add_argument("--val-pairs", type=int, default=3)
    ap.add_argument("--steps1", type=int, default=12000)
    ap.add_argument("--steps2", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--agree", type=float, default=0.125, help="confidence label tolerance (px)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--xs-small", type=int, default=32, help="xsearch of the shrunken search")
    ap.add_argument("--diag", action="store_true",
                    help="also report how often the shrunken search agrees given plain xcorr's own offsets")
    args = ap.parse_args()

    import torch
    xg = _package()
    import importlib
    prop = importlib.import_module("gmtsar_gpu_amd.proposal")
    synth = importlib.import_module("gmtsar_gpu_amd.synth")
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(args.seed)
    t0 = time.time()
    net = prop.ProposalNet(max_batch=args.batch)
    net.set_params(init_params(args.seed))
    xs, labels, offs, wins, pairs = make_dataset(xg, prop, synth, torch, net, args.pairs + args.val_pairs, rng, dev)
    tr, va = slice(0, args.pairs), slice(args.pairs, args.pairs + args.val_pairs)
    x_tr = torch.cat(xs[tr])
    lab_tr, off_tr = np.concatenate(labels[tr]), np.concatenate(offs[tr])
    log("dataset: %d training and %d held-out windows in %.1f s" % (
        x_tr.shape[0], sum(len(l) for l in labels[va]), time.time() - t0))
    curve = []
    if args.diag:
        for xsm in (16, 32):
            perfect = [(l[:, 0], l[:, 1], None) for l in labels[va]]
            ag = np.concatenate(agreement(xg, prop, torch, pairs[va], wins[va], perfect, labels[va], args.agree, xsm))
            good = np.concatenate([l[:, 2] for l in labels[va]])
            log("diag: shrunken search xsearch %d at plain's own offsets agrees within %.3f px on %.1f %% of "
                "held-out windows (%.1f %% of those with plain corr > 30)"
                % (xsm, args.agree, 100 * ag.mean(), 100 * ag[good > 30].mean()))
    t1 = time.time()
    train(net, x_tr, targets(lab_tr, off_tr), args.steps1, args.lr, args.batch, rng, torch, dev, curve, "regression")
    torch.cuda.synchronize()
    t_stage1 = time.time() - t1
    # confidence labels from the shrunken search at the current predictions
    preds = [predict(net, x, o, torch) for x, o in zip(xs, offs)]
    agree = agreement(xg, prop, torch, pairs, wins, preds, labels, args.agree, args.xs_small)
    conf_tr = np.concatenate(agree[tr])
    log("stage 2: %.1f %% of training windows agree with plain xcorr at the prediction" % (100 * conf_tr.mean()))
    t2 = time.time()
    train(net, x_tr, targets(lab_tr, off_tr, conf_tr, 1.0), args.steps2, args.lr * 0.5, args.batch, rng, torch,
          dev, curve, "confidence")
    torch.cuda.synchronize()
    t_stage2 = time.time() - t2
    params = net.get_params()
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, params)
    # held-out accuracy
    rep = {"training_windows": int(x_tr.shape[0]), "steps": [args.steps1, args.steps2], "batch": args.batch,
           "train_seconds": [round(t_stage1, 2), round(t_stage2, 2)], "curve": curve,
           "agree_px": args.agree, "train_agree_fraction": round(float(conf_tr.mean()), 4)}
    vp = [predict(net, x, o, torch) for x, o in zip(xs[va], offs[va])]
    va_agree = agreement(xg, prop, torch, pairs[va], wins[va], vp, labels[va], args.agree, args.xs_small)
    pdx = np.concatenate([v[0] for v in vp])
    pdy = np.concatenate([v[1] for v in vp])
    pc = np.concatenate([v[2] for v in vp])
    lab = np.concatenate(labels[va])
    ag = np.concatenate(va_agree)
    good = lab[:, 2] > 20
    rep["heldout"] = {
        "windows": int(len(pc)),
        "median_abs_err_px": [round(float(np.median(np.abs(pdx - lab[:, 0])[good])), 3),
                              round(float(np.median(np.abs(pdy - lab[:, 1])[good])), 3)],
        "within_4px": round(float(np.mean((np.abs(pdx - lab[:, 0]) < 4) & (np.abs(pdy - lab[:, 1]) < 8))), 4),
        "agree_fraction": round(float(ag.mean()), 4),
    }
    for tau in (0.5, 0.7, 0.9):
        sel = pc > tau
        rep["heldout"]["tau_%.1f" % tau] = {"confident": round(float(sel.mean()), 4),
                                            "agree_among_confident": round(float(ag[sel].mean()) if sel.any() else 0.0, 4)}
    log(json.dumps(rep["heldout"]))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(rep, f, indent=1)
    print(json.dumps({"weights": args.out, "heldout": rep["heldout"], "train_seconds": rep["train_seconds"]}),
          flush=True)
    net.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
