"""Learned offset proposal beside plain xcorr (include/xcorr_cnn.h).

``ProposalNet`` binds the CNN kernels of libxcorr_hip.so (BF16 matrix-core
training and inference).  ``Proposer`` is the proposal form of the xcorr path
for 256 x 256 windows (-xsearch 64 -ysearch 64):

1. 4 x 4 box-downsampled amplitude images of the pair (one pass);
2. the network predicts each window's offset and a confidence from the 64 x 64
   crops of the two downsampled images;
3. windows with confidence above ``tau`` are correlated with the search
   shrunk to xsearch = ysearch = ``xs_small`` (32: 64 x 64 lags, 128 x 128
   windows; 16: 32 x 32 lags, 64 x 64 windows) centred on the rounded
   prediction; above ``tau_skip`` the high-res refinement is also skipped; all
   other windows run plain xcorr (the fallback);
4. offsets are composed as print_results.c:37-38 does, with the prediction
   taking the place of the PRM shift.

Plain xcorr produces the training labels (train_proposal.py) and judges the
proposal form's error.  The reference has no learned component; this module
has no counterpart there.
"""
import ctypes
import os

import numpy as np

from . import (CINT16, DEVICE_PTRS, HOST_WINDOWS, NO_SYNC, RESULT_DTYPE, WINDOW_DTYPE, Engine, XcorrError,
               load_library, make_params, make_windows)

NPARAM = 28451
DS = 4            # box-downsampling factor
CROP = 64         # crop size in downsampled pixels (256 / DS)


def _lib():
    lib = load_library()
    if getattr(lib, "_cnn_bound", False):
        return lib
    P, I = ctypes.c_void_p, ctypes.c_int
    lib.xcorr_cnn_create.argtypes = [I, I, ctypes.POINTER(P)]
    lib.xcorr_cnn_destroy.argtypes = [P]
    lib.xcorr_cnn_destroy.restype = None
    lib.xcorr_cnn_last_error.argtypes = [P]
    lib.xcorr_cnn_last_error.restype = ctypes.c_char_p
    lib.xcorr_cnn_stream.argtypes = [P]
    lib.xcorr_cnn_stream.restype = P
    lib.xcorr_cnn_sync.argtypes = [P]
    lib.xcorr_cnn_set_stream.argtypes = [P, P]
    lib.xcorr_cnn_set_params.argtypes = [P, P]
    lib.xcorr_cnn_get_params.argtypes = [P, P]
    lib.xcorr_cnn_box4.argtypes = [P, P, I, I, I, P]
    lib.xcorr_cnn_crops.argtypes = [P, P, I, I, P, I, I, P, I, P]
    lib.xcorr_cnn_infer.argtypes = [P, P, I, P]
    lib.xcorr_cnn_forward.argtypes = [P, P, I, P]
    lib.xcorr_cnn_propose.argtypes = [P, P, I, I, P, I, I, P, I, P]
    lib.xcorr_cnn_gradient.argtypes = [P, P, P, I, P, P, P]
    lib.xcorr_cnn_train_step.argtypes = [P, P, P, I, ctypes.c_float, P]
    lib._cnn_bound = True
    return lib


def _dp(t):
    return ctypes.c_void_p(t.data_ptr())


class ProposalNet:
    """One network context (xcorr_cnn_create ... xcorr_cnn_destroy).  Tensors are
    torch device tensors; bf16 inputs are carried in int16 tensors.  Every call
    queues its kernels on torch's current stream, so they are ordered with the
    torch work that produces their inputs and reads their outputs."""

    def __init__(self, max_batch=512, device=-1):
        self.lib = _lib()
        h = ctypes.c_void_p()
        rc = self.lib.xcorr_cnn_create(int(device), int(max_batch), ctypes.byref(h))
        if rc != 0:
            raise XcorrError("xcorr_cnn_create failed (%d): %s" % (rc, self.lib.xcorr_cnn_last_error(None).decode()))
        self.h = h
        self.max_batch = max_batch

    def _check(self, rc, what):
        if rc < 0:
            raise XcorrError("%s failed (%d): %s" % (what, rc, self.lib.xcorr_cnn_last_error(self.h).decode()))
        return rc

    def _on_torch_stream(self):
        import torch
        self._check(self.lib.xcorr_cnn_set_stream(self.h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
                    "set_stream")

    def set_params(self, params):
        self._on_torch_stream()
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == NPARAM
        self._check(self.lib.xcorr_cnn_set_params(self.h, p.ctypes.data_as(ctypes.c_void_p)), "set_params")

    def get_params(self):
        self._on_torch_stream()
        p = np.zeros(NPARAM, np.float32)
        self._check(self.lib.xcorr_cnn_get_params(self.h, p.ctypes.data_as(ctypes.c_void_p)), "get_params")
        return p

    def sync(self):
        self._on_torch_stream()
        self._check(self.lib.xcorr_cnn_sync(self.h), "sync")

    def box4(self, img, out=None):
        self._on_torch_stream()
        import torch
        ny, nx = img.shape[0], img.shape[1]
        fmt = CINT16 if img.dim() == 3 else 1
        if out is None:
            out = torch.empty(ny // DS, nx // DS, dtype=torch.float32, device=img.device)
        self._check(self.lib.xcorr_cnn_box4(self.h, _dp(img), nx, ny, fmt, _dp(out)), "box4")
        return out

    def crops(self, dsm, dss, corners, out=None):
        self._on_torch_stream()
        import torch
        n = corners.shape[0]
        if out is None:
            out = torch.empty(n, CROP, CROP, 2, dtype=torch.int16, device=dsm.device)
        rc = self.lib.xcorr_cnn_crops(self.h, _dp(dsm), dsm.shape[1], dsm.shape[0], _dp(dss), dss.shape[1],
                                      dss.shape[0], _dp(corners), n, _dp(out))
        self._check(rc, "crops")
        return out

    def _out(self, n, out, device):
        import torch
        return torch.empty(n, 3, dtype=torch.float32, device=device) if out is None else out

    def infer(self, x0, out=None):
        self._on_torch_stream()
        out = self._out(x0.shape[0], out, x0.device)
        self._check(self.lib.xcorr_cnn_infer(self.h, _dp(x0), x0.shape[0], _dp(out)), "infer")
        return out

    def forward(self, x0, out=None):
        self._on_torch_stream()
        out = self._out(x0.shape[0], out, x0.device)
        self._check(self.lib.xcorr_cnn_forward(self.h, _dp(x0), x0.shape[0], _dp(out)), "forward")
        return out

    def propose(self, dsm, dss, corners, out=None):
        self._on_torch_stream()
        n = corners.shape[0]
        out = self._out(n, out, dsm.device)
        rc = self.lib.xcorr_cnn_propose(self.h, _dp(dsm), dsm.shape[1], dsm.shape[0], _dp(dss), dss.shape[1],
                                        dss.shape[0], _dp(corners), n, _dp(out))
        self._check(rc, "propose")
        return out

    def gradient(self, x0, tgt, out=None):
        self._on_torch_stream()
        g = np.zeros(NPARAM, np.float32)
        loss = ctypes.c_float()
        rc = self.lib.xcorr_cnn_gradient(self.h, _dp(x0), _dp(tgt), x0.shape[0], ctypes.byref(loss),
                                         g.ctypes.data_as(ctypes.c_void_p), None if out is None else _dp(out))
        self._check(rc, "gradient")
        return float(loss.value), g

    def train_step(self, x0, tgt, lr, want_loss=False):
        self._on_torch_stream()
        loss = ctypes.c_float()
        rc = self.lib.xcorr_cnn_train_step(self.h, _dp(x0), _dp(tgt), x0.shape[0], float(lr),
                                           ctypes.byref(loss) if want_loss else None)
        self._check(rc, "train_step")
        return float(loss.value) if want_loss else None

    def close(self):
        if getattr(self, "h", None):
            self.lib.xcorr_cnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def corners_for(windows):
    """Downsampled crop corners (mx, my, sx, sy) of 256 x 256 windows, and the
    full-resolution offset of the slave crop from the master crop (x, y)."""
    w = np.asarray(windows)
    c = np.stack([w["mx0"] // DS, w["my0"] // DS, w["sx0"] // DS, w["sy0"] // DS], 1).astype(np.int32)
    return c, (DS * (c[:, 2] - c[:, 0])).astype(np.float64), (DS * (c[:, 3] - c[:, 1])).astype(np.float64)


def sigmoid(v):
    return 1.0 / (1.0 + np.exp(-np.asarray(v, np.float64)))


def compose(res, windows, ri):
    """dx, dy of xcorr results: the lag offsets plus the slave window's offset
    from the master window (print_results.c:37-38 with rshift = sx0 - mx0 and
    ashift + ishft = sy0 - my0; float32 arithmetic)."""
    f32 = np.float32
    w = np.asarray(windows)
    dx = res["xoff"] / f32(ri) - res["xfrac"] / f32(ri) + (w["sx0"] - w["mx0"]).astype(f32)
    dy = (res["yoff"] - res["yfrac"]) + (w["sy0"] - w["my0"]).astype(f32)
    return dx.astype(f32), dy.astype(f32)


def shrunk_windows(windows, rx, ry, xs_small=16, npx=256):
    """Windows of the shrunken search: xsearch = ysearch = xs_small (4 xs_small
    square) centred on the same location as the npx x npx window, the slave
    window displaced by the integer prediction (rx, ry)."""
    w = np.asarray(windows)
    c = (npx - 4 * xs_small) // 2
    out = np.zeros(w.shape[0], WINDOW_DTYPE)
    out["mx0"] = w["mx0"] + c
    out["my0"] = w["my0"] + c
    out["sx0"] = out["mx0"] + np.asarray(rx, np.int64)
    out["sy0"] = out["my0"] + np.asarray(ry, np.int64)
    return out


class Proposer:
    """Proposal form of the xcorr path for 256 x 256 windows on one image pair
    (torch device tensors).  ``run`` returns the composed offsets, correlation
    and the path every window took (0 plain fallback, 1 shrunken search with
    high-res, 2 shrunken search without high-res, 3 shrunken search rejected by
    the verification -- peak on the border of the shrunken lag window or
    correlation below corr_min -- and redone with the plain search)."""

    def __init__(self, params, tau=0.9, tau_skip=1.01, max_batch=8192, device=0, interp_factor=16, xs_small=32,
                 verify=False, corr_min=0.0):
        import torch
        self.torch = torch
        self.net = ProposalNet(max_batch=512, device=device)
        self.net.set_params(params)
        self.tau, self.tau_skip, self.xs_small = tau, tau_skip, xs_small
        self.verify, self.corr_min = verify, corr_min
        self.full = Engine(make_params(64, 64, interp_factor=interp_factor), max_batch=max_batch, device=device)
        self.small = Engine(make_params(xs_small, xs_small, interp_factor=interp_factor), max_batch=max_batch,
                            device=device)
        self.small_ni = Engine(make_params(xs_small, xs_small, interp_flag=0), max_batch=max_batch, device=device)
        self.dev = torch.device("cuda", device if device >= 0 else torch.cuda.current_device())

    def set_images(self, master, slave):
        for e in (self.full, self.small, self.small_ni):
            e.set_images(master, slave)
        self.dsm = self.net.box4(master)
        self.dss = self.net.box4(slave)

    def predict(self, windows):
        """CNN predictions: full-resolution offsets (dx, dy) and confidence."""
        torch = self.torch
        corners, offx, offy = corners_for(windows)
        cd = torch.from_numpy(corners).to(self.dev)
        o = self.net.propose(self.dsm, self.dss, cd).cpu().numpy().astype(np.float64)
        return DS * o[:, 0] + offx, DS * o[:, 1] + offy, sigmoid(o[:, 2])

    def run(self, windows, pred=None):
        torch = self.torch
        w = np.asarray(windows)
        n = w.shape[0]
        if pred is None:
            pred = self.predict(w)
        pdx, pdy, conf = pred
        path = np.zeros(n, np.int8)
        path[conf > self.tau] = 1
        path[conf > self.tau_skip] = 2
        rx, ry = np.rint(pdx).astype(np.int64), np.rint(pdy).astype(np.int64)
        out = np.zeros(n, RESULT_DTYPE)
        dx = np.zeros(n, np.float32)
        dy = np.zeros(n, np.float32)
        jobs = []
        for k, eng, ri in ((0, self.full, 2), (1, self.small, 2), (2, self.small_ni, 2)):
            idx = np.nonzero(path == k)[0]
            if idx.size == 0:
                continue
            ww = w[idx] if k == 0 else shrunk_windows(w[idx], rx[idx], ry[idx], self.xs_small)
            res = torch.empty(idx.size * RESULT_DTYPE.itemsize, dtype=torch.uint8, device=self.dev)
            eng.correlate(ww, out=res, flags=NO_SYNC)        # the three engines' streams run concurrently
            jobs.append((idx, ww, res, eng, ri))
        for idx, ww, res, eng, ri in jobs:
            eng.sync()
            r = res.cpu().numpy().view(RESULT_DTYPE)
            out[idx] = r
            dx[idx], dy[idx] = compose(r, ww, ri)
        if self.verify:
            # a shrunken search whose integer peak lies on the border of its lag
            # window (the true peak may lie outside it) or whose correlation is
            # below corr_min is redone with the plain search (path 3)
            lag = 2 * self.xs_small
            sh = path > 0
            bad = sh & ((out["ip"] <= 0) | (out["ip"] >= lag - 1) | (out["jp"] <= 0) | (out["jp"] >= lag - 1) |
                        (out["corr"] < self.corr_min))
            idx = np.nonzero(bad)[0]
            if idx.size:
                r = self.full.correlate(w[idx])
                out[idx] = r
                dx[idx], dy[idx] = compose(r, w[idx], 2)
                path[idx] = 3
        return dx, dy, out, path

    def close(self):
        for e in (self.full, self.small, self.small_ni):
            e.close()
        self.net.close()
