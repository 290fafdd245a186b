"""Learned offset proposal beside plain xcorr (include/xcorr_cnn.h).

``ProposalNet`` binds the CNN kernels of libxcorr_hip.so (BF16 matrix-core
training and inference).  ``Proposer`` is the proposal form of the xcorr path
for 256 x 256 windows (-xsearch 64 -ysearch 64):

1. 4 x 4 box-downsampled amplitude images of the pair (one pass);
2. the network predicts each window's offset and a confidence from the 64 x 64
   crops of the two downsampled images;
3. windows with confidence above ``tau`` are correlated with the search
   shrunk to xsearch = ysearch = 16 (32 x 32 lags, 64 x 64 windows) centred on
   the rounded prediction; above ``tau_skip`` the high-res refinement is also
   skipped; all other windows run plain xcorr (the fallback);
4. offsets are composed as print_results.c:37-38 does, with the prediction
   taking the place of the PRM shift.

Plain xcorr produces the training labels (``labels_from_plain``) and judges the
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
    torch device tensors; bf16 inputs are carried in int16 tensors."""

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

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == NPARAM
        self._check(self.lib.xcorr_cnn_set_params(self.h, p.ctypes.data_as(ctypes.c_void_p)), "set_params")

    def get_params(self):
        p = np.zeros(NPARAM, np.float32)
        self._check(self.lib.xcorr_cnn_get_params(self.h, p.ctypes.data_as(ctypes.c_void_p)), "get_params")
        return p

    def sync(self):
        self._check(self.lib.xcorr_cnn_sync(self.h), "sync")

    def box4(self, img, out=None):
        import torch
        ny, nx = img.shape[0], img.shape[1]
        fmt = CINT16 if img.dim() == 3 else 1
        if out is None:
            out = torch.empty(ny // DS, nx // DS, dtype=torch.float32, device=img.device)
        self._check(self.lib.xcorr_cnn_box4(self.h, _dp(img), nx, ny, fmt, _dp(out)), "box4")
        return out

    def crops(self, dsm, dss, corners, out=None):
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
        out = self._out(x0.shape[0], out, x0.device)
        self._check(self.lib.xcorr_cnn_infer(self.h, _dp(x0), x0.shape[0], _dp(out)), "infer")
        return out

    def forward(self, x0, out=None):
        out = self._out(x0.shape[0], out, x0.device)
        self._check(self.lib.xcorr_cnn_forward(self.h, _dp(x0), x0.shape[0], _dp(out)), "forward")
        return out

    def propose(self, dsm, dss, corners, out=None):
        n = corners.shape[0]
        out = self._out(n, out, dsm.device)
        rc = self.lib.xcorr_cnn_propose(self.h, _dp(dsm), dsm.shape[1], dsm.shape[0], _dp(dss), dss.shape[1],
                                        dss.shape[0], _dp(corners), n, _dp(out))
        self._check(rc, "propose")
        return out

    def gradient(self, x0, tgt, out=None):
        g = np.zeros(NPARAM, np.float32)
        loss = ctypes.c_float()
        rc = self.lib.xcorr_cnn_gradient(self.h, _dp(x0), _dp(tgt), x0.shape[0], ctypes.byref(loss),
                                         g.ctypes.data_as(ctypes.c_void_p), None if out is None else _dp(out))
        self._check(rc, "gradient")
        return float(loss.value), g

    def train_step(self, x0, tgt, lr, want_loss=False):
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
