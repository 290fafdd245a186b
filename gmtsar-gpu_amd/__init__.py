"""MI355X-native xcorr offset tracking (GMTSAR ``xcorr`` hot path) -- Python host binding.

The engine is the HIP library ``libxcorr_hip.so`` built from ``csrc/`` for gfx950;
its C ABI is ``include/xcorr_hip.h``.  This module binds it with ctypes and
mirrors the reference's driver-level interface (xcorr.cu:299-378): parameters as
``struct xcorr`` names them, one call per batch of locations, results as
``struct locs`` fields.  There is no CPU fallback: if the library is missing or
no GPU is present, ``Engine`` raises.

Loaded by path (the directory name is not an identifier), e.g.::

    spec = importlib.util.spec_from_file_location("gmtsar_gpu_amd", ".../gmtsar-gpu_amd/__init__.py")
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libxcorr_hip.so")
HOST_LIB_PATH = os.path.join(HERE, "libxcorr_host.so")

TIME, TIME4, FREQ = 0, 1, 2
CINT16, FLOAT32 = 0, 1
DEVICE_PTRS, NO_SYNC, TIMING, HOST_WINDOWS, SERIAL, KEEP_SURFACE = 1, 2, 4, 8, 64, 128
TUNE, TUNE_RADIX = 16, 32

EXPORTED_SYMBOLS = (
    "xcorr_hip_abi_version", "xcorr_hip_default_params", "xcorr_hip_create", "xcorr_hip_set_images",
    "xcorr_hip_correlate", "xcorr_hip_sync", "xcorr_hip_stream", "xcorr_hip_kernel_times",
    "xcorr_hip_dump", "xcorr_hip_destroy", "xcorr_hip_last_error", "xcorr_hip_correlate_stream",
    "xcorr_hip_get_tuning", "xcorr_hip_set_tuning", "xcorr_hip_autotune", "xcorr_hip_multi_create",
    "xcorr_hip_multi_run", "xcorr_hip_multi_sync", "xcorr_hip_multi_stream", "xcorr_hip_multi_classes",
    "xcorr_hip_multi_destroy", "xcorr_hip_multi_last_error", "xcorr_hip_correlate_noalloc",
)

# int reader(void *user, int image, int first, int count, void *dst)
LINE_READER = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_void_p)


class Params(ctypes.Structure):
    """``xcorr_hip_params`` (field names of struct xcorr, xcorr_d.h:23-77)."""
    _fields_ = [(n, ctypes.c_int32) for n in (
        "xsearch", "ysearch", "nx_corr", "ny_corr", "npx", "npy", "nxc", "nyc", "ri",
        "interp_flag", "interp_factor", "n2x", "n2y", "corr_flag", "format")]


class Tuning(ctypes.Structure):
    """``xcorr_hip_tuning``: launch settings of a context."""
    _fields_ = [("batch", ctypes.c_int32), ("rows_per_block", ctypes.c_int32), ("cols_per_block", ctypes.c_int32),
                ("threads", ctypes.c_int32), ("radix_x", ctypes.c_int32 * 4), ("radix_y", ctypes.c_int32 * 4),
                ("ms_per_batch", ctypes.c_float), ("candidates", ctypes.c_int32)]

    def as_dict(self):
        rx = [int(v) for v in self.radix_x if v]
        ry = [int(v) for v in self.radix_y if v]
        return {"batch": self.batch, "rows_per_block": self.rows_per_block, "cols_per_block": self.cols_per_block,
                "threads": self.threads, "radix_x": rx, "radix_y": ry,
                "ms_per_batch": round(float(self.ms_per_batch), 4), "candidates": self.candidates}


class Pair(ctypes.Structure):
    """``xcorr_hip_pair``: one SLC pair of a multi-pair workload."""
    _fields_ = [("params", Params), ("master", ctypes.c_void_p), ("m_nx", ctypes.c_int32), ("m_ny", ctypes.c_int32),
                ("slave", ctypes.c_void_p), ("s_nx", ctypes.c_int32), ("s_ny", ctypes.c_int32),
                ("win", ctypes.c_void_p), ("n", ctypes.c_int32), ("out", ctypes.c_void_p)]


WINDOW_DTYPE = np.dtype([("mx0", "<i4"), ("my0", "<i4"), ("sx0", "<i4"), ("sy0", "<i4")])
RESULT_DTYPE = np.dtype([("cave", "<f8"), ("xoff", "<f4"), ("yoff", "<f4"), ("corr", "<f4"),
                         ("xfrac", "<f4"), ("yfrac", "<f4"), ("cmax", "<f4"), ("ip", "<i4"),
                         ("jp", "<i4"), ("hip", "<i4"), ("hjp", "<i4")])

_lib = None


def library_path():
    return LIB_PATH


def load_library():
    """Load libxcorr_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libxcorr_hip.so is not built (run __graft_entry__.build() or make -C gmtsar-gpu_amd)")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    lib.xcorr_hip_abi_version.restype = ctypes.c_int
    lib.xcorr_hip_default_params.argtypes = [ctypes.POINTER(Params), ctypes.c_int, ctypes.c_int]
    lib.xcorr_hip_default_params.restype = None
    lib.xcorr_hip_create.argtypes = [ctypes.POINTER(Params), ctypes.c_int, ctypes.c_int, ctypes.POINTER(P)]
    lib.xcorr_hip_set_images.argtypes = [P, P, ctypes.c_int, ctypes.c_int, P, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_uint]
    lib.xcorr_hip_correlate.argtypes = [P, P, ctypes.c_int, P, ctypes.c_uint]
    lib.xcorr_hip_correlate_noalloc.argtypes = [P, P, ctypes.c_int, P, ctypes.c_uint]
    lib.xcorr_hip_sync.argtypes = [P]
    lib.xcorr_hip_stream.argtypes = [P]
    lib.xcorr_hip_stream.restype = P
    lib.xcorr_hip_kernel_times.argtypes = [P, ctypes.POINTER(ctypes.c_char_p), P, P, ctypes.c_int]
    lib.xcorr_hip_dump.argtypes = [P, ctypes.c_int, ctypes.c_int, P, ctypes.c_size_t]
    lib.xcorr_hip_dump.restype = ctypes.c_long
    lib.xcorr_hip_destroy.argtypes = [P]
    lib.xcorr_hip_destroy.restype = None
    lib.xcorr_hip_correlate_stream.argtypes = [P, LINE_READER, P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_int, P, ctypes.c_int, P]
    lib.xcorr_hip_last_error.argtypes = [P]
    lib.xcorr_hip_last_error.restype = ctypes.c_char_p
    T = ctypes.POINTER(Tuning)
    lib.xcorr_hip_get_tuning.argtypes = [P, T]
    lib.xcorr_hip_set_tuning.argtypes = [P, T]
    lib.xcorr_hip_autotune.argtypes = [P, P, ctypes.c_int, ctypes.c_uint, T]
    lib.xcorr_hip_multi_create.argtypes = [ctypes.POINTER(Pair), ctypes.c_int, ctypes.c_int, ctypes.c_uint,
                                           ctypes.POINTER(P)]
    lib.xcorr_hip_multi_run.argtypes = [P, ctypes.c_uint]
    lib.xcorr_hip_multi_sync.argtypes = [P]
    lib.xcorr_hip_multi_stream.argtypes = [P]
    lib.xcorr_hip_multi_stream.restype = P
    lib.xcorr_hip_multi_classes.argtypes = [P, ctypes.POINTER(Params), T, P, ctypes.c_int]
    lib.xcorr_hip_multi_destroy.argtypes = [P]
    lib.xcorr_hip_multi_destroy.restype = None
    lib.xcorr_hip_multi_last_error.argtypes = [P]
    lib.xcorr_hip_multi_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def make_params(xsearch=64, ysearch=64, ri=2, interp_flag=1, interp_factor=16, corr_flag=FREQ,
                format=CINT16):
    """Parameters with the reference's derived sizes (parse_xcorr_input.c:139-158)
    and defaults (set_defaults, parse_xcorr_input.c:38-88)."""
    p = Params()
    load_library().xcorr_hip_default_params(ctypes.byref(p), xsearch, ysearch)
    p.ri, p.interp_flag, p.interp_factor = ri, interp_flag, interp_factor
    p.corr_flag, p.format = corr_flag, format
    return p


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _as_windows(windows):
    w = np.ascontiguousarray(windows)
    if w.dtype != WINDOW_DTYPE:
        w = np.ascontiguousarray(w.astype(np.int32).reshape(-1, 4)).view(WINDOW_DTYPE).reshape(-1)
    return w


def _is_torch(x):
    return type(x).__module__.startswith("torch")


class XcorrError(RuntimeError):
    pass


class Engine:
    """One engine context (xcorr_hip_create ... xcorr_hip_destroy)."""

    def __init__(self, params, max_batch=0, device=-1):
        self.lib = load_library()
        self.params = params
        h = ctypes.c_void_p()
        rc = self.lib.xcorr_hip_create(ctypes.byref(params), int(max_batch), int(device), ctypes.byref(h))
        if rc != 0:
            raise XcorrError("xcorr_hip_create failed (%d): %s" % (rc, self.lib.xcorr_hip_last_error(None).decode()))
        self.h = h
        self._keep = []

    def _check(self, rc, what):
        if rc < 0:
            raise XcorrError("%s failed (%d): %s" % (what, rc, self.lib.xcorr_hip_last_error(self.h).decode()))
        return rc

    def set_images(self, master, slave):
        """master/slave: numpy (host copy) or torch device tensors shaped (ny, nx, 2) int16
        or (ny, nx) float32."""
        dev = _is_torch(master)
        if dev:
            m, s = master.contiguous(), slave.contiguous()
            self._keep = [m, s]
            mp, sp = ctypes.c_void_p(m.data_ptr()), ctypes.c_void_p(s.data_ptr())
            shape_m, shape_s = tuple(m.shape), tuple(s.shape)
        else:
            m, s = np.ascontiguousarray(master), np.ascontiguousarray(slave)
            mp, sp = _ptr(m), _ptr(s)
            shape_m, shape_s = m.shape, s.shape
        rc = self.lib.xcorr_hip_set_images(self.h, mp, shape_m[1], shape_m[0], sp, shape_s[1], shape_s[0],
                                           DEVICE_PTRS if dev else 0)
        self._check(rc, "xcorr_hip_set_images")

    def correlate_stream(self, master, slave, windows, band_lines=0):
        """Correlate with the images streamed in bands of lines.  master / slave
        are host arrays (numpy, e.g. np.memmap of the SLC files) of shape
        (lines, samples[, 2]) with 4-byte samples; they are read line band by
        line band through a callback."""
        imgs = [np.ascontiguousarray(a).reshape(np.shape(a)[0], -1).view(np.uint32) for a in (master, slave)]
        err = []

        def reader(user, image, first, count, dst):
            try:
                src = imgs[image]
                nx = src.shape[1]
                out = np.ctypeslib.as_array(ctypes.cast(dst, ctypes.POINTER(ctypes.c_uint32)), (count, nx))
                out[:] = src[first:first + count]
                return 0
            except Exception as e:          # reported through the engine's error path
                err.append(e)
                return 1

        cb = LINE_READER(reader)
        w = _as_windows(windows)
        res = np.zeros(w.shape[0], RESULT_DTYPE)
        (m_ny, m_nx), (s_ny, s_nx) = imgs[0].shape, imgs[1].shape
        rc = self.lib.xcorr_hip_correlate_stream(self.h, cb, None, m_nx, m_ny, s_nx, s_ny, int(band_lines), _ptr(w),
                                                 w.shape[0], _ptr(res))
        if err:
            raise err[0]
        self._check(rc, "xcorr_hip_correlate_stream")
        return res

    def correlate(self, windows, out=None, flags=0):
        """windows: numpy WINDOW_DTYPE array (or (n,4) int32) -> numpy RESULT_DTYPE array.
        With torch device tensors for windows/out the call stays on the device; host
        windows with a device `out` tensor write the results on the device."""
        if _is_torch(out) and not _is_torch(windows):
            w = _as_windows(windows)
            self._keep = w            # read until the queued work has consumed it
            rc = self.lib.xcorr_hip_correlate(self.h, _ptr(w), w.shape[0], ctypes.c_void_p(out.data_ptr()),
                                              flags | DEVICE_PTRS | HOST_WINDOWS)
            self._check(rc, "xcorr_hip_correlate")
            return out
        if _is_torch(windows):
            n = windows.shape[0]
            rc = self.lib.xcorr_hip_correlate(self.h, ctypes.c_void_p(windows.data_ptr()), n,
                                              ctypes.c_void_p(out.data_ptr()), flags | DEVICE_PTRS)
            self._check(rc, "xcorr_hip_correlate")
            return out
        w = _as_windows(windows)
        n = w.shape[0]
        res = np.zeros(n, RESULT_DTYPE) if out is None else out
        rc = self.lib.xcorr_hip_correlate(self.h, _ptr(w), n, _ptr(res), flags)
        self._check(rc, "xcorr_hip_correlate")
        return res

    def correlate_noalloc(self, windows, out, flags=0):
        """xcorr_hip_correlate_noalloc: windows (n x 4 int32) and out (n result
        records as bytes) are torch device tensors; nothing is allocated."""
        n = windows.shape[0]
        rc = self.lib.xcorr_hip_correlate_noalloc(self.h, ctypes.c_void_p(windows.data_ptr()), n,
                                                  ctypes.c_void_p(out.data_ptr()), flags)
        self._check(rc, "xcorr_hip_correlate_noalloc")
        return out

    def sync(self):
        self._check(self.lib.xcorr_hip_sync(self.h), "xcorr_hip_sync")

    def get_tuning(self):
        t = Tuning()
        self._check(self.lib.xcorr_hip_get_tuning(self.h, ctypes.byref(t)), "xcorr_hip_get_tuning")
        return t

    def set_tuning(self, **kw):
        """Change launch settings (batch, rows_per_block, cols_per_block, threads,
        radix_x / radix_y as lists of radices); unnamed ones are kept."""
        t = self.get_tuning()
        for k, v in kw.items():
            if k in ("radix_x", "radix_y"):
                arr = getattr(t, k)
                for i in range(4):
                    arr[i] = v[i] if v is not None and i < len(v) else 0
            else:
                setattr(t, k, int(v))
        self._check(self.lib.xcorr_hip_set_tuning(self.h, ctypes.byref(t)), "xcorr_hip_set_tuning")
        return t

    def autotune(self, windows, radix=False):
        w = _as_windows(windows)
        t = Tuning()
        rc = self.lib.xcorr_hip_autotune(self.h, _ptr(w), w.shape[0], TUNE_RADIX if radix else 0, ctypes.byref(t))
        self._check(rc, "xcorr_hip_autotune")
        return t

    def stream(self):
        return self.lib.xcorr_hip_stream(self.h)

    def kernel_times(self):
        cap = 32
        names = (ctypes.c_char_p * cap)()
        ms = np.zeros(cap, np.float32)
        launches = np.zeros(cap, np.int32)
        n = self._check(self.lib.xcorr_hip_kernel_times(self.h, names, _ptr(ms), _ptr(launches), cap),
                        "xcorr_hip_kernel_times")
        return {names[i].decode(): (float(ms[i]), int(launches[i])) for i in range(n)}

    def dump(self, what, slot=0):
        p = self.params
        if what in (0, 1):
            a = np.zeros((p.npy, p.npx), np.float32)
        elif what == 2:
            a = np.zeros((p.nyc, p.nxc), np.float32 if p.corr_flag == FREQ else np.float64)
        else:
            a = np.zeros(2, np.float64)
        self._check(self.lib.xcorr_hip_dump(self.h, what, slot, _ptr(a), a.nbytes), "xcorr_hip_dump")
        return a

    def close(self):
        if getattr(self, "h", None):
            self.lib.xcorr_hip_destroy(self.h)
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


class Multi:
    """Multi-pair workload run as one HIP graph launch (xcorr_hip_multi_*).

    pairs: list of dicts with keys params (Params), master, slave (torch device
    tensors, (lines, samples, 2) int16 or (lines, samples) float32), windows
    (WINDOW_DTYPE array) and out (torch device uint8 tensor of
    n * RESULT_DTYPE.itemsize bytes)."""

    def __init__(self, pairs, device=-1, tune=False, tune_radix=False):
        self.lib = load_library()
        arr = (Pair * max(1, len(pairs)))()
        self._keep = []
        for i, q in enumerate(pairs):
            w = _as_windows(q["windows"])
            m, s = q["master"], q["slave"]
            self._keep += [w, m, s, q["out"]]
            a = arr[i]
            a.params = q["params"]
            a.master, a.m_nx, a.m_ny = m.data_ptr(), m.shape[1], m.shape[0]
            a.slave, a.s_nx, a.s_ny = s.data_ptr(), s.shape[1], s.shape[0]
            a.win, a.n = w.ctypes.data, w.shape[0]
            a.out = q["out"].data_ptr()
            if q["out"].numel() < w.shape[0] * RESULT_DTYPE.itemsize:
                raise ValueError("pair %d: out holds fewer than n results" % i)
        flags = (TUNE if tune or tune_radix else 0) | (TUNE_RADIX if tune_radix else 0)
        h = ctypes.c_void_p()
        rc = self.lib.xcorr_hip_multi_create(arr, len(pairs), int(device), flags, ctypes.byref(h))
        if rc != 0:
            raise XcorrError("xcorr_hip_multi_create failed (%d): %s"
                             % (rc, self.lib.xcorr_hip_multi_last_error(None).decode()))
        self.h = h
        self.npairs = len(pairs)

    def _check(self, rc, what):
        if rc < 0:
            raise XcorrError("%s failed (%d): %s" % (what, rc, self.lib.xcorr_hip_multi_last_error(self.h).decode()))
        return rc

    def run(self, flags=0):
        self._check(self.lib.xcorr_hip_multi_run(self.h, flags), "xcorr_hip_multi_run")

    def sync(self):
        self._check(self.lib.xcorr_hip_multi_sync(self.h), "xcorr_hip_multi_sync")

    def stream(self):
        return self.lib.xcorr_hip_multi_stream(self.h)

    def classes(self):
        """[(params, tuning dict)] per class and the class of every pair."""
        cap = 64
        ps = (Params * cap)()
        ts = (Tuning * cap)()
        pc = np.zeros(max(1, self.npairs), np.int32)
        n = self._check(self.lib.xcorr_hip_multi_classes(self.h, ps, ts, _ptr(pc), cap), "xcorr_hip_multi_classes")
        return [(ps[k], ts[k].as_dict()) for k in range(min(n, cap))], pc[:self.npairs]

    def close(self):
        if getattr(self, "h", None):
            self.lib.xcorr_hip_multi_destroy(self.h)
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


# ---------------------------------------------------------------------------
# boundary helpers (the reference's host-side formulas)
# ---------------------------------------------------------------------------
def get_locations(m_nx, m_ny, nxl, nyl, xsearch=64, ysearch=64):
    """Location grid, get_locations.c:14-45 (x_inc, y_inc integer division;
    i = 2..nxl+1, j = 1..nyl).  Returns (x, y) int32 arrays in output order."""
    nx_corr, ny_corr = 2 * xsearch, 2 * ysearch
    npx, npy = 4 * xsearch, 4 * ysearch
    x_inc = int((m_nx - 2 * (xsearch + nx_corr)) / (nxl + 3))
    y_inc = int((m_ny - 2 * (ysearch + ny_corr)) / (nyl + 1))
    j, i = np.meshgrid(np.arange(1, nyl + 1), np.arange(2, nxl + 2), indexing="ij")
    return (npx + i * x_inc).astype(np.int32).ravel(), (npy + j * y_inc).astype(np.int32).ravel()


def make_windows(x, y, npx, npy, rshift=0, ashift=0, astretcha=0.0):
    """Window corners as xcorr.cu:332-333 and read_xcorr_data.c:68-75 place them."""
    x = np.asarray(x, np.int64)
    y = np.asarray(y, np.int64)
    ishft = np.trunc(y * astretcha).astype(np.int64)
    w = np.zeros(x.shape[0], WINDOW_DTYPE)
    w["mx0"] = x - npx // 2
    w["my0"] = y - npy // 2
    w["sx0"] = x + rshift - npx // 2
    w["sy0"] = y + ashift + ishft - npy // 2
    return w


def offsets(res, ri, rshift=0, ashift=0, y=None, astretcha=0.0):
    """dx, dy as print_results.c:37-38 composes them (float32 arithmetic)."""
    f32 = np.float32
    ishft = np.zeros(res.shape[0], np.int64) if y is None else np.trunc(np.asarray(y) * astretcha).astype(np.int64)
    dx = (res["xoff"] / f32(ri)) - (res["xfrac"] / f32(ri)) + f32(rshift)
    dy = (res["yoff"] - res["yfrac"]) + f32(ashift) + ishft.astype(f32)
    return dx.astype(f32), dy.astype(f32)


# ---------------------------------------------------------------------------
# host library (libxcorr_host.so): the drop-in's CLI / PRM / location / output code
# ---------------------------------------------------------------------------
class HostConfig(ctypes.Structure):
    """``xh_config`` of host/xcorr_host.h."""
    _fields_ = [(n, ctypes.c_int) for n in (
        "format", "corr_flag", "offset_flag", "interp_flag", "interp_factor", "nx_corr", "ny_corr", "xsearch",
        "ysearch", "m_nx", "m_ny", "s_nx", "s_ny", "nxc", "nyc", "npx", "npy", "nxl", "nyl", "x_offset",
        "y_offset", "nlocs", "x_inc", "y_inc", "ri", "n2x", "n2y", "verbose")] + [
        ("astretcha", ctypes.c_double), ("data1_name", ctypes.c_char * 256), ("data2_name", ctypes.c_char * 256),
        ("filename", ctypes.c_char * 128)]


_hlib = None


def host_library():
    global _hlib
    if _hlib is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise ImportError("libxcorr_host.so is not built")
        h = ctypes.CDLL(HOST_LIB_PATH)
        C = ctypes.POINTER(HostConfig)
        h.xh_set_defaults.argtypes = [C]
        h.xh_parse_command_line.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), C, ctypes.c_char_p]
        h.xh_handle_prm.argtypes = [ctypes.c_char_p, ctypes.c_char_p, C, ctypes.c_char_p, ctypes.c_size_t]
        h.xh_get_locations.argtypes = [C, ctypes.c_void_p, ctypes.c_int]
        I = ctypes.POINTER(ctypes.c_int32)
        h.xh_window.argtypes = [C, ctypes.c_int, ctypes.c_int, I, I, I, I]
        h.xh_window.restype = None
        h.xh_format_result.argtypes = [C, ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 5 + [ctypes.c_char_p,
                                                                                                  ctypes.c_size_t]
        _hlib = h
    return _hlib
