"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (liboracle.so).

Loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as
the checker.  The product never imports this module.  See xcorr_oracle.h for
what the two restatements (xo32 FP32 spec path, xo64 FP64 textbook path)
follow in the reference.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "npx", "npy", "xsearch", "ysearch", "nx_corr", "ny_corr", "nxc", "nyc",
        "ri", "interp_flag", "interp_factor", "n2x", "n2y", "corr_flag")]


class Result(ctypes.Structure):
    _fields_ = [("cave", ctypes.c_double), ("xoff", ctypes.c_float), ("yoff", ctypes.c_float),
                ("corr", ctypes.c_float), ("xfrac", ctypes.c_float), ("yfrac", ctypes.c_float),
                ("cmax", ctypes.c_float), ("ip", ctypes.c_int), ("jp", ctypes.c_int),
                ("hip", ctypes.c_int), ("hjp", ctypes.c_int)]


RESULT_DTYPE = np.dtype([("cave", "<f8"), ("xoff", "<f4"), ("yoff", "<f4"), ("corr", "<f4"),
                         ("xfrac", "<f4"), ("yfrac", "<f4"), ("cmax", "<f4"), ("ip", "<i4"),
                         ("jp", "<i4"), ("hip", "<i4"), ("hjp", "<i4")])
assert RESULT_DTYPE.itemsize == ctypes.sizeof(Result)


class Dump32(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("amp1", "amp2", "mean", "surf", "md", "hres")]


Dump64 = Dump32


def build(quiet=True):
    """Compile liboracle.so with the oracle's own Makefile."""
    out = subprocess.run(["make", "-C", HERE, "liboracle.so"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        _lib.xo32_fft.argtypes = [P, ctypes.c_int, ctypes.c_int]
        _lib.xo64_fft.argtypes = [P, ctypes.c_int, ctypes.c_int]
        _lib.xo32_range_interp_row.argtypes = [P, ctypes.c_int, ctypes.c_int, P]
        _lib.xo64_range_interp_row.argtypes = [P, ctypes.c_int, ctypes.c_int, P]
        _lib.xo_twiddle64.argtypes = [ctypes.c_int, ctypes.c_long, P, P]
        _lib.xo32_set_fft_plan.argtypes = [ctypes.c_int, ctypes.c_int, P]
        for f in ("xo32_patch", "xo64_patch"):
            getattr(_lib, f).argtypes = [ctypes.POINTER(Params), P, P, ctypes.POINTER(Result), P]
        for f in ("xo32_run", "xo64_run"):
            getattr(_lib, f).argtypes = [ctypes.POINTER(Params), ctypes.c_int, P, ctypes.c_int,
                                         ctypes.c_int, P, ctypes.c_int, ctypes.c_int, P,
                                         ctypes.c_int, P, ctypes.c_int]
    return _lib


def make_params(xsearch=64, ysearch=64, ri=2, interp_flag=1, interp_factor=16, corr_flag=2,
                n2x=8, n2y=8):
    """Derived sizes as parse_xcorr_input.c:139-158 sets them."""
    p = Params()
    p.xsearch, p.ysearch = xsearch, ysearch
    p.nx_corr, p.ny_corr = 2 * xsearch, 2 * ysearch
    p.npx, p.npy = 4 * xsearch, 4 * ysearch
    p.nxc, p.nyc = 2 * xsearch, 2 * ysearch
    p.ri, p.interp_flag, p.interp_factor = ri, interp_flag, interp_factor
    p.n2x, p.n2y, p.corr_flag = n2x, n2y, corr_flag
    return p


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_fft_plan(n, radices=None):
    """Radix plan of every n-point FP32 transform (None: the default plan)."""
    r = np.zeros(4, np.int32)
    k = 0 if not radices else len(radices)
    r[:k] = radices or []
    if lib().xo32_set_fft_plan(n, k, _ptr(r)) != 0:
        raise ValueError("invalid radix plan %s for length %d" % (radices, n))


class fft_plans:
    """Context manager: radix plans {n: [radices]} for the FP32 restatement."""

    def __init__(self, plans):
        self.plans = {int(n): list(r) for n, r in (plans or {}).items() if r}

    def __enter__(self):
        for n, r in self.plans.items():
            set_fft_plan(n, r)
        return self

    def __exit__(self, *a):
        for n in self.plans:
            set_fft_plan(n, None)


def fft32(x, direction=-1):
    a = np.ascontiguousarray(x, dtype=np.complex64).copy()
    if lib().xo32_fft(_ptr(a), a.size, direction) != 0:
        raise ValueError("unsupported length %d" % a.size)
    return a


def fft64(x, direction=-1):
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    if lib().xo64_fft(_ptr(a), a.size, direction) != 0:
        raise ValueError("unsupported length %d" % a.size)
    return a


def range_interp32(row, ri):
    x = np.ascontiguousarray(row, dtype=np.complex64)
    out = np.zeros(x.size, np.float32)
    lib().xo32_range_interp_row(_ptr(x), x.size, ri, _ptr(out))
    return out


def range_interp64(row, ri):
    x = np.ascontiguousarray(row, dtype=np.complex128)
    out = np.zeros(x.size, np.float64)
    lib().xo64_range_interp_row(_ptr(x), x.size, ri, _ptr(out))
    return out


def twiddle64(n, m):
    c, s = ctypes.c_double(), ctypes.c_double()
    lib().xo_twiddle64(n, m, ctypes.byref(c), ctypes.byref(s))
    return c.value, s.value


def patch(params, mwin, swin, precision=32, dump=False):
    """Run one patch.  mwin/swin: (npy, npx) complex windows.  Returns (result
    record, dumps dict or None)."""
    mw = np.ascontiguousarray(mwin, dtype=np.complex64)
    sw = np.ascontiguousarray(swin, dtype=np.complex64)
    assert mw.shape == (params.npy, params.npx) and sw.shape == mw.shape
    res = Result()
    d = None
    dp = None
    if dump:
        ft = np.float32 if precision == 32 else np.float64
        L = params.interp_factor
        d = {"amp1": np.zeros(mw.shape, ft), "amp2": np.zeros(mw.shape, ft),
             "mean": np.zeros(2, np.float64),
             "surf": np.zeros((params.nyc, params.nxc), ft),
             "md": np.zeros((params.n2y, params.n2x), ft),
             "hres": np.zeros((params.n2y * L, params.n2x * L), ft)}
        dp = Dump32(*[d[k].ctypes.data for k in ("amp1", "amp2", "mean", "surf", "md", "hres")])
    f = lib().xo32_patch if precision == 32 else lib().xo64_patch
    rc = f(ctypes.byref(params), _ptr(mw), _ptr(sw), ctypes.byref(res),
           ctypes.byref(dp) if dp is not None else None)
    if rc != 0:
        raise RuntimeError("oracle patch failed rc=%d" % rc)
    rec = np.frombuffer(bytearray(res), dtype=RESULT_DTYPE)[0]
    return rec, d


def run(params, master, slave, locs, precision=32, fmt=0, nthreads=0):
    """Batch over locations straight from the images.  master/slave: (ny, nx, 2)
    int16 for fmt 0 or (ny, nx) float32 for fmt 1.  locs: (n, 4) int32 window
    corners (mx0, my0, sx0, sy0)."""
    m = np.ascontiguousarray(master)
    s = np.ascontiguousarray(slave)
    lc = np.ascontiguousarray(locs, dtype=np.int32)
    n = lc.shape[0]
    out = np.zeros(n, RESULT_DTYPE)
    f = lib().xo32_run if precision == 32 else lib().xo64_run
    rc = f(ctypes.byref(params), fmt, _ptr(m), m.shape[1], m.shape[0], _ptr(s), s.shape[1],
           s.shape[0], _ptr(lc), n, _ptr(out), nthreads)
    if rc != 0:
        raise RuntimeError("oracle run failed rc=%d" % rc)
    return out


def interp2d(md, ifc, precision=32):
    """Real part of the ifc-fold 2-D zero-pad interpolation of a real 8x8 (or
    n2y x n2x) neighbourhood."""
    ft = np.float32 if precision == 32 else np.float64
    a = np.ascontiguousarray(md, dtype=ft)
    n2y, n2x = a.shape
    out = np.zeros((n2y * ifc, n2x * ifc), ft)
    f = lib().xo32_interp2d if precision == 32 else lib().xo64_interp2d
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    if f(_ptr(a), n2x, n2y, ifc, _ptr(out)) != 0:
        raise RuntimeError("interp2d failed")
    return out
