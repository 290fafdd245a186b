"""C ABI surface: the library loads and exports every entry point declared in
include/xcorr_hip.h; parameter defaults and validation (no GPU needed)."""
import ctypes
import os
import re

from conftest import ROOT, gpu_present


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "xcorr_hip.h")).read()
    return sorted(set(re.findall(r"\b(xcorr_hip_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_declared_symbols(xg):
    lib = ctypes.CDLL(xg.LIB_PATH)
    declared = _declared_symbols()
    assert declared, "no declarations found"
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    assert set(declared) == set(xg.EXPORTED_SYMBOLS)


def test_library_exports_cnn_symbols(xg):
    text = open(os.path.join(ROOT, "include", "xcorr_cnn.h")).read()
    declared = sorted(set(re.findall(r"\b(xcorr_cnn_[a-z_0-9]+)\s*\(", text)))
    assert len(declared) >= 14
    lib = ctypes.CDLL(xg.LIB_PATH)
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing


def test_abi_version(xg):
    assert xg.load_library().xcorr_hip_abi_version() == 1


def test_default_params_follow_reference(xg):
    # parse_xcorr_input.c:38-88 and 139-158: npx = nx_corr + 2*xsearch, nxc = 2*xsearch
    p = xg.make_params(64, 32)
    assert (p.nx_corr, p.ny_corr, p.npx, p.npy, p.nxc, p.nyc) == (128, 64, 256, 128, 128, 64)
    assert (p.ri, p.interp_flag, p.interp_factor, p.n2x, p.n2y, p.corr_flag) == (2, 1, 16, 8, 8, 2)


def test_create_rejects_bad_params_or_fails_loudly_without_gpu(xg):
    p = xg.make_params(48, 64)          # not a power of two
    try:
        xg.Engine(p)
    except xg.XcorrError as e:
        assert "power" in str(e) or "device" in str(e).lower()
    else:
        raise AssertionError("bad parameters accepted")
    if not gpu_present():
        try:
            xg.Engine(xg.make_params(64, 64))
        except xg.XcorrError as e:
            assert "device" in str(e).lower()
        else:
            raise AssertionError("engine created without a GPU")
