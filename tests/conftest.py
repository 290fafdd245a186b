"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on CPU."""
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def load_package():
    if "gmtsar_gpu_amd" in sys.modules:
        return sys.modules["gmtsar_gpu_amd"]
    pkg_dir = os.path.join(ROOT, "gmtsar-gpu_amd")
    spec = importlib.util.spec_from_file_location("gmtsar_gpu_amd", os.path.join(pkg_dir, "__init__.py"),
                                                  submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gmtsar_gpu_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def load_synth():
    load_package()
    import importlib
    return importlib.import_module("gmtsar_gpu_amd.synth")


@pytest.fixture(scope="session")
def xg():
    return load_package()


@pytest.fixture(scope="session")
def synth():
    return load_synth()


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


def gpu_present():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
