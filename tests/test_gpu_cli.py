"""The drop-in xcorr executable on the GPU: same command line, PRM/SLC inputs and
output file as the reference.  Its output must equal, byte for byte, the golden
file written by the reference's own print_results.c from the FP32 CPU
restatement's results (tests/golden/make_golden.py)."""
import json
import os
import shutil
import subprocess

import pytest

from conftest import ROOT, gpu_present

pytestmark = pytest.mark.gpu

GOLD = os.path.join(ROOT, "tests", "golden", "cli")
XCORR = os.path.join(ROOT, "gmtsar-gpu_amd", "xcorr")
CASES = json.load(open(os.path.join(GOLD, "cases.json")))


def _run_case(tmp_path, case, env=None):
    for f in os.listdir(GOLD):
        if f.endswith((".PRM", ".SLC")):
            shutil.copy(os.path.join(GOLD, f), tmp_path)
    return subprocess.run([XCORR] + case["args"], cwd=tmp_path, capture_output=True, text=True, timeout=300,
                          env=env)


@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_cli_output_matches_golden(tmp_path, case):
    if not gpu_present():
        pytest.skip("no GPU")
    r = _run_case(tmp_path, case)
    assert r.returncode == 0, r.stderr
    assert "elapsed time" in r.stdout
    got = open(os.path.join(tmp_path, case["output"])).read()
    exp = open(os.path.join(GOLD, "expected", case["name"] + ".dat")).read()
    assert got == exp


@pytest.mark.parametrize("name", ["freq_default", "freq_shift_stretch", "time"])
def test_cli_small_bands_match_golden(tmp_path, name):
    """The SLC lines are streamed in bands; forcing many small bands must not
    change a byte of the output."""
    if not gpu_present():
        pytest.skip("no GPU")
    case = [c for c in CASES if c["name"] == name][0]
    env = dict(os.environ, XCORR_HIP_BAND_LINES="100")
    r = _run_case(tmp_path, case, env)
    assert r.returncode == 0, r.stderr
    got = open(os.path.join(tmp_path, case["output"])).read()
    exp = open(os.path.join(GOLD, "expected", case["name"] + ".dat")).read()
    assert got == exp
