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


@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_cli_output_matches_golden(tmp_path, case):
    if not gpu_present():
        pytest.skip("no GPU")
    for f in os.listdir(GOLD):
        if f.endswith((".PRM", ".SLC")):
            shutil.copy(os.path.join(GOLD, f), tmp_path)
    r = subprocess.run([XCORR] + case["args"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "elapsed time" in r.stdout
    got = open(os.path.join(tmp_path, case["output"])).read()
    exp = open(os.path.join(GOLD, "expected", case["name"] + ".dat")).read()
    assert got == exp
