"""bench.py's trace summary (--graph-occupancy) on small hand-made kernel traces:
the timed region between the last two markers, the busy union, the time at
each concurrency and the workgroup fill are checked against values worked out
by hand."""
import csv
import importlib.util
import os

import pytest

from conftest import ROOT

FIELDS = ["Kind", "Agent_Id", "Queue_Id", "Stream_Id", "Thread_Id", "Dispatch_Id", "Kernel_Id", "Kernel_Name",
          "Correlation_Id", "Start_Timestamp", "End_Timestamp", "LDS_Block_Size", "Scratch_Size", "VGPR_Count",
          "Accum_VGPR_Count", "SGPR_Count", "Workgroup_Size_X", "Workgroup_Size_Y", "Workgroup_Size_Z",
          "Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z"]


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _trace(path, kernels):
    """kernels: (dispatch id, name, start ns, end ns, workgroups of 256 threads)."""
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for did, name, s, e, wg in kernels:
            w.writerow({"Kind": "KERNEL_DISPATCH", "Agent_Id": "Agent 1", "Queue_Id": 1, "Stream_Id": 0,
                        "Thread_Id": 1, "Dispatch_Id": did, "Kernel_Id": 1, "Kernel_Name": name,
                        "Correlation_Id": did, "Start_Timestamp": s, "End_Timestamp": e, "LDS_Block_Size": 0,
                        "Scratch_Size": 0, "VGPR_Count": 64, "Accum_VGPR_Count": 0, "SGPR_Count": 32,
                        "Workgroup_Size_X": 256, "Workgroup_Size_Y": 1, "Workgroup_Size_Z": 1,
                        "Grid_Size_X": 256 * wg, "Grid_Size_Y": 1, "Grid_Size_Z": 1})


SPIN = "void at::cuda::(anonymous namespace)::spin_kernel(long)"


def test_occupancy_of_the_last_marked_region(bench, tmp_path):
    path = str(tmp_path / "t.csv")
    ms = 1_000_000
    _trace(path, [
        (1, "void xc::k_stripA256<0, false>(xc::DevParams)", 0, 5 * ms, 4096),      # before the markers: ignored
        (2, SPIN, 6 * ms, 6 * ms + 10, 1),
        (3, SPIN, 7 * ms, 7 * ms + 10, 1),
        (4, "void xc::k_stripA256<0, false>(xc::DevParams)", 10 * ms, 14 * ms, 4096),
        (5, "xc::k_finish(xc::DevParams)", 12 * ms, 16 * ms, 512),                 # overlaps the first
        (6, "void xc::k_colpass256<true>(xc::DevParams)", 18 * ms, 20 * ms, 2048),  # after a 2 ms hole
        (7, "void at::native::fill_kernel(float)", 20 * ms, 21 * ms, 64),          # not an engine kernel
        (8, SPIN, 22 * ms, 22 * ms + 10, 1),
        (9, "xc::k_finish(xc::DevParams)", 30 * ms, 31 * ms, 512),                 # after the last marker
    ])
    r = bench.graph_occupancy(path)
    assert r["kernels"] == 3
    assert r["span_ms"] == 10.0                     # 10 .. 20 ms
    assert r["busy_ms"] == 8.0                      # 10 .. 16 and 18 .. 20
    assert r["busy_frac"] == 0.8
    assert r["ms_at_concurrency"] == {"0": 2.0, "1": 6.0, "2": 2.0}
    # fill: 10..12 full (4096 wg), 12..14 full, 14..16 512/1024, 16..18 idle, 18..20 full
    assert r["fill_frac"] == pytest.approx((2 + 2 + 1 + 0 + 2) / 10.0)
    assert r["kernel_ms"]["xc::k_stripA256<0, false>"] == {"launches": 1, "ms": 4.0}
    assert r["kernel_ms"]["xc::k_finish"] == {"launches": 1, "ms": 4.0}


def test_occupancy_needs_two_markers(bench, tmp_path):
    path = str(tmp_path / "t.csv")
    _trace(path, [(1, SPIN, 0, 10, 1), (2, "xc::k_finish(xc::DevParams)", 100, 200, 1)])
    assert "error" in bench.graph_occupancy(path)
