#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/cli/ (run from the repo root,
in the container that has the reference tree; needs oracle/_ref/ref_host, built
by `make -C oracle ref`).

For a small synthetic SLC pair (written as GMTSAR SLC + PRM files) and several
xcorr option sets, it records:
  * the location grid and parameters produced by the reference's own host code
    (parse_xcorr_input.c, sio_struct.c, get_locations.c via ref_host locs);
  * the master/slave windows of one location as the reference reads them
    (read_xcorr_data.c via ref_host read);
  * the expected output file: correlation results of the FP32 CPU restatement
    (oracle) formatted by the reference's print_results.c (ref_host print).
"""
import importlib
import importlib.util
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(ROOT, "tests", "golden", "cli")
REF_HOST = os.path.join(ROOT, "oracle", "_ref", "ref_host")
sys.path.insert(0, os.path.join(ROOT, "oracle"))

N_RANGE, N_AZ = 256, 320

CASES = [
    # name, aligned PRM, extra options, output file
    ("freq_default", "aligned.PRM", ["-nx", "5", "-ny", "4", "-xsearch", "16", "-ysearch", "16"], "freq_xcorr.dat"),
    ("freq_xs8_ys16_interp8", "aligned.PRM", ["-nx", "6", "-ny", "5", "-xsearch", "8", "-ysearch", "16",
                                              "-interp", "8"], "freq_xcorr.dat"),
    ("freq_nointerp", "aligned.PRM", ["-nx", "4", "-ny", "3", "-xsearch", "16", "-ysearch", "16", "-nointerp"],
     "freq_xcorr.dat"),
    ("freq_norange", "aligned.PRM", ["-nx", "4", "-ny", "3", "-xsearch", "16", "-ysearch", "16", "-norange"],
     "freq_xcorr.dat"),
    ("freq_ri4", "aligned.PRM", ["-nx", "4", "-ny", "3", "-xsearch", "16", "-ysearch", "16", "-range_interp", "4"],
     "freq_xcorr.dat"),
    ("freq_shift_stretch", "aligned_shift.PRM", ["-nx", "4", "-ny", "4", "-xsearch", "16", "-ysearch", "16"],
     "freq_xcorr.dat"),
    ("freq_noshift", "aligned_shift.PRM", ["-nx", "4", "-ny", "4", "-xsearch", "16", "-ysearch", "16", "-noshift"],
     "freq_xcorr.dat"),
    ("time", "aligned.PRM", ["-nx", "3", "-ny", "2", "-xsearch", "16", "-ysearch", "16", "-time"], "time_xcorr.dat"),
    ("time4", "aligned.PRM", ["-nx", "3", "-ny", "2", "-xsearch", "8", "-ysearch", "8", "-time4"],
     "time_xcorr_Gatelli.dat"),
    ("real", "aligned_real.PRM", ["-nx", "4", "-ny", "3", "-xsearch", "16", "-ysearch", "16", "-real"],
     "freq_xcorr.dat"),
]


def load_synth():
    pkg = os.path.join(ROOT, "gmtsar-gpu_amd")
    spec = importlib.util.spec_from_file_location("gmtsar_gpu_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gmtsar_gpu_amd"] = mod
    spec.loader.exec_module(mod)
    return importlib.import_module("gmtsar_gpu_amd.synth")


def ref(args, stdin=None):
    r = subprocess.run([REF_HOST] + args, cwd=OUT, capture_output=True, text=True, input=stdin)
    if r.returncode != 0:
        raise RuntimeError("ref_host %s failed: %s" % (args, r.stderr))
    return r.stdout


def main():
    import oracle
    synth = load_synth()
    os.makedirs(os.path.join(OUT, "expected"), exist_ok=True)
    m, s = synth.constant_shift_pair(N_RANGE, N_AZ, 1.6, -2.3, 0.85, seed=21, noise_seed=22)
    synth.write_slc(os.path.join(OUT, "master.SLC"), m)
    synth.write_slc(os.path.join(OUT, "aligned.SLC"), s)
    mr = np.ascontiguousarray(m[..., 0].astype(np.float32))
    sr = np.ascontiguousarray(s[..., 0].astype(np.float32))
    mr.astype("<f4").tofile(os.path.join(OUT, "master_real.SLC"))
    sr.astype("<f4").tofile(os.path.join(OUT, "aligned_real.SLC"))
    synth.write_prm(os.path.join(OUT, "master.PRM"), "master.SLC", N_RANGE, N_AZ)
    synth.write_prm(os.path.join(OUT, "aligned.PRM"), "aligned.SLC", N_RANGE, N_AZ)
    synth.write_prm(os.path.join(OUT, "aligned_shift.PRM"), "aligned.SLC", N_RANGE, N_AZ, rshift=2, ashift=-3,
                    prf=1700.0)
    synth.write_prm(os.path.join(OUT, "master_real.PRM"), "master_real.SLC", N_RANGE, N_AZ)
    synth.write_prm(os.path.join(OUT, "aligned_real.PRM"), "aligned_real.SLC", N_RANGE, N_AZ)
    manifest = []
    for name, aprm, opts, outfile in CASES:
        mprm = "master_real.PRM" if aprm == "aligned_real.PRM" else "master.PRM"
        text = ref(["locs", mprm, aprm] + opts)
        lines = text.strip().splitlines()
        P = lines[0].split()[1:]
        (m_nx, m_ny, s_nx, s_ny, x_off, y_off) = [int(v) for v in P[:6]]
        astretcha = float(P[6])
        (npx, npy, nxc, nyc, nx_corr, ny_corr, xsearch, ysearch, ri, interp_flag, ifc, corr_flag, nxl, nyl,
         fmt) = [int(v) for v in P[7:]]
        xy = np.array([[int(v) for v in l.split()[1:]] for l in lines[1:]], np.int32)
        x, y = xy[:, 0], xy[:, 1]
        ishft = np.trunc(y.astype(np.float64) * astretcha).astype(np.int64)
        locs = np.stack([x - npx // 2, y - npy // 2, x + x_off - npx // 2, y + y_off + ishft - npy // 2],
                        1).astype(np.int32)
        op = oracle.make_params(xsearch, ysearch, ri=ri, interp_flag=interp_flag, interp_factor=ifc,
                                corr_flag=corr_flag)
        img_m, img_s = (mr, sr) if fmt == 1 else (m, s)
        res = oracle.run(op, img_m, img_s, locs, precision=32, fmt=fmt)
        feed = "".join("%d %d %.9g %.9g %.9g %.9g %.9g\n" % (x[i], y[i], res["xoff"][i], res["yoff"][i],
                                                              res["xfrac"][i], res["yfrac"][i], res["corr"][i])
                       for i in range(len(x)))
        expected = ref(["print", str(ri), str(interp_flag), str(ifc), str(x_off), str(y_off), repr(astretcha)],
                       stdin=feed)
        with open(os.path.join(OUT, "expected", name + ".dat"), "w") as f:
            f.write(expected)
        with open(os.path.join(OUT, "expected", name + ".locs"), "w") as f:
            f.write(text)
        manifest.append({"name": name, "args": [mprm, aprm] + opts, "output": outfile})
    # window read by the reference for one location (pins the gather: xcorr.cu:330-337)
    text = ref(["read", "master.PRM", "aligned_shift.PRM", "5", "-nx", "4", "-ny", "4", "-xsearch", "16",
                "-ysearch", "16"])
    with open(os.path.join(OUT, "expected", "read_window_loc5.txt"), "w") as f:
        f.write(text)
    with open(os.path.join(OUT, "cases.json"), "w") as f:
        json.dump(manifest, f, indent=1)
    print("wrote %d cases to %s" % (len(manifest), OUT))


if __name__ == "__main__":
    main()
