#!/usr/bin/env python3
"""Benchmark: xcorr correlation patches/s on one MI355X.

Workload (the headline strip workload): a 10800 (range) x 27648 (azimuth) complex-int16
SLC pair generated on the GPU (speckle model, offset field dx = 2.1 + 1e-4 x,
dy = -7.6 + 3e-4 y, coherence per 512 x 512 block uniform in [0.3, 0.95]),
locations from get_locations with -nx 100 -ny 400 -xsearch 64 -ysearch 64
(40,000 patches, 256 x 256 windows, range interpolation x2, 16x high-res).

One step = the whole hot path over all 40,000 locations with the images already
resident in HBM and the results left in HBM.  The window list is host memory,
as in xcorr: each step builds the strip plan on the host (sort, merge, upload
of a few hundred KB per batch) while the previous batches run on the GPU.

Also reported:
  roofline      dominant kernel: algorithmic bytes per launch (524,308 B per
                patch: two 256 x 256 complex-int16 windows + 20 B of output)
                / its mean launch time, measured with HIP events on the
                engine's stream, against 8 TB/s
  cpu_baseline  the FP32 CPU restatement (oracle) on a bounded sample of the
                same workload, on the host cores of this machine
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

METRIC = "correlation patches/sec on 1 GPU (128², 16× interp); HBM roofline %; proposal vs plain"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
BYTES_PER_PATCH = 2 * 256 * 256 * 4 + 20          # 524,308
FLOP_PER_PATCH = 33.39e6                          # 5 N log2 N over the path's FFTs


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """Threads for the CPU baseline and a description of the host: every CPU in
    this process's affinity set, capped by OMP_NUM_THREADS when the environment
    sets one (a machine that allots a CPU share per GPU says so there)."""
    aff = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(aff, int(omp)) if omp.isdigit() and int(omp) > 0 else aff
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return threads, {"nproc": aff, "os_cpu_count": os.cpu_count(), "omp_num_threads": omp or None,
                     "cpu_model": model}


def committed_traffic():
    """Per-launch HBM bytes per kernel from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (the JSON line under "Per-launch
    traffic"), or {} when it is absent."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_xcorr_strip.md")
    try:
        lines = open(path).read().splitlines()
    except OSError:
        return {}
    for i, line in enumerate(lines):
        if line.startswith("Per-launch traffic"):
            for nxt in lines[i + 1:]:
                if nxt.strip().startswith("{"):
                    return json.loads(nxt.strip())
    return {}


def trace_marker(torch):
    """A one-cycle marker kernel (torch.cuda._sleep, `spin_kernel` in a trace)
    launched right before and right after each timed loop, with the device
    idle on both sides, so a kernel trace of the run shows where the timed
    steps are."""
    torch.cuda.synchronize()
    torch.cuda._sleep(1)
    torch.cuda.synchronize()


def graph_occupancy(path, slots_per_cu=4, cus=256):
    """Device occupancy over the timed steps in a rocprofv3 kernel trace of a
    bench.py run: the engine's kernels (`xc::` names) dispatched between the
    last two trace markers (see trace_marker).  Reported: the span from the
    first start to the last end, the union of the kernels' busy intervals,
    the time at each number of concurrently running kernels, and a fill
    estimate that weights time by min(1, running workgroups / (slots_per_cu x
    cus)), i.e. how much of it could have held a full device of 256-thread
    workgroups.  Times are those of the profiled run (the tracer slows each
    launch and may serialise a graph's branches)."""
    import csv
    import collections
    rows = []
    marks = []
    with open(path) as f:
        for d in csv.DictReader(f):
            did = int(d["Dispatch_Id"])
            if "spin_kernel" in d["Kernel_Name"]:
                marks.append(did)
                continue
            if "xc::" not in d["Kernel_Name"]:
                continue
            wg = (int(d["Grid_Size_X"]) * int(d["Grid_Size_Y"]) * int(d["Grid_Size_Z"]) //
                  max(1, int(d["Workgroup_Size_X"]) * int(d["Workgroup_Size_Y"]) * int(d["Workgroup_Size_Z"])))
            name = d["Kernel_Name"].replace("void ", "").split("(")[0]
            rows.append((did, int(d["Start_Timestamp"]), int(d["End_Timestamp"]), name, wg))
    marks.sort()
    if len(marks) < 2:
        return {"error": "fewer than two trace markers in %s" % path}
    lo, hi = marks[-2], marks[-1]
    rep = sorted((r[1], r[2], r[3], r[4]) for r in rows if lo < r[0] < hi)
    if not rep:
        return {"error": "no engine kernels between the last two markers in %s" % path}
    t0, t1 = rep[0][0], max(r[1] for r in rep)
    busy, cs, ce = 0, rep[0][0], rep[0][1]
    for s, e, _, _ in rep[1:]:
        if s > ce:
            busy += ce - cs
            cs, ce = s, e
        else:
            ce = max(ce, e)
    busy += ce - cs
    events = sorted([(r[0], 1, r[3]) for r in rep] + [(r[1], -1, -r[3]) for r in rep])
    conc = collections.Counter()
    running = wg = 0
    last = t0
    fill = 0.0
    for t, dk, dw in events:
        conc[running] += t - last
        fill += min(1.0, wg / float(slots_per_cu * cus)) * (t - last)
        last = t
        running += dk
        wg += dw
    per = collections.defaultdict(lambda: [0, 0])
    for s, e, name, _ in rep:
        per[name][0] += 1
        per[name][1] += e - s
    span = t1 - t0
    return {"trace": os.path.basename(path), "kernels": len(rep), "span_ms": round(span / 1e6, 3),
            "busy_ms": round(busy / 1e6, 3), "busy_frac": round(busy / span, 4),
            "fill_frac": round(fill / span, 4),
            "ms_at_concurrency": {str(k): round(v / 1e6, 3) for k, v in sorted(conc.items())},
            "kernel_ms": {k: {"launches": v[0], "ms": round(v[1] / 1e6, 3)}
                          for k, v in sorted(per.items(), key=lambda kv: -kv[1][1])}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=8192, help="locations per engine batch")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-others", action="store_true",
                    help="strip workload only: skip the compact batched / proposal / dense records")
    ap.add_argument("--serial", action="store_true",
                    help="strip batches one after another on one stream (default: consecutive batches overlap on "
                         "two contexts); rocprof kernel durations are then those of each kernel alone (strip and "
                         "time workloads)")
    ap.add_argument("--small", action="store_true", help="the small workload (1024 x 2048, 1000 patches) instead")
    ap.add_argument("--workload", default="strip",
                    choices=("strip", "small", "batched", "proposal", "dense", "time"),
                    help="strip: 10800 x 27648 pair, 40,000 patches (default); small: 1024 x 2048, 1000 patches; "
                         "batched: 64 uneven pairs with mixed patch sizes in one graph launch; proposal: the "
                         "learned-proposal form beside plain xcorr on the strip workload; dense: 291k 128 x 128 "
                         "windows on a 32-px grid, half time-domain, half frequency, one graph launch; time: "
                         "-time and -time4 at xsearch 32 and 64 on the strip pair")
    ap.add_argument("--tau", type=float, default=0.9, help="proposal: confidence threshold of the shrunken search")
    ap.add_argument("--tau-skip", type=float, default=1.01, help="proposal: confidence above which high-res is skipped")
    ap.add_argument("--xs-small", type=int, default=32, help="proposal: xsearch of the shrunken search")
    ap.add_argument("--verify", action="store_true",
                    help="proposal: redo shrunken searches whose peak is on the lag-window border with plain xcorr")
    ap.add_argument("--corr-min", type=float, default=0.0, help="proposal: verification also redoes corr below this")
    ap.add_argument("--weights", default=os.path.join(ROOT, "gmtsar-gpu_amd", "weights", "proposal_cnn.npy"))
    ap.add_argument("--pairs", type=int, default=64, help="batched: number of SLC pairs")
    ap.add_argument("--no-tune", action="store_true", help="batched: skip the autotuner")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (HBM counter calibration)")
    ap.add_argument("--copy-probe-mode", default="plain", choices=("plain", "transpose"),
                    help="transpose: the copy reads a (MIB*4096 x 32) complex64 matrix column by column "
                         "(strided 8-byte reads, each byte read once) into its transpose")
    ap.add_argument("--traffic", default="", metavar="JSON",
                    help="per-launch HBM bytes of the dominant kernel from rocprofv3 FETCH_SIZE/WRITE_SIZE runs")
    ap.add_argument("--graph-occupancy", default="", metavar="CSV",
                    help="only summarise a rocprofv3 kernel trace (kernel_trace.csv) of --workload batched or "
                         "dense: how busy the device is during the last graph replay")
    args = ap.parse_args()
    if args.graph_occupancy:
        print(json.dumps(graph_occupancy(args.graph_occupancy)), flush=True)
        return 0
    if args.copy_probe:
        import torch
        nbytes = args.copy_probe << 20
        if args.copy_probe_mode == "plain":
            src = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
            dst = torch.empty_like(src)
            for _ in range(3):
                dst.copy_(src)
        else:
            rows = nbytes // (32 * 8)
            src = torch.ones(rows, 32, dtype=torch.complex64, device="cuda")
            dst = torch.empty(32, rows, dtype=torch.complex64, device="cuda")
            for _ in range(3):
                dst.copy_(src.t())
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": nbytes, "mode": args.copy_probe_mode}), flush=True)
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the xcorr path has independent patches and this build "
              "runs on a single MI355X (--gpus %d requested)" % args.gpus, flush=True)
        return 2

    import torch
    import importlib
    import __graft_entry__
    xg = __graft_entry__.load_package()
    synth = importlib.import_module("gmtsar_gpu_amd.synth")
    if not torch.cuda.is_available():
        print("no GPU available", file=sys.stderr)
        return 1
    dev = torch.device("cuda:0")
    others = {"batched": bench_batched, "proposal": bench_proposal, "dense": bench_dense, "time": bench_time}
    if args.workload in others:
        print(json.dumps(others[args.workload](args, xg, synth, torch, dev)), flush=True)
        return 0

    t0 = time.time()
    if args.small or args.workload == "small":
        n_range, n_az, nxl, nyl = 1024, 2048, 20, 50
        mh, sh = synth.constant_shift_pair(n_range, n_az, 3.37, -12.81, 0.8)
        m, s = torch.from_numpy(mh).to(dev), torch.from_numpy(sh).to(dev)
        workload = "plain xcorr, 1024 x 2048 cint16 pair, -nx 20 -ny 50 -xsearch 64 -ysearch 64, interp 16"
    else:
        n_range, n_az, nxl, nyl = 10800, 27648, 100, 400
        m, s = synth.strip_pair_torch(n_range, n_az, lambda x: 2.1 + 1e-4 * x, lambda y: -7.6 + 3e-4 * y,
                                      device="cuda")
        workload = "plain xcorr, 10800 x 27648 cint16 pair, -nx 100 -ny 400 -xsearch 64 -ysearch 64, interp 16"
    torch.cuda.synchronize()
    log("synthetic pair ready in %.1f s" % (time.time() - t0))

    p = xg.make_params(64, 64)
    x, y = xg.get_locations(n_range, n_az, nxl, nyl, 64, 64)
    w = xg.make_windows(x, y, p.npx, p.npy)
    n = len(w)
    dout = torch.zeros(n * xg.RESULT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    eng = xg.Engine(p, max_batch=min(args.batch, n))
    eng.set_images(m, s)

    step_flags = xg.NO_SYNC | (xg.SERIAL if args.serial else 0)
    for _ in range(args.warmup):
        eng.correlate(w, out=dout, flags=step_flags)
    eng.sync()
    torch.cuda.synchronize()
    log("warmup done")

    # HIP events on the engine's stream between the steps give the per-step spread
    est = torch.cuda.ExternalStream(eng.stream())
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    trace_marker(torch)
    t1 = time.perf_counter()
    evs[0].record(est)
    for k in range(args.steps):
        eng.correlate(w, out=dout, flags=step_flags)
        evs[k + 1].record(est)
    eng.sync()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    trace_marker(torch)
    step_s = (t2 - t1) / args.steps
    value = n / step_s
    in_order = [evs[k].elapsed_time(evs[k + 1]) for k in range(args.steps)]
    per_step = sorted(in_order)
    step_stats = {"median_ms": round(per_step[len(per_step) // 2], 4), "min_ms": round(per_step[0], 4),
                  "max_ms": round(per_step[-1], 4), "ms": [round(v, 3) for v in in_order],
                  "source": "HIP events on the engine stream"}
    log("timed %d steps: %.3f ms/step, %.0f patches/s (median step %.3f ms)"
        % (args.steps, step_s * 1e3, value, step_stats["median_ms"]))

    # per-kernel HIP-event timing on the engine's stream (separate, untimed step)
    eng.correlate(w, out=dout, flags=xg.TIMING)
    kt = eng.kernel_times()
    dom = max(kt.items(), key=lambda kv: kv[1][0])
    dom_name, (dom_ms, dom_launches) = dom
    per_launch_ms = dom_ms / dom_launches
    patches_per_launch = n / dom_launches
    achieved = BYTES_PER_PATCH * patches_per_launch / (per_launch_ms * 1e-3) / 1e9
    tr = {}
    if args.traffic:
        tr = json.loads(args.traffic) if args.traffic.lstrip().startswith("{") else json.load(open(args.traffic))
    elif not args.small:
        tr = committed_traffic()
    traffic = tr.get(dom_name)
    # measured device-copy bandwidth (the practical HBM ceiling) and the host-to-device
    # rate of the pair itself (the PCIe-inclusive cost of handing over host buffers)
    cbytes = 1 << 30
    src = torch.empty(cbytes, dtype=torch.uint8, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    tc = time.perf_counter()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    copy_gbs = 2 * cbytes * 5 / (time.perf_counter() - tc) / 1e9
    del src, dst
    pair_bytes = 2 * m.numel() * m.element_size()
    hbuf = torch.empty(m.numel() * m.element_size(), dtype=torch.uint8, pin_memory=True)
    gbuf = torch.empty_like(hbuf, device=dev)
    gbuf.copy_(hbuf, non_blocking=True)
    torch.cuda.synchronize()
    th = time.perf_counter()
    for _ in range(2):
        gbuf.copy_(hbuf, non_blocking=True)
    torch.cuda.synchronize()
    h2d_s = (time.perf_counter() - th) / 2 * 2          # both images
    del hbuf, gbuf
    h2d = {"bytes": pair_bytes, "ms": round(h2d_s * 1e3, 3), "GBs": round(pair_bytes / h2d_s / 1e9, 2),
           "pcie_inclusive_patches_per_s": round(n / (step_s + h2d_s), 2)}
    log("device copy %.0f GB/s; pair upload %.1f ms (%.1f GB/s)" % (copy_gbs, h2d_s * 1e3, h2d["GBs"]))

    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "traffic_source": ("rocprofv3 FETCH_SIZE + WRITE_SIZE, profiles/hbm_traffic_xcorr_strip.md"
                                   if traffic and not args.traffic else None),
                "kernel": dom_name, "launch_ms": round(per_launch_ms, 4),
                "patches_per_launch": patches_per_launch, "copy_peak_GBs": round(copy_gbs, 1),
                "frac_of_copy": round(achieved / copy_gbs, 5)}
    kernels = {k: {"ms_total": round(v[0], 4), "launches": v[1]} for k, v in kt.items()}
    path = {"hbm_GBs": round(BYTES_PER_PATCH * value / 1e9, 2),
            "hbm_frac": round(BYTES_PER_PATCH * value / 1e9 / HBM_PEAK_GBS, 5),
            "hbm_frac_of_copy": round(BYTES_PER_PATCH * value / 1e9 / copy_gbs, 5),
            "fp32_TFLOPs": round(FLOP_PER_PATCH * value / 1e12, 3),
            "fp32_frac": round(FLOP_PER_PATCH * value / 1e12 / FP32_PEAK_TFLOPS, 5)}
    log("kernels:", json.dumps(kernels))

    cpu = None
    parity = None
    if not args.no_cpu:
        import oracle
        oracle.lib()
        got = dout.cpu().numpy().view(xg.RESULT_DTYPE)
        nthreads, host = host_cpu()
        # host copy of the rows the sample needs
        op = oracle.make_params(64, 64)
        probe = min(n, max(nthreads, 16))
        rows_lo = int(min(w["my0"][:probe].min(), w["sy0"][:probe].min()))
        rows_hi = int(max(w["my0"][:probe].max(), w["sy0"][:probe].max())) + p.npy
        lo = max(rows_lo, 0)
        mh = m[lo:rows_hi].cpu().numpy()
        shh = s[lo:rows_hi].cpu().numpy()
        locs = np.stack([w["mx0"], w["my0"] - lo, w["sx0"], w["sy0"] - lo], 1).astype(np.int32)
        tp = time.perf_counter()
        oracle.run(op, mh, shh, locs[:probe], precision=32, nthreads=nthreads)
        per = (time.perf_counter() - tp) / probe
        sample = int(min(n, max(probe, args.cpu_seconds / per)))
        rows_hi = int(max(w["my0"][:sample].max(), w["sy0"][:sample].max())) + p.npy
        mh = m[lo:rows_hi].cpu().numpy()
        shh = s[lo:rows_hi].cpu().numpy()
        tp = time.perf_counter()
        ref = oracle.run(op, mh, shh, locs[:sample], precision=32, nthreads=nthreads)
        cpu_s = time.perf_counter() - tp
        # one thread on a shorter sample of the same locations
        one = int(min(sample, max(4, args.cpu_seconds / 4 / (per * nthreads))))
        tp = time.perf_counter()
        oracle.run(op, mh, shh, locs[:one], precision=32, nthreads=1)
        one_s = time.perf_counter() - tp
        cpu = {"value": round(sample / cpu_s, 2), "unit": "patches/s", "cores": nthreads, "kind": "port",
               "sample": "first %d of the %d locations (FP32 CPU restatement, OpenMP over patches, %.1f s); "
                         "1 thread: first %d (%.1f s)" % (sample, n, cpu_s, one, one_s),
               "threads_1_patches_per_s": round(one / one_s, 2),
               "threads_n_patches_per_s": round(sample / cpu_s, 2), **host}
        same = True
        for f in ("xoff", "yoff", "corr", "xfrac", "yfrac", "ip", "jp"):
            a, b = got[f][:sample], ref[f]
            if a.dtype.kind == "f":
                same &= bool(np.array_equal(a.view(np.uint32), b.view(np.uint32)))
            else:
                same &= bool(np.array_equal(a, b))
        parity = {"sample": sample, "bit_exact_vs_cpu": same}
        log("cpu baseline: %.1f patches/s on %d threads; parity %s" % (cpu["value"], nthreads, same))

    out = {"metric": METRIC, "value": round(value, 2), "unit": "patches/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(step_s * 1e3, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": workload, "patches": n, "xsearch": 64, "ysearch": 64, "ri": 2,
                      "interp_factor": 16, "batch": min(args.batch, n),
                      "batches": "serial" if args.serial else "overlapped on two contexts"},
           "roofline": roofline, "cpu_baseline": cpu, "path_roofline": path, "kernels": kernels,
           "step_stats": step_stats, "h2d": h2d, "parity": parity}
    eng.close()
    del m, s, dout
    torch.cuda.empty_cache()
    if not (args.small or args.workload == "small" or args.no_others):
        out["workloads"] = other_workloads(args, xg, synth, torch, dev)
    print(json.dumps(out), flush=True)
    return 0


def other_workloads(args, xg, synth, torch, dev):
    """Compact records of the batched, proposal, dense and time workloads, run
    after the headline one with few steps and a short CPU sample each (their
    full records: --workload batched / proposal / dense / time)."""
    sub = argparse.Namespace(**vars(args))
    sub.steps, sub.warmup = min(args.steps, 3), 1
    sub.cpu_seconds = max(1.0, args.cpu_seconds / 4)
    recs = {}
    for name, fn in (("batched", bench_batched), ("proposal", bench_proposal), ("dense", bench_dense),
                     ("time", bench_time)):
        t0 = time.time()
        r = fn(sub, xg, synth, torch, dev)
        torch.cuda.empty_cache()
        rec = {"workload": r["config"]["workload"], "patches": r["config"]["patches"], "value": r["value"],
               "unit": r["unit"], "ms_per_step": r["ms_per_step"], "steps": r["steps"],
               "cpu_baseline": None if not r.get("cpu_baseline") else
               {k: r["cpu_baseline"][k] for k in ("value", "cores", "kind", "sample")},
               "parity": r.get("parity"), "wall_s": round(time.time() - t0, 1)}
        if "accuracy" in r:
            rec["accuracy"] = r["accuracy"]
        if "cases" in r:
            rec["cases"] = [{k: c[k] for k in ("mode", "xsearch", "windows_per_s", "ms_per_step", "bit_exact_vs_cpu",
                                               "cpu_windows_per_s")} for c in r["cases"]]
        if "proposal" in r:
            pp = r["proposal"]
            rec["proposal"] = {k: pp[k] for k in ("plain_patches_per_s", "proposal_patches_per_s", "speedup",
                                                  "tau", "shrunken_fraction", "rmse_dx_px", "rmse_dy_px",
                                                  "cnn_ms")}
        recs[name] = rec
        log("%s: %.0f patches/s (%.1f s)" % (name, r["value"], rec["wall_s"]))
    return recs


def path_flops(npx, npy, ri, ifc):
    """FFT work of one patch, 5 N log2 N per transform: range interpolation
    (npx forward + ri*npx inverse per line, both images), three npx x npy 2-D
    transforms, and the 8 x 8 -> 8 ifc x 8 ifc high-res interpolation."""
    lg = math.log2
    rng = 2 * npy * (5 * npx * lg(npx) + (5 * ri * npx * lg(ri * npx) if ri > 1 else 0))
    corr = 3 * 5 * npx * npy * lg(npx * npy)
    L = 8 * ifc
    hres = (8 + L) * (5 * 8 * 3 + 5 * L * lg(L))
    return rng + corr + hres


def batched_workload(xg, synth, torch, dev, npairs, seed=3):
    """The batched uneven workload: npairs complex-int16 pairs, sizes uniform in [2048, 12000] range x
    [4096, 28000] azimuth, each with its own constant-plus-linear offset field
    (|d| <= 20 px; the constant part at the image centre goes into rshift /
    ashift as a GMTSAR PRM would carry it), coherence in [0.3, 0.95], patch
    size xsearch = ysearch in {32, 64, 128} (64^2, 128^2, 256^2 templates),
    interpolation factor 8 or 16, grids of 20-200 x 50-400 locations."""
    rng = np.random.default_rng(seed)
    work = []
    for i in range(npairs):
        nx, ny = int(rng.integers(2048, 12001)), int(rng.integers(4096, 28001))
        xs, ifc = int(rng.choice([32, 64, 128])), int(rng.choice([8, 16]))
        nxl, nyl = int(rng.integers(20, 201)), int(rng.integers(50, 401))
        coh = float(rng.uniform(0.3, 0.95))
        a, b = float(rng.uniform(-12, 12)), float(rng.uniform(-8, 8)) / nx
        c, e = float(rng.uniform(-12, 12)), float(rng.uniform(-8, 8)) / ny
        m, s = synth.strip_pair_torch(nx, ny, lambda x, a=a, b=b: a + b * x, lambda y, c=c, e=e: c + e * y,
                                      coh_range=(coh, coh), seed=100 + i, device=dev)
        rshift, ashift = int(round(a + b * nx / 2)), int(round(c + e * ny / 2))
        p = xg.make_params(xs, xs, interp_factor=ifc)
        x, y = xg.get_locations(nx, ny, nxl, nyl, xs, xs)
        w = xg.make_windows(x, y, p.npx, p.npy, rshift, ashift)
        out = torch.zeros(len(w) * xg.RESULT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
        work.append(dict(params=p, master=m, slave=s, windows=w, out=out, rshift=rshift, ashift=ashift,
                         truth=(a, b, c, e), coh=coh))
        torch.cuda.empty_cache()
    return work


def bench_batched(args, xg, synth, torch, dev):
    t0 = time.time()
    work = batched_workload(xg, synth, torch, dev, args.pairs)
    torch.cuda.synchronize()
    n = sum(len(q["windows"]) for q in work)
    log("%d synthetic pairs, %d patches, ready in %.1f s" % (len(work), n, time.time() - t0))
    t0 = time.time()
    mu = xg.Multi(work, tune=not args.no_tune, tune_radix=not args.no_tune)
    classes, pc = mu.classes()
    log("graph built (%d classes%s) in %.1f s" % (len(classes), ", autotuned" if not args.no_tune else "",
                                                 time.time() - t0))
    for _ in range(args.warmup):
        mu.run(flags=xg.NO_SYNC)
    mu.sync()
    trace_marker(torch)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        mu.run(flags=xg.NO_SYNC)
    mu.sync()
    torch.cuda.synchronize()
    step_s = (time.perf_counter() - t1) / args.steps
    trace_marker(torch)
    value = n / step_s
    log("timed %d steps: %.3f ms/step, %.0f patches/s" % (args.steps, step_s * 1e3, value))
    nbytes = sum(len(q["windows"]) * (2 * q["params"].npx * q["params"].npy * 4 + 20) for q in work)
    flops = sum(len(q["windows"]) * path_flops(q["params"].npx, q["params"].npy, q["params"].ri,
                                                q["params"].interp_factor) for q in work)
    per_class = []
    for k, (p, t) in enumerate(classes):
        nk = int(sum(len(work[i]["windows"]) for i in range(len(work)) if pc[i] == k))
        per_class.append({"xsearch": p.xsearch, "interp_factor": p.interp_factor, "pairs": int((pc == k).sum()),
                          "patches": nk, "tuning": t})
    achieved = nbytes / step_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "kernel": "whole graph (all classes)", "launch_ms": round(step_s * 1e3, 4)}
    path = {"hbm_GBs": round(achieved, 2), "hbm_frac": round(achieved / HBM_PEAK_GBS, 5),
            "fp32_TFLOPs": round(flops / step_s / 1e12, 3),
            "fp32_frac": round(flops / step_s / 1e12 / FP32_PEAK_TFLOPS, 5)}
    # offsets against the injected fields (a property of the whole run)
    errs = []
    for q in work:
        got = q["out"].cpu().numpy().view(xg.RESULT_DTYPE)
        p = q["params"]
        dx, dy = xg.offsets(got, p.ri, q["rshift"], q["ashift"])
        w = q["windows"]
        xc = w["mx0"] + p.npx // 2
        yc = w["my0"] + p.npy // 2
        a, b, c, e = q["truth"]
        ok = got["corr"] > 20
        errs.append(np.stack([(dx - (a + b * xc))[ok], (dy - (c + e * yc))[ok]], 1))
    errs = np.concatenate(errs)
    accuracy = {"median_abs_dx_err": round(float(np.median(np.abs(errs[:, 0]))), 4),
                "median_abs_dy_err": round(float(np.median(np.abs(errs[:, 1]))), 4), "patches_corr_gt_20": len(errs)}
    cpu, parity = None, None
    if not args.no_cpu:
        cpu, parity = batched_cpu_baseline(args, xg, work, classes, pc)
    out = {"metric": METRIC, "value": round(value, 2), "unit": "patches/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(step_s * 1e3, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "batched uneven: %d cint16 pairs (2048-12000 x 4096-28000), xsearch 32/64/128, "
                                  "interp 8/16, one HIP graph launch per step" % len(work),
                      "patches": n, "classes": per_class},
           "roofline": roofline, "cpu_baseline": cpu, "path_roofline": path, "accuracy": accuracy, "parity": parity}
    mu.close()
    return out


def bench_proposal(args, xg, synth, torch, dev):
    """Learned-proposal form beside plain xcorr on the 10800 x 27648 strip
    (40,000 windows of 256 x 256).  Both are timed end to end per step with
    host results (plain: one correlate call; proposal: box-downsampling, CNN
    proposals, the shrunken and fallback searches on concurrent streams, host
    composition of the offsets)."""
    import importlib
    prop = importlib.import_module("gmtsar_gpu_amd.proposal")
    params = np.load(args.weights, allow_pickle=False)
    n_range, n_az, nxl, nyl = 10800, 27648, 100, 400
    m, s = synth.strip_pair_torch(n_range, n_az, lambda x: 2.1 + 1e-4 * x, lambda y: -7.6 + 3e-4 * y, device="cuda")
    torch.cuda.synchronize()
    p = xg.make_params(64, 64)
    x, y = xg.get_locations(n_range, n_az, nxl, nyl, 64, 64)
    w = xg.make_windows(x, y, p.npx, p.npy)
    n = len(w)
    pr = prop.Proposer(params, tau=args.tau, tau_skip=args.tau_skip, max_batch=min(args.batch, n),
                       xs_small=args.xs_small, verify=args.verify, corr_min=args.corr_min)
    pr.set_images(m, s)
    torch.cuda.synchronize()

    def timed(fn):
        for _ in range(args.warmup):
            r = fn()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t1) / args.steps, r

    t_plain, plain = timed(lambda: pr.full.correlate(w))
    pdx, pdy = prop.compose(plain, w, 2)
    t_prop, (dx, dy, res, path) = timed(lambda: pr.run(w))
    t_pred, pred = timed(lambda: pr.predict(w))
    ex, ey = dx - pdx, dy - pdy
    summary = {}
    for tau in (0.3, 0.5, 0.7, 0.9):
        for mode in ("", "_skip", "_verify"):
            pr.tau, pr.tau_skip = tau, (tau if mode == "_skip" else 1.01)
            pr.verify, pr.corr_min = mode == "_verify", args.corr_min
            t3 = time.perf_counter()
            dxx, dyy, _, pth = pr.run(w, pred=pred)
            t3 = time.perf_counter() - t3 + t_pred
            e = np.concatenate([dxx - pdx, dyy - pdy])
            summary["tau%.1f%s" % (tau, mode)] = {
                "shrunken": round(float(((pth == 1) | (pth == 2)).mean()), 4),
                "rmse_px": round(float(np.sqrt(np.mean(e ** 2))), 4),
                "speedup": round(t_plain / t3, 3)}
    pr.tau, pr.tau_skip, pr.verify, pr.corr_min = args.tau, args.tau_skip, args.verify, args.corr_min
    conf = pred[2]
    sel = (path == 1) | (path == 2)
    pe = np.abs(pred[0] - pdx), np.abs(pred[1] - pdy)
    proposal = {
        "plain_patches_per_s": round(n / t_plain, 2), "proposal_patches_per_s": round(n / t_prop, 2),
        "speedup": round(t_plain / t_prop, 4), "cnn_ms": round(t_pred * 1e3, 3),
        "tau": args.tau, "tau_skip": args.tau_skip, "verify": args.verify, "corr_min": args.corr_min,
        "shrunken_fraction": round(float(((path == 1) | (path == 2)).mean()), 4),
        "verified_redone_fraction": round(float((path == 3).mean()), 4),
        "skipped_highres_fraction": round(float((path == 2).mean()), 4),
        "rmse_dx_px": round(float(np.sqrt(np.mean(ex ** 2))), 4),
        "rmse_dy_px": round(float(np.sqrt(np.mean(ey ** 2))), 4),
        "max_abs_err_px": round(float(max(np.abs(ex).max(), np.abs(ey).max())), 4),
        "same_integer_offsets": round(float(np.mean((np.rint(2 * dx) == np.rint(2 * pdx)) &
                                                    (np.rint(dy) == np.rint(pdy)))), 4),
        "cnn_median_abs_err_px": [round(float(np.median(pe[0])), 3), round(float(np.median(pe[1])), 3)],
        "mean_confidence": round(float(conf.mean()), 4), "tau_sweep": summary}
    log("plain %.0f patches/s, proposal %.0f patches/s (x%.2f), shrunken %.1f %%, rmse %.4f / %.4f px"
        % (n / t_plain, n / t_prop, t_plain / t_prop, 100 * sel.mean(), proposal["rmse_dx_px"],
           proposal["rmse_dy_px"]))
    out = {"metric": METRIC, "value": round(n / t_prop, 2), "unit": "patches/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(t_prop * 1e3, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32 (CNN: bf16 MFMA, f32 accumulate)",
           "data": "synthetic",
           "config": {"workload": "learned proposal beside plain xcorr, 10800 x 27648 cint16 pair, -nx 100 -ny 400 "
                                  "-xsearch 64 -ysearch 64, interp 16; shrunken search xsearch %d" % args.xs_small,
                      "patches": n, "weights": os.path.relpath(args.weights, ROOT)},
           "roofline": None, "cpu_baseline": None, "proposal": proposal}
    pr.close()
    return out


def bench_dense(args, xg, synth, torch, dev):
    """Dense offset tracking: a 10800 x 27648 pair whose offset field is a sum
    of 8 random 2-D sinusoids per component (|d| <= 15 px, wavelengths 133-667
    px = 2-10 km at 15 m pixels) with spatially correlated coherence in
    [0.15, 0.95], the slave resampled per sample (synth.field_pair_torch);
    -xsearch 32 -ysearch 32 (128 x 128 windows, 64 x 64 templates) on the grid
    -nx 337 -ny 864 (31-px steps, 291,168 locations), interp 8; the exact
    time-domain search (-time) on one checkerboard half of the locations and
    the frequency path on the other, both in one graph."""
    rng = np.random.default_rng(5)
    field = synth.sinusoid_field(rng)
    coh = synth.smooth_coherence(rng)
    n_range, n_az, nxl, nyl = 10800, 27648, 337, 864
    t0 = time.time()
    m, s = synth.field_pair_torch(n_range, n_az, field, coh, device=dev)
    torch.cuda.synchronize()
    log("dense pair ready in %.1f s" % (time.time() - t0))
    x, y = xg.get_locations(n_range, n_az, nxl, nyl, 32, 32)
    i, j = np.meshgrid(np.arange(nyl), np.arange(nxl), indexing="ij")
    half = ((i + j) & 1).ravel() == 1
    work = []
    for cf, sel in ((xg.TIME, half), (xg.FREQ, ~half)):
        p = xg.make_params(32, 32, interp_factor=8, corr_flag=cf)
        w = xg.make_windows(x[sel], y[sel], p.npx, p.npy)
        out = torch.zeros(len(w) * xg.RESULT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
        work.append(dict(params=p, master=m, slave=s, windows=w, out=out, rshift=0, ashift=0,
                         xy=(x[sel], y[sel])))
    n = sum(len(q["windows"]) for q in work)
    t0 = time.time()
    mu = xg.Multi(work, tune=not args.no_tune, tune_radix=not args.no_tune)
    classes, pc = mu.classes()
    log("graph built (%d classes) in %.1f s" % (len(classes), time.time() - t0))
    for _ in range(args.warmup):
        mu.run(flags=xg.NO_SYNC)
    mu.sync()
    trace_marker(torch)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        mu.run(flags=xg.NO_SYNC)
    mu.sync()
    torch.cuda.synchronize()
    step_s = (time.perf_counter() - t1) / args.steps
    trace_marker(torch)
    value = n / step_s
    log("timed %d steps: %.3f ms/step, %.0f patches/s" % (args.steps, step_s * 1e3, value))
    # per-class time from one single-class graph each
    per_class = []
    for k, (p, t) in enumerate(classes):
        q = work[int(np.nonzero(pc == k)[0][0])]
        with xg.Multi([q], tune=False) as one:
            one.run()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            for _ in range(max(1, args.steps // 2)):
                one.run(flags=xg.NO_SYNC)
            one.sync()
            one_eng_t = (time.perf_counter() - t2) / max(1, args.steps // 2)
        per_class.append({"corr_flag": p.corr_flag, "patches": len(q["windows"]), "ms": round(one_eng_t * 1e3, 3),
                          "patches_per_s": round(len(q["windows"]) / one_eng_t, 1), "tuning": t})
    accuracy = dense_accuracy(xg, synth, field, work)
    log("accuracy:", json.dumps(accuracy))
    cpu, parity = None, None
    if not args.no_cpu:
        cpu, parity = batched_cpu_baseline(args, xg, work, classes, pc)
    nbytes = n * (2 * 128 * 128 * 4 + 20)
    out = {"metric": METRIC, "value": round(value, 2), "unit": "patches/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(step_s * 1e3, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32 (time search: exact integer sums in f64)",
           "data": "synthetic",
           "config": {"workload": "dense offset tracking, 10800 x 27648 cint16 pair, smooth sinusoid field "
                                  "<= 15 px, -xsearch 32 -ysearch 32 -nx 337 -ny 864, interp 8, half -time half "
                                  "frequency, one HIP graph launch per step", "patches": n, "classes": per_class},
           "roofline": {"bound": "hbm", "achieved": round(nbytes / step_s / 1e9, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(nbytes / step_s / 1e9 / HBM_PEAK_GBS, 5), "traffic": None,
                        "kernel": "whole graph (both classes)", "launch_ms": round(step_s * 1e3, 4)},
           "cpu_baseline": cpu, "accuracy": accuracy, "parity": parity}
    mu.close()
    return out


def bench_time(args, xg, synth, torch, dev):
    """The time-domain searches (do_time_int_xcorr.cu:103-189: -time, the exact
    integer NCC over every lag, and -time4, Gatelli's fourth-power form) on the
    strip workload's 10800 x 27648 pair at -nx 100 -ny 400: -xsearch/-ysearch
    32 (128 x 128 windows, interp 8) and 64 (256 x 256 windows, the reference's
    default, interp 16), each on the int8 matrix cores.  Per case: windows/s
    over the timed steps, per-kernel times, and the first windows against the
    FP32 CPU restatement bit for bit (its rate is the CPU baseline).  `value`
    is the -time rate at xsearch 64."""
    n_range, n_az, nxl, nyl = 10800, 27648, 100, 400
    t0 = time.time()
    m, s = synth.strip_pair_torch(n_range, n_az, lambda x: 2.1 + 1e-4 * x, lambda y: -7.6 + 3e-4 * y,
                                  device="cuda")
    torch.cuda.synchronize()
    log("time-mode pair ready in %.1f s" % (time.time() - t0))
    nthreads, host = host_cpu()
    cases = []
    for xs, ifc in ((32, 8), (64, 16)):
        for cf, mode in ((xg.TIME, "-time"), (xg.TIME4, "-time4")):
            p = xg.make_params(xs, xs, interp_factor=ifc, corr_flag=cf)
            x, y = xg.get_locations(n_range, n_az, nxl, nyl, xs, xs)
            w = xg.make_windows(x, y, p.npx, p.npy)
            n = len(w)
            dout = torch.zeros(n * xg.RESULT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
            flags = xg.NO_SYNC | (xg.SERIAL if args.serial else 0)
            with xg.Engine(p) as eng:
                eng.set_images(m, s)
                for _ in range(args.warmup):
                    eng.correlate(w, out=dout, flags=flags)
                eng.sync()
                torch.cuda.synchronize()
                trace_marker(torch)
                t1 = time.perf_counter()
                for _ in range(args.steps):
                    eng.correlate(w, out=dout, flags=flags)
                eng.sync()
                step_s = (time.perf_counter() - t1) / args.steps
                trace_marker(torch)
                eng.correlate(w, out=dout, flags=xg.TIMING)
                kt = {k: round(v[0], 3) for k, v in eng.kernel_times().items()}
            got = dout.cpu().numpy().view(xg.RESULT_DTYPE)
            rec = {"mode": mode, "xsearch": xs, "interp": ifc, "windows": n, "ms_per_step": round(step_s * 1e3, 3),
                   "windows_per_s": round(n / step_s, 1), "kernel_ms": kt}
            if not args.no_cpu:
                rec.update(time_cpu_sample(args, xg, p, w, got, m, s, nthreads))
            log("%s xsearch %d: %.0f windows/s%s" % (mode, xs, n / step_s,
                ", CPU %.1f windows/s, bit-exact %s" % (rec["cpu_windows_per_s"], rec["bit_exact_vs_cpu"])
                if "cpu_windows_per_s" in rec else ""))
            rec.setdefault("cpu_windows_per_s", None)
            rec.setdefault("bit_exact_vs_cpu", None)
            cases.append(rec)
            del dout
    del m, s
    head = next(c for c in cases if c["mode"] == "-time" and c["xsearch"] == 64)
    same = [c["bit_exact_vs_cpu"] for c in cases if c["bit_exact_vs_cpu"] is not None]
    cpu = None
    if head["cpu_windows_per_s"]:
        cpu = {"value": head["cpu_windows_per_s"], "unit": "patches/s", "cores": nthreads, "kind": "port",
               "sample": "first %d windows of -time at xsearch 64 (FP32 CPU restatement, OpenMP over patches); "
                         "every case's own sample is in cases" % head["cpu_sample"], **host}
    return {"metric": METRIC, "value": head["windows_per_s"], "unit": "patches/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms_per_step"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "int8 matrix-core limb products, exact integer sums", "data": "synthetic",
            "config": {"workload": "time-domain search (-time, -time4) on the 10800 x 27648 cint16 strip pair, "
                                   "-nx 100 -ny 400, -xsearch/-ysearch 32 (interp 8) and 64 (interp 16)",
                       "patches": head["windows"]},
            "cases": cases, "cpu_baseline": cpu,
            "parity": {"sample": sum(c.get("cpu_sample", 0) for c in cases),
                       "bit_exact_vs_cpu": all(same) if same else None}}


def time_cpu_sample(args, xg, p, w, got, m, s, nthreads):
    """The first windows of one time-mode case on the FP32 CPU restatement:
    a probe of one window per thread sizes a sample of about a quarter of the
    CPU budget; bit-exactness of the peaks, offsets and correlations."""
    import oracle
    oracle.lib()
    op = oracle.make_params(p.xsearch, p.ysearch, ri=p.ri, interp_flag=p.interp_flag,
                            interp_factor=p.interp_factor, corr_flag=p.corr_flag)

    def run(cnt):
        lo = max(int(min(w["my0"][:cnt].min(), w["sy0"][:cnt].min())), 0)
        hi = int(max(w["my0"][:cnt].max(), w["sy0"][:cnt].max())) + p.npy
        mh, sh = m[lo:hi].cpu().numpy(), s[lo:hi].cpu().numpy()
        locs = np.stack([w["mx0"], w["my0"] - lo, w["sx0"], w["sy0"] - lo], 1).astype(np.int32)[:cnt]
        tp = time.perf_counter()
        ref = oracle.run(op, mh, sh, locs, precision=32, nthreads=nthreads)
        return ref, time.perf_counter() - tp

    probe = min(len(w), nthreads)
    ref, ts = run(probe)
    cnt = int(min(len(w), max(probe, (args.cpu_seconds / 4) / (ts / probe))))
    if cnt > probe:
        ref, ts = run(cnt)
    same = True
    for f in ("xoff", "yoff", "corr", "xfrac", "yfrac", "ip", "jp"):
        a, b = got[f][:cnt], ref[f]
        same &= bool(np.array_equal(a.view(np.uint32), b.view(np.uint32)) if a.dtype.kind == "f"
                     else np.array_equal(a, b))
    return {"cpu_sample": cnt, "cpu_windows_per_s": round(cnt / ts, 2), "bit_exact_vs_cpu": same}


def dense_accuracy(xg, synth, field, work):
    """Offsets against the generator's exact displacement, averaged over each
    window's 64 x 64 template (synth.window_truth), per class (-time and
    frequency), for windows with corr > 20; also split by how much the field
    itself changes across the template (the specified field changes by a
    median 5.5 px in range across 64 px, which smears any window match)."""
    out = {}
    for q in work:
        got = q["out"].cpu().numpy().view(xg.RESULT_DTYPE)
        dx, dy = xg.offsets(got, 2)
        tx, ty, sx, sy = synth.window_truth(field, q["xy"][0], q["xy"][1], 32)
        ok = got["corr"] > 20
        ex, ey = np.abs(dx - tx), np.abs(dy - ty)
        spread = np.maximum(sx, sy)
        rec = {"windows": int(len(got)), "corr_gt_20": int(ok.sum()),
               "median_abs_dx_err": round(float(np.median(ex[ok])), 4),
               "median_abs_dy_err": round(float(np.median(ey[ok])), 4),
               "median_field_spread_px": round(float(np.median(spread)), 3)}
        for lo, hi in ((0, 2), (2, 5), (5, 1e9)):
            sel = ok & (spread >= lo) & (spread < hi)
            if sel.any():
                rec["spread_%g_%g" % (lo, min(hi, 99))] = {
                    "windows": int(sel.sum()), "median_abs_dx_err": round(float(np.median(ex[sel])), 4),
                    "median_abs_dy_err": round(float(np.median(ey[sel])), 4)}
        out["time" if q["params"].corr_flag == xg.TIME else "freq"] = rec
    return out


def batched_cpu_baseline(args, xg, work, classes, pc):
    """FP32 CPU restatement per class on a sample of the class's largest pair
    (with the class's tuned radix plans), weighted by the class's share of the
    workload; the sample doubles as a bit-exactness check."""
    import oracle
    oracle.lib()
    nthreads, host = host_cpu()
    total_cpu_s, total_n, sample_n, same = 0.0, 0, 0, True
    budget = args.cpu_seconds / max(1, len(classes))
    for k, (p, t) in enumerate(classes):
        idx = [i for i in range(len(work)) if pc[i] == k]
        q = work[max(idx, key=lambda i: len(work[i]["windows"]))]
        w = q["windows"]
        op = oracle.make_params(p.xsearch, p.ysearch, ri=p.ri, interp_flag=p.interp_flag,
                                interp_factor=p.interp_factor, corr_flag=p.corr_flag)
        got = q["out"].cpu().numpy().view(xg.RESULT_DTYPE)
        plans = {p.npx: t["radix_x"]}
        plans[p.npy] = t["radix_y"]

        def run_sample(cnt):
            lo = max(int(min(w["my0"][:cnt].min(), w["sy0"][:cnt].min())), 0)
            hi = int(max(w["my0"][:cnt].max(), w["sy0"][:cnt].max())) + p.npy
            mh, sh = q["master"][lo:hi].cpu().numpy(), q["slave"][lo:hi].cpu().numpy()
            locs = np.stack([w["mx0"], w["my0"] - lo, w["sx0"], w["sy0"] - lo], 1).astype(np.int32)[:cnt]
            with oracle.fft_plans(plans):
                tp = time.perf_counter()
                ref = oracle.run(op, mh, sh, locs, precision=32, nthreads=nthreads)
                return ref, time.perf_counter() - tp

        probe = min(len(w), nthreads)
        _, ts = run_sample(probe)
        cnt = int(min(len(w), max(probe, budget / (ts / probe))))
        ref, ts = run_sample(cnt)
        for f in ("xoff", "yoff", "corr", "xfrac", "yfrac", "ip", "jp"):
            a, b = got[f][:cnt], ref[f]
            same &= bool(np.array_equal(a.view(np.uint32), b.view(np.uint32)) if a.dtype.kind == "f"
                         else np.array_equal(a, b))
        nk = sum(len(work[i]["windows"]) for i in idx)
        total_cpu_s += nk * ts / cnt
        total_n += nk
        sample_n += cnt
    cpu = {"value": round(total_n / total_cpu_s, 2), "unit": "patches/s", "cores": nthreads, "kind": "port",
           "sample": "%d patches over %d classes (largest pair of each), per-class CPU time weighted by the class's "
                     "share of the %d patches (FP32 CPU restatement, OpenMP over patches)"
                     % (sample_n, len(classes), total_n), **host}
    log("cpu baseline: %.1f patches/s on %d threads; parity %s" % (cpu["value"], nthreads, same))
    return cpu, {"sample": sample_n, "bit_exact_vs_cpu": same}


if __name__ == "__main__":
    sys.exit(main())
