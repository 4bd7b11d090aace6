"""Benchmark of the pattern-offset hot path on one MI355X.

Default workload (fft_f32_873): one reference and 256 deformed 873 x 873 uint16
band patterns (synthetic), 400 ROIs on a 20 x 20 grid, radius 32 (64 x 64 slices,
128 x 128 FFT), FFT policy in float, fitsize 5, batch 64.  The 256 patterns are
resident in HBM before timing; one step is one batch of 64 patterns through the
whole pipeline (prepare, FFT correlation, arg-max, gather, fit, offsets) via the
device-resident entry point.  Throughput counts patterns (and ROI offsets) per
second.

Also reported:
  roofline      the FFT pipeline kernel's algorithmic FLOP rate (HIP events on its
                stream) against the FP32 peak; HBM traffic per launch from the
                committed rocprofv3 counter summary when present;
  cpu_baseline  the CPU oracle (direct correlation, FP64, OpenMP on the cores this
                process is granted, and on one core) on a bounded sample of the same
                workload on this host, with the host's CPU model, nproc and SMT, and
                a CPU FFT form (scipy pocketfft) for FFT workloads;
  accuracy      max |GPU - FP64 oracle| offset over the sample's well-conditioned
                ROIs (FP64 peak margin above float rounding, fitted maximum inside
                the neighbourhood: the parity tests' criterion), the fraction of ROIs
                that covers, the integer-peak agreement and the all-ROI maximum;
  served        (default workload) the same plan behind the request server: per-pattern
                p50/p99 latency at 2000 requests/s offered, saturation patterns/s, and
                the plain FP64 brute form served the same way;
  precision_form (default workload) the precision form's shape (100 ROIs of 128 x 128,
                fitsize 7, batch 128): refined fp16 matrix-core brute (the default
                mode) against the plain FP64 brute form;
  batch latency p50/p99 time for one batch to complete when submitted alone.
"""
import argparse
import json
import math
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "patterns/s + ROI offsets/s at ≤0.01 px err vs FP64; p50/p99 pattern latency"
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
MFMA_16BIT_PEAK_TFLOPS = 2500.0  # dense bf16 / fp16
MFMA_MODES = {"bf16": (0, 1), "bf16_split": (1, 3), "f16": (2, 1), "f16_split": (3, 3)}  # code, products


def fft_flops_per_roi(m):
    # 2 transforms x 2.5 M^2 log2(M^2) + 6 flops per half-spectrum product
    return 2 * 2.5 * m * m * math.log2(m * m) + 6 * (m // 2 + 1) * m


def brute_flops_per_roi(s, cross):
    # 2 flops per multiply-add over every overlapping pixel pair of every shift
    sx, sy = s
    rx, ry = (cross[0] - 1) // 2, (cross[1] - 1) // 2
    ox = sum(sx - abs(d) for d in range(-rx, cross[0] - rx) if abs(d) < sx)
    oy = sum(sy - abs(d) for d in range(-ry, cross[1] - ry) if abs(d) < sy)
    return 2.0 * ox * oy


def read_traffic(workload, mode=None):
    """HBM bytes per launch of the dominant kernel from profiles/traffic_<workload>[_<mode>].txt
    (FETCH_SIZE and WRITE_SIZE in KiB per dispatch, calibrated with a copy of known
    size).  None when no summary is committed."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}{'_' + mode if mode else ''}.txt")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    m = re.search(r"^hbm_bytes_per_launch:\s*([0-9.eE+]+)", text, re.M)
    return float(m.group(1)) if m else None


def cpu_threads():
    """Every core this process may run on, capped by OMP_NUM_THREADS when the host sets
    it (a GPU host grants each GPU its share of the cores that way)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def host_info():
    """nproc, the cores this process may run on, the CPU model and its SMT state."""
    info = {"nproc": os.cpu_count(), "affinity_cores": len(os.sched_getaffinity(0)), "cpu_model": None,
            "smt_active": None, "threads_per_core": None}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    info["cpu_model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        with open("/sys/devices/system/cpu/smt/active") as f:
            info["smt_active"] = f.read().strip() == "1"
    except OSError:
        pass
    try:
        with open("/sys/devices/system/cpu/cpu0/topology/thread_siblings_list") as f:
            n = 0
            for part in f.read().strip().split(","):
                a, _, b = part.partition("-")
                n += int(b) - int(a) + 1 if b else 1
            info["threads_per_core"] = n
    except (OSError, ValueError):
        pass
    return info


def cpu_baseline(cfg, ps, begins, slice_size, cross, seconds, threads):
    """The CPU oracle (the reference's CPU path restated: sums, prepare, direct
    correlation, arg-max, gather, subpixel fit; FP64, OpenMP over ROIs) on whole
    patterns of the workload until `seconds` are used.  Returns the patterns done, the
    time, and the oracle's offsets, coefficients, arg-max indices and peak statistics."""
    from oracle import oracle as O

    O.set_threads(threads)
    policy = O.POLICY_FFT if cfg["policy"] == "fft" else O.POLICY_BRUTE
    done, t_used, offs, coefs, idx, stats = 0, 0.0, [], [], [], []
    while done < ps.pics.shape[0] and (done == 0 or t_used < seconds):
        t0 = time.perf_counter()
        o, c, m, _, st = O.get_offset(ps.pics[done], ps.ref, ps.size, begins, slice_size, cross, 1, cfg["fit"], policy,
                                      np.float64, return_intermediates=True, return_stats=True)
        t_used += time.perf_counter() - t0
        offs.append(o)
        coefs.append(c)
        idx.append(m["index"])
        stats.append(st)
        done += 1
    return done, t_used, np.concatenate(offs), np.concatenate(coefs), np.concatenate(idx), np.concatenate(stats)


def cpu_single_thread(cfg, ps, begins, slice_size, cross, seconds):
    """The same CPU oracle on one thread, over ROIs of the first pattern in groups of 20
    until `seconds` are used; patterns/s scaled from the ROIs done."""
    from oracle import oracle as O

    O.set_threads(1)
    policy = O.POLICY_FFT if cfg["policy"] == "fft" else O.POLICY_BRUTE
    done, t_used = 0, 0.0
    while done < len(begins) and (done == 0 or t_used < seconds):
        sub = begins[done:done + 20]
        t0 = time.perf_counter()
        O.get_offset(ps.pics[0], ps.ref, ps.size, sub, slice_size, cross, 1, cfg["fit"], policy, np.float64)
        t_used += time.perf_counter() - t0
        done += len(sub)
    return done / len(begins) / t_used, done, t_used


def cpu_fft_baseline(cfg, ps, begins, slice_size, seconds, threads):
    """A CPU FFT form of the pipeline (the reference's CPU path has no FFT): the
    oracle's sums and preparation, scipy's pocketfft R2C of every 2s x 2s zero-padded
    slice (reference spectra computed once), product with the conjugate spectrum,
    C2R, then the oracle's arg-max, gather (FFT position policy) and fit; FP64,
    `threads` FFT workers.  Returns patterns/s and the patterns done."""
    import scipy.fft as sf
    from oracle import oracle as O

    O.set_threads(threads)
    sx, sy = slice_size
    mx, my = 2 * sx, 2 * sy
    nb, fit = len(begins), cfg["fit"]
    hx, hy = O.hann(sx), O.hann(sy)
    sr = O.sums(ps.ref.ravel(), begins, ps.size, slice_size, 1)
    r = O.prepare(ps.ref.ravel(), hx, hy, sr, begins, ps.size, slice_size, None, 1).reshape(nb, sy, sx)
    R = np.conj(sf.rfft2(r, s=(my, mx), workers=threads))
    done, t_used = 0, 0.0
    while done < ps.pics.shape[0] and (done == 0 or t_used < seconds):
        t0 = time.perf_counter()
        pic = ps.pics[done].ravel()
        sp = O.sums(pic, begins, ps.size, slice_size, 1)
        p = O.prepare(pic, hx, hy, sp, begins, ps.size, slice_size, None, 1).reshape(nb, sy, sx)
        c = sf.irfft2(sf.rfft2(p, s=(my, mx), workers=threads) * R, s=(my, mx), workers=threads)
        c = np.ascontiguousarray((c * (mx * my)).reshape(nb, -1))  # unnormalised, as a C2R
        mxi = O.maxarg(c, mx * my, nb)
        nbh = O.extract(c, mxi, fit, (mx - 1, my - 1), nb, O.POLICY_FFT)
        O.subpixel(nbh, fit)
        t_used += time.perf_counter() - t0
        done += 1
    return done / t_used, done


ACCURACY_DEFINITION = ("max |GPU - FP64 oracle| offset (px) over the well-conditioned ROIs: the FP64 "
                       "correlation peak beats the runner-up by more than 1e-4 of the largest magnitude and "
                       "the FP64 quadratic fit has its maximum inside the F x F neighbourhood "
                       "(oracle.well_conditioned, the parity tests' criterion)")


def accuracy(gpu_offs, gpu_index, cpu_offs, cpu_coefs, cpu_index, stats, fit):
    """The offset error against FP64 under the one definition the parity tests use, the
    fraction of ROIs it covers, the integer-peak agreement and the all-ROI maximum."""
    from oracle import oracle as O

    ok = O.well_conditioned(cpu_coefs, fit, stats)
    err = np.abs(gpu_offs - cpu_offs).max(axis=1)
    same = gpu_index == cpu_index
    return {"definition": ACCURACY_DEFINITION, "rois": int(len(err)),
            "max_err_px": float(err[ok].max()) if ok.any() else None,
            "mean_err_px": float(err[ok].mean()) if ok.any() else None,
            "well_conditioned_roi_fraction": round(float(ok.mean()), 4),
            "integer_peak_match_fraction": round(float(same.mean()), 4),
            "integer_peak_match_fraction_well_conditioned": round(float(same[ok].mean()), 4) if ok.any() else None,
            "max_err_px_all_rois": float(np.nanmax(err))}


def fp64_fft_peak_stats(pics, ref, src, begins, slice_size, threads):
    """Per slice the largest value, the runner-up and the largest magnitude of the FP64
    FFT-policy correlation ((2s)^2 circular buffer) of every pattern in `pics`, from
    the oracle's preparation and scipy's double FFT - the statistics
    oracle.well_conditioned reads, at FFT speed for large subsets."""
    import scipy.fft as sf
    from oracle import oracle as O

    sx, sy = slice_size
    nb = len(begins)
    hx, hy = O.hann(sx), O.hann(sy)
    sr = O.sums(ref.ravel(), begins, src, slice_size, 1)
    R = np.conj(sf.rfft2(O.prepare(ref.ravel(), hx, hy, sr, begins, src, slice_size, None, 1).reshape(nb, sy, sx),
                         s=(2 * sy, 2 * sx), workers=threads))
    out = []
    for pic in pics:
        sp = O.sums(pic.ravel(), begins, src, slice_size, 1)
        p = O.prepare(pic.ravel(), hx, hy, sp, begins, src, slice_size, None, 1).reshape(nb, sy, sx)
        c = sf.irfft2(sf.rfft2(p, s=(2 * sy, 2 * sx), workers=threads) * R, s=(2 * sy, 2 * sx),
                      workers=threads).reshape(nb, -1)
        top = np.partition(c, -2, axis=1)[:, -2:]
        out.append(np.stack([top.max(axis=1), top.min(axis=1), np.abs(c).max(axis=1)], axis=1))
    return np.concatenate(out)


def copy_calibration(mb):
    import ctypes
    import emida_amd
    from emida_amd import _capi as C

    n = int(mb * 1024 * 1024)
    a, b = emida_amd.DeviceBuffer(n), emida_amd.DeviceBuffer(n)
    a.zero()
    for _ in range(5):
        C.check(C.lib().emida_memcpy_d2d(b.ptr, a.ptr, n))
    emida_amd.synchronize()
    print(json.dumps({"copy_calibration_bytes": n, "copies": 5}))


def run_load(srv, pids, pics, n, rate, seed, submitters=1):
    """n single-pattern requests to the plans `pids` (chosen uniformly) of a request
    server; Poisson arrivals at `rate` per second from one thread, or back to back from
    `submitters` threads when rate is None.  Returns per-request latencies (ms), the
    span from the first submission to the last completion (s), the outputs and the
    mean batch size."""
    import queue
    import threading

    rng = np.random.default_rng(seed)
    choice = rng.integers(0, len(pids), n)
    pat = rng.integers(0, pics.shape[0], n)
    gaps = rng.exponential(1.0 / rate, n) if rate else np.zeros(n)
    arrival = np.cumsum(gaps)
    rids, outs, infos = [None] * n, [None] * n, [None] * n
    done_q = queue.Queue()

    def waiter():
        for c in range(n):
            i = done_q.get()
            infos[i] = srv.wait(rids[i])
            if c % 2000 == 1999:
                print(f"# {c + 1}/{n} requests done", file=sys.stderr, flush=True)

    def submitter(idx, t0):
        for i in idx:
            if rate:
                while True:
                    d = t0 + arrival[i] - time.perf_counter()
                    if d <= 0:
                        break
                    time.sleep(d if d > 2e-4 else 0)
            rid, offs, _ = srv.submit(pids[choice[i]], pics[pat[i]])
            rids[i], outs[i] = rid, (choice[i], pat[i], offs)
            done_q.put(i)

    th = threading.Thread(target=waiter)
    th.start()
    t0 = time.perf_counter()
    subs = [threading.Thread(target=submitter, args=(range(k, n, submitters), t0)) for k in range(submitters)]
    for t in subs:
        t.start()
    for t in subs:
        t.join()
    th.join()
    lat = np.array([inf["latency_ms"] for inf in infos])
    t_first = min(inf["submit_ns"] for inf in infos)
    t_last = max(inf["done_ns"] for inf in infos)
    return dict(lat=lat, span_s=(t_last - t_first) * 1e-9, outs=outs,
                batch_mean=float(np.mean([inf["batch"] for inf in infos])))


def serve_plans(cfg, plain_fp64=False):
    """(radius, k, T, policy, mfma mode) of every plan of the serving workload."""
    import emida_amd

    out = []
    for r in cfg["radii"]:
        for k in cfg["grids"]:
            if plain_fp64:
                out.append((r, k, np.float64, emida_amd.CROSS_POLICY_BRUTE, 0))
            elif r == min(cfg["radii"]) or 2 * r > emida_amd.fft_max_slice(np.float32):
                out.append((r, k, np.float32, emida_amd.CROSS_POLICY_BRUTE_MFMA, emida_amd.MFMA_AUTO))
            else:
                out.append((r, k, np.float32, emida_amd.CROSS_POLICY_FFT, 0))
    return out


def serve_main(args, cfg):
    """Serving workload: a request loop over resident plans (one per ROI radius and
    grid), fed by a Poisson arrival process; reports p50/p99 per-pattern latency at
    the offered load, then the saturation throughput of the same request mix, and the
    same mix served by the plain FP64 brute-force form."""
    import threading

    import emida_amd
    from emida_amd import synth

    t_gen = time.perf_counter()
    ps = synth.PatternSet(args.workload, cfg["size"], args.unique, t_range=cfg["t_range"])
    t_gen = time.perf_counter() - t_gen
    dev_name, cus = emida_amd.device_query()
    print(f"# {dev_name}, {cus} CUs; generated {args.unique} distinct patterns in {t_gen:.1f}s", file=sys.stderr,
          flush=True)
    fit = cfg["fit"]

    def make_server(plain_fp64):
        srv = emida_amd.Server(hbm_budget=cfg["hbm_budget"], max_batch=args.serve_batch, streams=4,
                               max_queued=4096, slots_per_plan=2)
        plans = []
        for r, k, T, pol, mode in serve_plans(cfg, plain_fp64):
            begins = synth.grid_begins(cfg["size"], k, r)
            pid = srv.add_plan(T, ps.ref, begins, (2 * r, 2 * r), (4 * r - 1, 4 * r - 1), pol, fit, mfma_mode=mode)
            plans.append(dict(pid=pid, r=r, k=k, T=T, pol=pol, begins=begins))
        return srv, plans

    srv, plans = make_server(False)
    pids = [p["pid"] for p in plans]
    # warm-up: every plan captures its graphs for a few batch sizes
    run_load(srv, pids, ps.pics, 40 * len(plans), None, 1, submitters=4)
    n_load = int(cfg["rate"] * args.duration)
    print(f"# offered load: {n_load} requests at {cfg['rate']:.0f}/s", file=sys.stderr, flush=True)
    load = run_load(srv, pids, ps.pics, n_load, cfg["rate"], 2)
    n_sat = args.saturation_requests
    sat = run_load(srv, pids, ps.pics, n_sat, None, 3, submitters=4)
    st = srv.stats()

    # accuracy against the FP64 oracle (the plan's policy) on one pattern per plan,
    # every 8th ROI, over the well-conditioned ROIs
    from oracle import oracle as O
    O.set_threads(cpu_threads())
    errs = {}
    for p in plans:
        rid, offs, coefs = srv.submit(p["pid"], ps.pics[0])
        srv.wait(rid)
        sub = p["begins"][::8]
        pol = O.POLICY_FFT if p["pol"] == emida_amd.CROSS_POLICY_FFT else O.POLICY_BRUTE
        eo, ec, _, _, est = O.get_offset(ps.pics[0], ps.ref, ps.size, sub, (2 * p["r"], 2 * p["r"]),
                                         (4 * p["r"] - 1, 4 * p["r"] - 1), 1, fit, pol, np.float64,
                                         return_intermediates=True, return_stats=True)
        ok = O.well_conditioned(ec, fit, est)
        e = np.abs(offs[::8] - eo).max(axis=1)
        errs[f"r{p['r']}_k{p['k']}"] = float(e[ok].max()) if ok.any() else None
    srv.close()

    # the same request mix through the plain FP64 brute-force form (saturation)
    srv64, plans64 = make_server(True)
    pids64 = [p["pid"] for p in plans64]
    run_load(srv64, pids64, ps.pics, len(plans64), None, 4)
    sat64 = run_load(srv64, pids64, ps.pics, args.fp64_requests, None, 3, submitters=4)
    srv64.close()

    sat_rate = n_sat / sat["span_s"]
    sat64_rate = args.fp64_requests / sat64["span_s"]
    out = {"metric": METRIC, "value": round(sat_rate, 2), "unit": "patterns/s", "n_gpus": 1,
           "steps": n_sat, "warmup": 40 * len(plans), "ms_per_step": round(1e3 / sat_rate, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 FFT (r32, r64) + f16 MFMA brute screen with exact f32 peak (r16), f32 accumulate", "data": "synthetic",
           "config": {"workload": args.workload, "pattern": f"{cfg['size']}x{cfg['size']} uint16",
                      "radii": list(cfg["radii"]), "roi_grids": [k * k for k in cfg["grids"]], "fitsize": fit,
                      "plans": [f"r{p['r']}/{p['k'] * p['k']} ROIs/" +
                                ("fft" if p["pol"] == emida_amd.CROSS_POLICY_FFT else "mfma-f16-refined")
                                for p in plans],
                      "offered_rate": cfg["rate"], "offered_seconds": args.duration,
                      "hbm_budget": cfg["hbm_budget"], "max_batch": args.serve_batch, "streams": 4,
                      "slots_per_plan": 2,
                      "distinct_patterns": args.unique, "device": dev_name},
           "offered_load": {"requests": n_load, "achieved_patterns_per_s": round(n_load / load["span_s"], 1),
                            "p50_pattern_latency_ms": round(float(np.percentile(load["lat"], 50)), 4),
                            "p99_pattern_latency_ms": round(float(np.percentile(load["lat"], 99)), 4),
                            "max_pattern_latency_ms": round(float(load["lat"].max()), 4),
                            "mean_batch": round(load["batch_mean"], 2)},
           "saturation": {"requests": n_sat, "patterns_per_s": round(sat_rate, 1),
                          "p50_pattern_latency_ms": round(float(np.percentile(sat["lat"], 50)), 4),
                          "mean_batch": round(sat["batch_mean"], 2)},
           "plain_fp64_brute_served_patterns_per_s": round(sat64_rate, 2),
           "served_over_plain_fp64": round(sat_rate / sat64_rate, 2),
           "max_err_px_vs_fp64_per_plan": errs, "accuracy_definition": ACCURACY_DEFINITION,
           "server_stats": {k: int(v) for k, v in st.items()}}
    print(json.dumps(out))
    return 0


def scan_main(args, cfg):
    """Scan workload: patterns streamed back to back through the request loop, once
    with the FFT policy in float and once with split-bf16 matrix-core brute
    correlation; offsets of a subset of distinct raster points (cfg["subset"], 500)
    compared with FP64 brute force on the GPU (the FP64 brute kernels reproduce the
    CPU oracle bit for bit) over the well-conditioned ROIs."""
    import queue
    import threading

    import emida_amd
    from emida_amd import synth

    t_gen = time.perf_counter()
    sc = synth.ScanSet(args.workload, cfg["size"], args.unique, t_range=cfg["t_range"])
    # the error subset: distinct raster points of their own (rendered before the GPU is
    # touched, in a process pool), apart from the points the timed stream cycles over
    sub = cfg["subset"]
    if sub > args.unique:
        more, _ = sc.render(range(args.unique, sub), workers=cpu_threads())
        sub_pics = np.concatenate([sc.pics, more])
    else:
        sub_pics = sc.pics[:sub]
    t_gen = time.perf_counter() - t_gen
    dev_name, cus = emida_amd.device_query()
    print(f"# {dev_name}, {cus} CUs; rendered {max(sub, args.unique)} raster points in {t_gen:.1f}s",
          file=sys.stderr, flush=True)
    r, fit = cfg["radius"], cfg["fit"]
    s2, cross = (2 * r, 2 * r), (4 * r - 1, 4 * r - 1)
    begins = synth.grid_begins(cfg["size"], cfg["k"], r)
    n = args.scan_patterns or cfg["patterns"]

    def stream(policy, mode, T=np.float32):
        srv = emida_amd.Server(hbm_budget=cfg["hbm_budget"], max_batch=args.serve_batch, streams=4,
                               max_queued=512, slots_per_plan=4)
        pid = srv.add_plan(T, sc.ref, begins, s2, cross, policy, fit, mfma_mode=mode)
        outs = {}
        for i in range(min(64, n)):  # warm-up: graphs of every batch size
            rid, o, _ = srv.submit(pid, sc.pics[i % args.unique])
            srv.wait(rid)
        done_q, infos = queue.Queue(), []

        def waiter():
            for _ in range(n):
                rid, i, o = done_q.get()
                infos.append(srv.wait(rid))
                if i < args.unique:
                    outs[i] = o.copy()
                if len(infos) % 5000 == 0:
                    print(f"# {len(infos)}/{n} patterns done", file=sys.stderr, flush=True)

        def submitter(k, parts):
            for i in range(k, n, parts):
                rid, o, _ = srv.submit(pid, sc.pics[i % args.unique])
                done_q.put((rid, i, o))

        th = threading.Thread(target=waiter)
        th.start()
        subs = [threading.Thread(target=submitter, args=(k, 4)) for k in range(4)]
        for t in subs:
            t.start()
        for t in subs:
            t.join()
        th.join()
        span = (max(i["done_ns"] for i in infos) - min(i["submit_ns"] for i in infos)) * 1e-9
        lat = np.array([i["latency_ms"] for i in infos])
        # the error subset through the same server (not timed)
        reqs = [srv.submit(pid, sub_pics[i]) for i in range(sub)]
        sub_out = []
        for rid, o, _ in reqs:
            srv.wait(rid)
            sub_out.append(o.copy())
        srv.close()
        return n / span, lat, sub_out

    fft_rate, fft_lat, fft_out = stream(emida_amd.CROSS_POLICY_FFT, 0)
    # matrix-core screen with the exact peak in double: the FP64 brute results
    mfma_rate, mfma_lat, mfma_out = stream(emida_amd.CROSS_POLICY_BRUTE_MFMA, emida_amd.MFMA_BF16_SPLIT, np.float64)
    f16_rate, f16_lat, f16_out = stream(emida_amd.CROSS_POLICY_BRUTE_MFMA, emida_amd.MFMA_AUTO, np.float64)

    # FP64 reference offsets of the subset, brute force on the GPU, and the FP64
    # correlation's peak statistics (the well-conditioned criterion)
    with emida_amd.GpuOffset(np.float64, np.uint16, sc.size, begins, s2, cross, 16, emida_amd.CROSS_POLICY_BRUTE,
                             fit) as g:
        g.allocate_memory(sc.ref)
        ref_offs, ref_coefs = [], []
        for b0 in range(0, sub, 16):
            pics = np.zeros((16,) + sc.pics.shape[1:], np.uint16)
            m = min(16, sub - b0)
            pics[:m] = sub_pics[b0:b0 + m]
            o, c = g.get_offset(pics)
            ref_offs.append(o.reshape(16, -1, 2)[:m])
            ref_coefs.append(c.reshape(16, -1, 6)[:m])
    ref_offs = np.concatenate(ref_offs)
    ref_coefs = np.concatenate(ref_coefs)
    from oracle import oracle as O
    stats = fp64_fft_peak_stats(sub_pics, sc.ref, sc.size, begins, s2, cpu_threads())
    ok = O.well_conditioned(ref_coefs.reshape(-1, 6), fit, stats).reshape(sub, -1)

    def err(outs):
        e = np.stack([np.abs(outs[i] - ref_offs[i]).max(axis=1) for i in range(sub)])
        return {"max_px": float(e[ok].max()), "mean_px": float(e[ok].mean()),
                "max_px_all_rois": float(np.nanmax(e))}

    out = {"metric": METRIC, "value": round(fft_rate, 2), "unit": "patterns/s", "n_gpus": 1, "steps": n,
           "warmup": min(64, n), "ms_per_step": round(1e3 / fft_rate, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32 FFT; bf16 split / f16 MFMA brute screen (f32 accumulate) + exact f64 peak",
           "data": "synthetic",
           "config": {"workload": args.workload, "pattern": f"{cfg['size']}x{cfg['size']} uint16",
                      "patterns": n, "rois": len(begins), "roi": f"{2 * r}x{2 * r}", "fitsize": fit,
                      "raster": "100 x 200 triangle, eps N(0, 2e-3) with 10-point correlation length",
                      "distinct_patterns": args.unique, "fp64_subset": sub, "max_batch": args.serve_batch,
                      "device": dev_name},
           "fft_f32": {"patterns_per_s": round(fft_rate, 1), "roi_offsets_per_s": round(fft_rate * len(begins)),
                       "p50_pattern_latency_ms": round(float(np.percentile(fft_lat, 50)), 3),
                       "err_vs_fp64": err(fft_out)},
           "mfma_bf16_split": {"patterns_per_s": round(mfma_rate, 1),
                               "roi_offsets_per_s": round(mfma_rate * len(begins)),
                               "p50_pattern_latency_ms": round(float(np.percentile(mfma_lat, 50)), 3),
                               "err_vs_fp64": err(mfma_out)},
           "mfma_f16": {"patterns_per_s": round(f16_rate, 1),
                              "roi_offsets_per_s": round(f16_rate * len(begins)),
                              "p50_pattern_latency_ms": round(float(np.percentile(f16_lat, 50)), 3),
                              "err_vs_fp64": err(f16_out)},
           "accuracy_definition": ACCURACY_DEFINITION, "fp64_subset_patterns": sub,
           "well_conditioned_roi_fraction": round(float(ok.mean()), 4),
           "pcie_floor_patterns_per_s": round(50e9 / (cfg["size"] ** 2 * 2), 1)}
    print(json.dumps(out))
    return 0


def files_main(args, cfg):
    """File-driver workload: TIFF files on the host (written first, so they are read
    from the page cache) through emida_process_files with 1..16 loader threads:
    read, upload, FFT pipeline, results back to the host, all timed."""
    import shutil
    import tempfile

    import emida_amd
    from emida_amd import files as F
    from emida_amd import synth

    t_gen = time.perf_counter()
    ps = synth.PatternSet(args.workload, cfg["size"], args.unique, t_range=cfg["t_range"])
    r, fit, batch = cfg["radius"], cfg["fit"], cfg["batch"]
    slice_size, cross = (2 * r, 2 * r), (4 * r - 1, 4 * r - 1)
    begins = synth.grid_begins(cfg["size"], cfg["k"], r)
    n_roi, n = len(begins), cfg["patterns"]
    d = tempfile.mkdtemp(prefix="emida_files_")
    try:
        paths = []
        for i in range(n):
            p = os.path.join(d, f"DEFORMED_{i:05d}.tif")
            F.write_tiff(p, ps.pics[i % args.unique], rows_per_strip=8)
            paths.append(p)
        t_gen = time.perf_counter() - t_gen
        file_bytes = os.path.getsize(paths[0])
        dev_name, cus = emida_amd.device_query()
        print(f"# {dev_name}, {cus} CUs; wrote {n} files ({file_bytes} bytes each) in {t_gen:.1f}s",
              file=sys.stderr, flush=True)
        sweep, best, ref_offs = [], None, None
        for w in cfg["workers"]:
            g = emida_amd.GpuOffset(np.float32, np.uint16, ps.size, begins, slice_size, cross, batch,
                                    emida_amd.CROSS_POLICY_FFT, fit, in_buffers_size=2 * w)
            g.allocate_memory(ps.ref)
            F.process_files(g, paths[:2 * batch * w], w)  # warm-up
            g.set_timing(True)
            t0 = time.perf_counter()
            res = F.process_files(g, paths, w)
            el = time.perf_counter() - t0
            g.set_timing(False)
            kname, k_ms, k_n = g.get_timing()
            g.close()
            if ref_offs is None:
                ref_offs = res.offsets
            same = bool(np.array_equal(res.offsets, ref_offs, equal_nan=True))
            row = {"loaders": w, "patterns_per_s": round(n / el, 1), "file_MBps": round(n * file_bytes / el / 1e6, 1),
                   "load_ms_per_file": round(res.stats["load_ms"] / n, 4),
                   "failed_batches": res.stats["failed_batches"], "same_results": same}
            print("# " + json.dumps(row), file=sys.stderr, flush=True)
            sweep.append(row)
            if best is None or n / el > best[0]:
                best = (n / el, el, kname, k_ms, k_n, w)
    finally:
        shutil.rmtree(d, ignore_errors=True)

    pats_per_s, el, kname, k_ms, k_n, w_best = best
    k_avg_ms = k_ms / max(k_n, 1)
    flops = fft_flops_per_roi(4 * r) * n_roi * batch
    achieved = flops / (k_avg_ms * 1e-3) / 1e12
    out = {"metric": METRIC, "value": round(pats_per_s, 2), "unit": "patterns/s", "n_gpus": 1,
           "steps": n // batch, "warmup": 2 * w_best, "ms_per_step": round(1e3 * el / (n // batch), 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic TIFF files (page cache)",
           "config": {"workload": args.workload, "pattern": f"{cfg['size']}x{cfg['size']} uint16 TIFF",
                      "patterns": n, "rois": n_roi, "roi": f"{2 * r}x{2 * r}", "policy": "fft", "fitsize": fit,
                      "batch": batch, "loaders": w_best, "device": dev_name},
           "roi_offsets_per_s": round(pats_per_s * n_roi, 1),
           "loader_sweep": sweep,
           "roofline": {"bound": "mfma", "pipe": "FP32 vector", "kernel": kname, "achieved": round(achieved, 3),
                        "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                        "traffic": None, "flops_per_launch": flops, "avg_launch_ms": round(k_avg_ms, 4),
                        "launches": k_n}}
    if not args.no_cpu_baseline:
        threads = min(16, len(os.sched_getaffinity(0)))
        done, t_used, _, _ = cpu_baseline(dict(cfg, policy="fft"), ps, begins, slice_size, cross,
                                          args.cpu_seconds, threads)
        out["cpu_baseline"] = {"value": round(done / t_used, 4), "unit": "patterns/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{done} pattern(s) x {n_roi} ROIs through the CPU oracle (direct "
                                         f"correlation, FP64, OpenMP; file reading not included), {t_used:.1f}s"}
    print(json.dumps(out))
    return 0


def post_main(args, cfg):
    """Post-step workload: per pattern the algebraic homography fit, ROI selection,
    median quality and polar decomposition (emida_fit_deformation) over ROI offsets
    and coefficients already in HBM; checked against the numpy post-processing on a
    subset."""
    import emida_amd
    from emida_amd import _capi as C
    from emida_amd import synth
    from oracle import post as P

    rng = np.random.default_rng(0)
    n, r = cfg["patterns"], cfg["radius"]
    mids = np.array([(x + r, y + r) for x, y in synth.grid_begins(cfg["size"], cfg["k"], r)], float)
    N = len(mids)
    calib = np.array([cfg["size"] // 2] * 3, float)
    base = 64
    offs = np.empty((base, N, 2))
    coefs = rng.normal(0, 1, (base, N, 6))
    for i in range(base):
        Ft = np.eye(3) + rng.normal(0, 2e-3, (3, 3))
        offs[i] = P.apply_transform(Ft, mids, calib) + rng.normal(0, 0.02, mids.shape) - mids
    coefs[:, :, 3] = -rng.uniform(0.5, 2, (base, N))
    coefs[:, :, 5] = -rng.uniform(-0.5, 2, (base, N))
    reps = -(-n // base)
    d_mids = emida_amd.DeviceBuffer.from_array(mids)
    d_offs = emida_amd.DeviceBuffer.from_array(np.tile(offs, (reps, 1, 1))[:n])
    d_coefs = emida_amd.DeviceBuffer.from_array(np.tile(coefs, (reps, 1, 1))[:n])
    d_mask = emida_amd.DeviceBuffer(4 * n)
    d_q = emida_amd.DeviceBuffer(8 * n)
    outs = [emida_amd.DeviceBuffer(72 * n) for _ in range(4)]
    cal = C.Calib(*calib)

    def step():
        C.check(C.lib().emida_fit_deformation(d_mids.ptr, 0, N, d_offs.ptr, d_coefs.ptr, n, cal, 1e-6, 10,
                                              d_mask.ptr, d_q.ptr, outs[0].ptr, outs[1].ptr, outs[2].ptr,
                                              outs[3].ptr, None))

    for _ in range(args.warmup):
        step()
    emida_amd.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    emida_amd.synchronize()
    el = (time.perf_counter() - t0) / args.steps
    # check a subset against the numpy post-processing
    sub = 64
    data = [np.column_stack([mids, offs[i], coefs[i]]) for i in range(sub)]
    mask, quality, H, F, eps, omega = P.post_process(data, calib)
    got_mask = d_mask.download(np.int32, sub).astype(bool)
    got_eps = outs[2].download(np.float64, 9 * sub).reshape(sub, 3, 3)
    err = float(np.nanmax(np.abs(got_eps[mask] - eps[mask]))) if mask.any() else 0.0
    t1 = time.perf_counter()
    for i in range(sub):
        P.post_process([data[i]], calib)
    cpu_rate = sub / (time.perf_counter() - t1)
    in_bytes = n * N * (16 + 48)
    achieved = in_bytes / el / 1e9
    out = {"metric": METRIC, "value": round(n / el, 1), "unit": "patterns/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(1e3 * el, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic ROI offsets and coefficients",
           "config": {"workload": args.workload, "patterns": n, "rois": N, "calib": calib.tolist()},
           "roofline": {"bound": "hbm", "kernel": "deform_fit_kernel", "achieved": round(achieved, 2),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": None, "algorithmic_bytes_per_launch": in_bytes},
           "mask_equal_to_numpy": bool(np.array_equal(got_mask, mask)),
           "max_abs_strain_diff_vs_numpy": err,
           "cpu_baseline": {"value": round(cpu_rate, 2), "unit": "patterns/s", "cores": 1, "kind": "port",
                            "sample": f"{sub} patterns x {N} ROIs through the numpy post-processing "
                                      "(jove/analyze.py restated, SVD per pattern)"}}
    print(json.dumps(out))
    return 0


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="fft_f32_873")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--unique", type=int, default=64, help="distinct synthetic patterns (tiled to the job size)")
    ap.add_argument("--copy-calibration-mb", type=float, default=0, help="only run a device copy of this size")
    ap.add_argument("--mfma", choices=sorted(MFMA_MODES), default=None,
                    help="brute workloads: matrix-core correlation in this mode instead of the plain form")
    ap.add_argument("--inflight", type=int, default=1,
                    help="batches in flight for the batch workloads: one offset object (own stream and "
                         "scratch) per batch, used in turn")
    ap.add_argument("--no-refine", action="store_true",
                    help="with --mfma: use the float screen values directly (T = float) instead of the exact peak")
    ap.add_argument("--duration", type=float, default=None, help="serving: seconds of offered load")
    ap.add_argument("--serve-batch", type=int, default=32, help="serving: patterns per launch at most")
    ap.add_argument("--saturation-requests", type=int, default=6000, help="serving: back-to-back requests")
    ap.add_argument("--fp64-requests", type=int, default=90, help="serving: requests through the FP64 form")
    ap.add_argument("--scan-patterns", type=int, default=0, help="scan: patterns streamed (0: the workload's)")
    ap.add_argument("--no-extras", action="store_true",
                    help="default workload: skip the served-form and precision-form blocks")
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="served block: offered requests per second")
    ap.add_argument("--serve-seconds", type=float, default=3.0, help="served block: seconds of offered load")
    ap.add_argument("--serve-saturation", type=int, default=3000, help="served block: back-to-back requests")
    ap.add_argument("--serve-fp64-requests", type=int, default=256, help="served block: requests to the FP64 form")
    ap.add_argument("--precision-steps", type=int, default=6, help="precision block: timed batches of 128")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported "
              "(patterns would shard across GPUs with no exchange between them)", file=sys.stderr)
        return 2
    if args.copy_calibration_mb:
        copy_calibration(args.copy_calibration_mb)
        return 0

    import emida_amd
    from emida_amd import synth

    cfg = synth.CONFIGS[args.workload]
    if cfg["policy"] == "post":
        return post_main(args, cfg)
    if cfg["policy"] == "files":
        args.unique = min(args.unique, 32)
        return files_main(args, cfg)
    if cfg["policy"] == "scan":
        args.unique = min(args.unique, 64)
        return scan_main(args, cfg)
    if cfg["policy"] == "serve":
        if args.duration is None:
            args.duration = cfg["duration"]
        args.unique = min(args.unique, 32)
        return serve_main(args, cfg)
    r = cfg["radius"]
    slice_size = (2 * r, 2 * r)
    cross = (4 * r - 1, 4 * r - 1)
    T = np.float32 if cfg["dtype"] == "f32" else np.float64
    policy = emida_amd.CROSS_POLICY_FFT if cfg["policy"] == "fft" else emida_amd.CROSS_POLICY_BRUTE
    mfma_code, mfma_products = MFMA_MODES[args.mfma] if args.mfma else (0, 1)
    if args.mfma:
        if cfg["policy"] != "brute":
            print("bench.py: --mfma applies to brute workloads only", file=sys.stderr)
            return 2
        policy = emida_amd.CROSS_POLICY_BRUTE_MFMA
        if args.no_refine:
            T = np.float32
    batch, n_pat = cfg["batch"], cfg["patterns"]
    n_batches = n_pat // batch

    t_gen = time.perf_counter()
    ps = synth.PatternSet(args.workload, cfg["size"], n_pat, t_range=cfg["t_range"], unique=args.unique)
    begins = synth.grid_begins(cfg["size"], cfg["k"], r)
    n_roi = len(begins)
    t_gen = time.perf_counter() - t_gen
    dev_name, cus = emida_amd.device_query()
    print(f"# {dev_name}, {cus} CUs; generated {args.unique} distinct patterns in {t_gen:.1f}s", file=sys.stderr,
          flush=True)

    g = emida_amd.GpuOffset(T, np.uint16, ps.size, begins, slice_size, cross, batch, policy, cfg["fit"],
                            mfma_mode=mfma_code, mfma_refine=not args.no_refine)
    g.allocate_memory(ps.ref)
    d_pics = emida_amd.DeviceBuffer.from_array(ps.pics)  # the whole job, resident in HBM
    pat_bytes = ps.size[0] * ps.size[1] * 2
    total = g.total_slices
    d_maxes = emida_amd.DeviceBuffer(total * g.data_index.itemsize)
    d_neigh = emida_amd.DeviceBuffer(g.neigh_size * np.dtype(T).itemsize)
    d_offs = emida_amd.DeviceBuffer(total * 16)
    d_coefs = emida_amd.DeviceBuffer(total * 48)

    # further batches in flight: their own offset objects (streams, scratch, results)
    gs, outs = [g], [(d_maxes, d_neigh, d_offs, d_coefs)]
    for _ in range(1, max(1, args.inflight)):
        g2 = emida_amd.GpuOffset(T, np.uint16, ps.size, begins, slice_size, cross, batch, policy, cfg["fit"],
                                 mfma_mode=mfma_code, mfma_refine=not args.no_refine)
        g2.allocate_memory(ps.ref)
        gs.append(g2)
        outs.append((emida_amd.DeviceBuffer(total * g.data_index.itemsize),
                     emida_amd.DeviceBuffer(g.neigh_size * np.dtype(T).itemsize),
                     emida_amd.DeviceBuffer(total * 16), emida_amd.DeviceBuffer(total * 48)))

    def step(i):
        ptr = d_pics.ptr.value + (i % n_batches) * batch * pat_bytes
        k = i % len(gs)
        gs[k].process_device(ptr, batch, *outs[k])

    for i in range(args.warmup):
        step(i)
    emida_amd.synchronize()
    g.set_timing(True)
    emida_amd.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(i)
    emida_amd.synchronize()
    t1 = time.perf_counter()
    g.set_timing(False)
    kname, k_ms, k_n = g.get_timing()
    elapsed = t1 - t0
    ms_per_step = 1e3 * elapsed / args.steps
    pats_per_s = args.steps * batch / elapsed

    # batch latency (p50/p99): time from submission to completion of one batch
    lat = []
    for i in range(min(50, max(10, args.steps // 4))):
        emida_amd.synchronize()
        s0 = time.perf_counter()
        step(i)
        emida_amd.synchronize()
        lat.append(1e3 * (time.perf_counter() - s0))
    lat = np.array(lat)

    # GPU offsets of the first batch for the error check
    step(0)
    emida_amd.synchronize()
    gpu_offs = d_offs.download(np.float64, 2 * total).reshape(-1, 2)

    # roofline of the dominant kernel
    k_avg_ms = k_ms / max(k_n, 1)
    if cfg["policy"] == "fft":
        flops = fft_flops_per_roi(4 * r) * total
        peak, pipe = FP32_PEAK_TFLOPS if T == np.float32 else FP64_PEAK_TFLOPS, "FP32 vector" if T == np.float32 else "FP64 vector"
    elif args.mfma:
        # matrix-core work: every product of the split form is issued (3 for hi+lo)
        flops = brute_flops_per_roi(slice_size, cross) * total * mfma_products
        peak, pipe = MFMA_16BIT_PEAK_TFLOPS, f"{args.mfma.split('_')[0]} MFMA, {mfma_products} product(s)"
    else:
        flops = brute_flops_per_roi(slice_size, cross) * total
        peak, pipe = ((FP32_PEAK_TFLOPS, "FP32 vector FMA") if T == np.float32 else
                      (FP64_PEAK_TFLOPS, "FP64 vector FMA"))
    achieved = flops / (k_avg_ms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "pipe": pipe, "kernel": kname, "achieved": round(achieved, 3), "peak": peak,
                "unit": "TFLOP/s", "frac": round(achieved / peak, 4), "traffic": read_traffic(args.workload, args.mfma),
                "flops_per_launch": flops, "avg_launch_ms": round(k_avg_ms, 4), "launches": k_n,
                "algorithmic_bytes_per_launch": batch * pat_bytes + total * (g.data_index.itemsize + 16 + 48)
                + g.neigh_size * np.dtype(T).itemsize}

    out = {"metric": METRIC, "value": round(pats_per_s, 2), "unit": "patterns/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": ((f"{args.mfma} screen (f32 accumulate)" +
                      ("" if args.no_refine else f", exact peak in {'f32' if T == np.float32 else 'f64'}"))
                     if args.mfma else "f32" if T == np.float32 else "f64"),
           "data": "synthetic",
           "config": {"workload": args.workload, "pattern": f"{cfg['size']}x{cfg['size']} uint16",
                      "patterns": n_pat, "rois": n_roi, "roi": f"{2 * r}x{2 * r}",
                      "policy": cfg["policy"] + (f"+mfma:{args.mfma}" + ("" if args.no_refine else "+refine")
                                                 if args.mfma else ""),
                      "fitsize": cfg["fit"], "batch": batch, "batches_in_flight": len(gs),
                      "distinct_patterns": args.unique, "device": dev_name},
           "roi_offsets_per_s": round(pats_per_s * n_roi, 1),
           # one batch of `batch` patterns submitted alone and waited for
           "batch_latency_p50_ms": round(float(np.percentile(lat, 50)), 4),
           "batch_latency_p99_ms": round(float(np.percentile(lat, 99)), 4),
           "roofline": roofline}

    gpu_index = d_maxes.download(g.data_index, total)["index"]
    for gi in gs:
        gi.close()
    del d_pics

    extras = args.workload == "fft_f32_873" and not args.no_extras
    if extras:
        # the serving form and the precision form at their own shapes, on this GPU
        out["served"] = served_block(args, cfg, ps, begins, slice_size, cross)
        # per-pattern latency: single-pattern requests to the request server at the offered load
        out["p50_pattern_latency_ms"] = out["served"]["p50_pattern_latency_ms"]
        out["p99_pattern_latency_ms"] = out["served"]["p99_pattern_latency_ms"]
        out["precision_form"] = precision_block(args, ps)

    if not args.no_cpu_baseline:
        threads = cpu_threads()
        done, t_used, cpu_offs, cpu_coefs, cpu_index, stats = cpu_baseline(cfg, ps, begins, slice_size, cross,
                                                                           args.cpu_seconds, threads)
        cpu_rate = done / t_used
        n = done * n_roi
        acc = accuracy(gpu_offs[:n], gpu_index[:n], cpu_offs, cpu_coefs, cpu_index, stats, cfg["fit"])
        one_rate, one_rois, one_t = cpu_single_thread(cfg, ps, begins, slice_size, cross, args.cpu_seconds / 3)
        out["cpu_baseline"] = {"value": round(cpu_rate, 4), "unit": "patterns/s", "cores": threads, "kind": "port",
                               "sample": f"{done} pattern(s) x {n_roi} ROIs of the same workload through the CPU "
                                         f"oracle (the reference's CPU path restated: direct correlation, FP64, "
                                         f"OpenMP over ROIs, -O3 -mavx2 -mfma), {t_used:.1f}s",
                               "single_thread": {"value": round(one_rate, 4), "unit": "patterns/s", "cores": 1,
                                                 "sample": f"{one_rois} ROIs of one pattern, {one_t:.1f}s"},
                               "host": host_info()}
        if cfg["policy"] == "fft":
            fft_rate, fft_done = cpu_fft_baseline(cfg, ps, begins, slice_size, args.cpu_seconds / 3, threads)
            out["cpu_baseline"]["cpu_fft"] = {
                "value": round(fft_rate, 3), "unit": "patterns/s", "cores": threads,
                "label": "CPU FFT form (no counterpart in the reference's CPU path): oracle sums/prepare, scipy "
                         "pocketfft R2C / product / C2R in FP64, oracle arg-max, gather and fit",
                "sample": f"{fft_done} pattern(s) x {n_roi} ROIs"}
            out["gpu_over_cpu_fft"] = round(pats_per_s / fft_rate, 1)
        out["accuracy"] = acc
        out["max_err_px_vs_fp64"] = acc["max_err_px"]
        out["well_conditioned_roi_fraction"] = acc["well_conditioned_roi_fraction"]
        out["integer_peak_match_fraction"] = acc["integer_peak_match_fraction"]
        out["max_err_px_vs_fp64_all_rois"] = acc["max_err_px_all_rois"]
        out["gpu_over_cpu"] = round(pats_per_s / cpu_rate, 1)
        out["gpu_over_cpu_single_thread"] = round(pats_per_s / one_rate, 1)
    print(json.dumps(out))
    return 0


def served_block(args, cfg, ps, begins, slice_size, cross):
    """The serving form at this workload's shape: single-pattern requests to a request
    server holding the plan resident - Poisson arrivals at args.serve_rate for
    args.serve_seconds (per-pattern p50/p99 latency), then back-to-back requests
    (saturation patterns/s) - and the same requests served by the plain FP64 brute form."""
    import emida_amd

    def server(T, policy):
        srv = emida_amd.Server(hbm_budget=16 << 30, max_batch=args.serve_batch, streams=4, max_queued=4096,
                               slots_per_plan=2)
        pid = srv.add_plan(T, ps.ref, begins, slice_size, cross, policy, cfg["fit"])
        return srv, [pid]

    srv, pids = server(np.float32, emida_amd.CROSS_POLICY_FFT)
    with srv:
        run_load(srv, pids, ps.pics, 256, None, 1, submitters=4)  # graphs of every batch size
        n_load = max(1, int(args.serve_rate * args.serve_seconds))
        load = run_load(srv, pids, ps.pics, n_load, args.serve_rate, 2)
        sat = run_load(srv, pids, ps.pics, args.serve_saturation, None, 3, submitters=4)
    srv64, pids64 = server(np.float64, emida_amd.CROSS_POLICY_BRUTE)
    with srv64:
        run_load(srv64, pids64, ps.pics, 8, None, 4, submitters=2)
        sat64 = run_load(srv64, pids64, ps.pics, args.serve_fp64_requests, None, 5, submitters=4)
    sat_rate = args.serve_saturation / sat["span_s"]
    sat64_rate = args.serve_fp64_requests / sat64["span_s"]
    return {"plan": f"{len(begins)} ROIs of {slice_size[0]}x{slice_size[1]}, FFT float, fitsize {cfg['fit']}",
            "offered_rate": args.serve_rate, "offered_requests": n_load,
            "achieved_patterns_per_s": round(n_load / load["span_s"], 1),
            "p50_pattern_latency_ms": round(float(np.percentile(load["lat"], 50)), 4),
            "p99_pattern_latency_ms": round(float(np.percentile(load["lat"], 99)), 4),
            "mean_batch": round(load["batch_mean"], 2),
            "saturation_patterns_per_s": round(sat_rate, 1), "saturation_requests": args.serve_saturation,
            "plain_fp64_brute_served_patterns_per_s": round(sat64_rate, 2),
            "plain_fp64_requests": args.serve_fp64_requests,
            "served_over_plain_fp64": round(sat_rate / sat64_rate, 2)}


def precision_block(args, ps):
    """The precision form at its own shape (873 x 873, 100 ROIs of 128 x 128, the full
    255 x 255 window, fitsize 7, batch 128): matrix-core screen in the default mode (fp16
    operands for uint16 patterns) with the exact double peak, against the plain FP64
    brute form on the same resident patterns; both give the same results (checked here
    on one batch)."""
    import emida_amd
    from emida_amd import synth

    cfg3 = synth.CONFIGS["brute_873_r64"]
    r, batch, fit = cfg3["radius"], cfg3["batch"], cfg3["fit"]
    begins = synth.grid_begins(cfg3["size"], cfg3["k"], r)
    s2, cross = (2 * r, 2 * r), (4 * r - 1, 4 * r - 1)
    pics = np.concatenate([ps.pics] * -(-batch // ps.pics.shape[0]))[:batch]
    d_pics = emida_amd.DeviceBuffer.from_array(pics)
    total = len(begins) * batch
    res = {}
    for name, policy, steps in (("refined_f16", emida_amd.CROSS_POLICY_BRUTE_MFMA, args.precision_steps),
                                ("plain_fp64_brute", emida_amd.CROSS_POLICY_BRUTE, max(1, args.precision_steps // 3))):
        with emida_amd.GpuOffset(np.float64, np.uint16, ps.size, begins, s2, cross, batch, policy, fit) as g:
            g.allocate_memory(ps.ref)
            d_offs = emida_amd.DeviceBuffer(total * 16)
            g.process_device(d_pics, batch, d_offsets=d_offs)  # warm-up
            emida_amd.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                g.process_device(d_pics, batch, d_offsets=d_offs)
            emida_amd.synchronize()
            el = time.perf_counter() - t0
            res[name] = (steps * batch / el, d_offs.download(np.float64, 2 * total))
    same = bool(np.array_equal(res["refined_f16"][1], res["plain_fp64_brute"][1], equal_nan=True))
    return {"shape": f"{cfg3['size']}x{cfg3['size']}, {len(begins)} ROIs of {2 * r}x{2 * r}, window "
                     f"{cross[0]}x{cross[1]}, fitsize {fit}, batch {batch}",
            "refined_f16_patterns_per_s": round(res["refined_f16"][0], 1),
            "plain_fp64_brute_patterns_per_s": round(res["plain_fp64_brute"][0], 2),
            "refined_over_plain_fp64": round(res["refined_f16"][0] / res["plain_fp64_brute"][0], 2),
            "results_equal_fp64_brute": same}


if __name__ == "__main__":
    sys.exit(main())
