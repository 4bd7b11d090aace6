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
  cpu_baseline  the CPU oracle (direct correlation, FP64, OpenMP) on a bounded
                sample of the same workload on this host;
  error         max |GPU - FP64 oracle| offset over the sample's ROIs whose FP64
                fit has its maximum inside the neighbourhood (the others are
                extrapolations of noise; their raw maximum is reported too);
  latency       p50/p99 time for a pattern's batch to complete (all patterns of a
                batch complete together).
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


def read_traffic(workload):
    """HBM bytes per launch of the dominant kernel from profiles/traffic_<workload>.txt
    (FETCH_SIZE and WRITE_SIZE in KiB per dispatch, calibrated with a copy of known
    size).  None when no summary is committed."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    m = re.search(r"^hbm_bytes_per_launch:\s*([0-9.eE+]+)", text, re.M)
    return float(m.group(1)) if m else None


def cpu_baseline(cfg, ps, begins, slice_size, cross, seconds, threads):
    from oracle import oracle as O

    O.set_threads(threads)
    policy = O.POLICY_FFT if cfg["policy"] == "fft" else O.POLICY_BRUTE
    done, t_used, offs, coefs = 0, 0.0, [], []
    while done < ps.pics.shape[0] and (done == 0 or t_used < seconds):
        t0 = time.perf_counter()
        o, c = O.get_offset(ps.pics[done], ps.ref, ps.size, begins, slice_size, cross, 1, cfg["fit"], policy,
                            np.float64)
        t_used += time.perf_counter() - t0
        offs.append(o)
        coefs.append(c)
        done += 1
    return done, t_used, np.concatenate(offs), np.concatenate(coefs)


def well_posed(coefs, fit):
    """ROIs whose FP64 quadratic fit has a maximum inside the F x F neighbourhood.
    Elsewhere (noise-dominated ROIs) the fitted peak is an extrapolation whose
    position is ill-conditioned in any precision."""
    b, c, d, e, f = (coefs[:, i] for i in range(1, 6))
    det = 4 * f * d - e * e
    with np.errstate(divide="ignore", invalid="ignore"):
        py = (b * e - 2 * c * d) / det
        px = (-e * py - b) / (2 * d)
    return (d < 0) & (det > 0) & (px >= 0) & (px <= fit - 1) & (py >= 0) & (py <= fit - 1)


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
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(patterns shard across GPUs with no exchange; see DESIGN.md)", file=sys.stderr)
        return 2
    if args.copy_calibration_mb:
        copy_calibration(args.copy_calibration_mb)
        return 0

    import emida_amd
    from emida_amd import synth

    cfg = synth.CONFIGS[args.workload]
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
        T, policy = np.float32, emida_amd.CROSS_POLICY_BRUTE_MFMA
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
                            mfma_mode=mfma_code)
    g.allocate_memory(ps.ref)
    d_pics = emida_amd.DeviceBuffer.from_array(ps.pics)  # the whole job, resident in HBM
    pat_bytes = ps.size[0] * ps.size[1] * 2
    total = g.total_slices
    d_maxes = emida_amd.DeviceBuffer(total * g.data_index.itemsize)
    d_neigh = emida_amd.DeviceBuffer(g.neigh_size * np.dtype(T).itemsize)
    d_offs = emida_amd.DeviceBuffer(total * 16)
    d_coefs = emida_amd.DeviceBuffer(total * 48)

    def step(i):
        ptr = d_pics.ptr.value + (i % n_batches) * batch * pat_bytes
        g.process_device(ptr, batch, d_maxes, d_neigh, d_offs, d_coefs)

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
                "unit": "TFLOP/s", "frac": round(achieved / peak, 4), "traffic": read_traffic(args.workload),
                "flops_per_launch": flops, "avg_launch_ms": round(k_avg_ms, 4), "launches": k_n,
                "algorithmic_bytes_per_launch": batch * pat_bytes + total * (g.data_index.itemsize + 16 + 48)
                + g.neigh_size * np.dtype(T).itemsize}

    out = {"metric": METRIC, "value": round(pats_per_s, 2), "unit": "patterns/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": (f"{args.mfma} (f32 accumulate)" if args.mfma else "f32" if T == np.float32 else "f64"),
           "data": "synthetic",
           "config": {"workload": args.workload, "pattern": f"{cfg['size']}x{cfg['size']} uint16",
                      "patterns": n_pat, "rois": n_roi, "roi": f"{2 * r}x{2 * r}",
                      "policy": cfg["policy"] + (f"+mfma:{args.mfma}" if args.mfma else ""),
                      "fitsize": cfg["fit"], "batch": batch,
                      "distinct_patterns": args.unique, "device": dev_name},
           "roi_offsets_per_s": round(pats_per_s * n_roi, 1),
           "p50_pattern_latency_ms": round(float(np.percentile(lat, 50)), 4),
           "p99_pattern_latency_ms": round(float(np.percentile(lat, 99)), 4),
           "roofline": roofline}

    if not args.no_cpu_baseline:
        threads = min(16, len(os.sched_getaffinity(0)))
        done, t_used, cpu_offs, cpu_coefs = cpu_baseline(cfg, ps, begins, slice_size, cross, args.cpu_seconds,
                                                         threads)
        cpu_rate = done / t_used
        err = np.abs(gpu_offs[:done * n_roi] - cpu_offs).max(axis=1)
        ok = well_posed(cpu_coefs, cfg["fit"])
        out["cpu_baseline"] = {"value": round(cpu_rate, 4), "unit": "patterns/s", "cores": threads, "kind": "port",
                               "sample": f"{done} pattern(s) x {n_roi} ROIs of the same workload through the CPU "
                                         f"oracle (direct correlation, FP64, OpenMP), {t_used:.1f}s"}
        out["max_err_px_vs_fp64"] = float(err[ok].max())
        out["well_posed_roi_fraction"] = round(float(ok.mean()), 4)
        out["max_err_px_vs_fp64_all_rois"] = float(np.nanmax(err))
        out["gpu_over_cpu"] = round(pats_per_s / cpu_rate, 1)
    print(json.dumps(out))
    g.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
