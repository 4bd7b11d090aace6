"""The serving and scan workloads under parity tests on the GPU.

Serving (873 x 873 uint16 patterns; ROI radius 16, 32 or 64 on 10 x 10, 15 x 15 or
20 x 20 grids; fitsize 5): a request server holds every plan resident - refined
matrix-core brute correlation (fp16 split) for radius 16, the FFT policy in float for
radius 32 and 64 - plus a refined double plan.  Patterns are submitted from several
threads at once.  Every served result must equal the plain pipeline's result for the
same pattern bit for bit (GpuOffset, batch 1), and the plans must agree with the CPU
oracle: refined plans bit for bit (they equal the brute policy, which reproduces the
oracle's operation order), float FFT plans within 2e-3 px with equal integer peaks on
the well-conditioned ROIs (oracle.well_conditioned: FP64 peak margin above float's
rounding, fitted maximum inside the neighbourhood).

Scan (1024 x 1024 patterns, 31 x 31 ROIs of 64 x 64, fitsize 9): one raster pattern
streamed through the request loop with the FFT policy in float and with refined
split-bf16 matrix-core brute correlation in double, checked the same way.

The reference has no server; its batch driver is src/process_files.hpp:136-181 and its
end-to-end test style test/get_offset_test.cpp:108-199.
"""
import threading

import numpy as np
import pytest

import emida_amd
from emida_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FFT, BRUTE, MFMA = emida_amd.CROSS_POLICY_FFT, emida_amd.CROSS_POLICY_BRUTE, emida_amd.CROSS_POLICY_BRUTE_MFMA
FIT = 5


def _oracle_fft_check(pattern, ref, size, begins, side, fit, offs, maxes_index):
    """FP64 oracle (FFT policy) on `begins`: on the well-conditioned ROIs (FP64 peak
    margin above float's rounding, fitted maximum inside the neighbourhood) the integer
    peaks must be equal and the offsets within 2e-3 px.  Returns (well-conditioned
    count, max error there, fraction of all ROIs with the FP64 integer peak)."""
    s2 = (side, side)
    cross = (2 * side - 1, 2 * side - 1)
    eo, ec, em, _, st = O.get_offset(pattern.ravel(), ref, size, begins, s2, cross, 1, fit, O.POLICY_FFT,
                                     np.float64, return_intermediates=True, return_stats=True)
    ok = O.well_conditioned(ec, fit, st)
    np.testing.assert_array_equal(maxes_index[ok], em["index"][ok])
    err = float(np.abs(offs - eo)[ok].max())
    assert err < 2e-3, err
    return int(ok.sum()), err, float((maxes_index == em["index"]).mean())


# ------------------------------------------------------------------ serving
SERVE_SIZE = 873
SERVE_PLANS = [  # (radius, grid k, T, policy, matrix-core mode): the serving workload's plans
    (r, k, np.float32, MFMA if r == 16 else FFT, emida_amd.MFMA_AUTO)
    for r in (16, 32, 64) for k in (10, 15, 20)
] + [(16, 10, np.float64, MFMA, emida_amd.MFMA_F16_SPLIT)]  # and a split-mode double plan


@pytest.fixture(scope="module")
def served():
    ps = synth.PatternSet("serve-c4", SERVE_SIZE, 2, t_range=3.0)
    plans = []
    for r, k, T, pol, mode in SERVE_PLANS:
        begins = synth.grid_begins(SERVE_SIZE, k, r)
        cross = (4 * r - 1, 4 * r - 1)
        plain = []
        with emida_amd.GpuOffset(T, np.uint16, ps.size, begins, (2 * r, 2 * r), cross, 1, pol, FIT,
                                 mfma_mode=mode) as g:
            g.allocate_memory(ps.ref)
            for i in range(2):
                maxes, neigh = g.new_result_buffers(pinned=False)
                idx = g.transfer_pic_to_device_async(ps.pics[i:i + 1])
                g.wait_uploads()
                g.get_offset_core(maxes, neigh, idx)
                offs, coefs = g.finalize(maxes, neigh)
                plain.append((offs, coefs, maxes.copy()))
        plans.append(dict(r=r, k=k, T=T, pol=pol, mode=mode, begins=begins, cross=cross, plain=plain))

    # one batch slot per plan: a plan's requests that arrive while its batch runs queue
    # behind it and share the next launch
    srv = emida_amd.Server(hbm_budget=8 << 30, max_batch=8, streams=4, max_queued=256, slots_per_plan=1)
    with srv:
        for p in plans:
            p["pid"] = srv.add_plan(p["T"], ps.ref, p["begins"], (2 * p["r"], 2 * p["r"]), p["cross"], p["pol"], FIT,
                                    mfma_mode=p["mode"])
        # two submitting threads, each sending both patterns to every plan in its own order
        results, lock = {}, threading.Lock()

        def worker(seed):
            rng = np.random.default_rng(seed)
            jobs = [(pi, i) for pi in range(len(plans)) for i in range(2)]
            mine = []
            for j in rng.permutation(len(jobs)):
                pi, i = jobs[j]
                rid, offs, coefs = srv.submit(plans[pi]["pid"], ps.pics[i])
                mine.append((rid, pi, i, offs, coefs))
            for rid, pi, i, offs, coefs in mine:
                info = srv.wait(rid)
                with lock:
                    results.setdefault((pi, i), []).append((offs, coefs, info))

        ts = [threading.Thread(target=worker, args=(s,)) for s in (1, 2)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        stats = srv.stats()
    return ps, plans, results, stats


def test_served_results_equal_plain_pipeline(served):
    ps, plans, results, stats = served
    assert stats["completed"] == stats["submitted"] == 2 * 2 * len(plans)
    assert max(inf["batch"] for v in results.values() for _, _, inf in v) > 1  # requests shared launches
    for (pi, i), runs in results.items():
        offs0, coefs0, _ = plans[pi]["plain"][i]
        assert len(runs) == 2
        for offs, coefs, _ in runs:
            np.testing.assert_array_equal(offs, offs0, err_msg=f"plan {pi} pattern {i}")
            np.testing.assert_array_equal(coefs, coefs0, err_msg=f"plan {pi} pattern {i}")


@pytest.mark.parametrize("pi", [i for i, p in enumerate(SERVE_PLANS) if p[3] == MFMA])
def test_served_refined_plans_equal_oracle(served, pi):
    # refined matrix-core plans give the brute policy's results, which are the oracle's
    # bit for bit (T = float and double); every ROI, both patterns
    ps, plans, results, _ = served
    p = plans[pi]
    r = p["r"]
    for i in range(2):
        eo, ec = O.get_offset(ps.pics[i].ravel(), ps.ref, ps.size, p["begins"], (2 * r, 2 * r), p["cross"], 1, FIT,
                              O.POLICY_BRUTE, p["T"])
        offs, coefs, _ = results[(pi, i)][0]
        np.testing.assert_array_equal(offs, eo)
        np.testing.assert_array_equal(coefs, ec)


@pytest.mark.parametrize("pi", [i for i, p in enumerate(SERVE_PLANS) if p[3] == FFT])
def test_served_fft_plans_vs_oracle(served, pi):
    # every 8th ROI (the FFT gather never leaves its own slice, so a subset of ROIs
    # gives the same per-ROI results); radius 64 on the first pattern only
    ps, plans, results, _ = served
    p = plans[pi]
    r = p["r"]
    sub = np.arange(0, len(p["begins"]), 8)
    begins = [p["begins"][j] for j in sub]
    for i in (range(1) if r == 64 else range(2)):
        offs, _, _ = results[(pi, i)][0]
        maxes = p["plain"][i][2]
        n_ok, err, match = _oracle_fft_check(ps.pics[i], ps.ref, ps.size, begins, 2 * r, FIT, offs[sub],
                                             maxes["index"][sub])
        assert n_ok >= 0.5 * len(sub), (n_ok, len(sub))
        print(f"r{r} k{p['k']} pattern {i}: {n_ok}/{len(sub)} well-conditioned ROIs, max error {err:.2e} px, "
              f"integer peaks equal on {match:.1%}")


# ------------------------------------------------------------------ scan
@pytest.fixture(scope="module")
def scan():
    sc = synth.ScanSet("scan-c5", 1024, 1, t_range=2.0)
    begins = synth.grid_begins(1024, 31, 32)
    out = {}
    srv = emida_amd.Server(hbm_budget=8 << 30, max_batch=4, streams=2, max_queued=64)
    with srv:
        fft = srv.add_plan(np.float32, sc.ref, begins, (64, 64), (127, 127), FFT, 9)
        mfma = srv.add_plan(np.float64, sc.ref, begins, (64, 64), (127, 127), MFMA, 9,
                            mfma_mode=emida_amd.MFMA_BF16_SPLIT)
        for name, pid in (("fft", fft), ("mfma", mfma)):
            rid, offs, coefs = srv.submit(pid, sc.pics[0])
            srv.wait(rid)
            out[name] = (offs, coefs)
    with emida_amd.GpuOffset(np.float32, np.uint16, sc.size, begins, (64, 64), (127, 127), 1, FFT, 9) as g:
        g.allocate_memory(sc.ref)
        maxes, neigh = g.new_result_buffers(pinned=False)
        idx = g.transfer_pic_to_device_async(sc.pics[:1])
        g.wait_uploads()
        g.get_offset_core(maxes, neigh, idx)
        offs, _ = g.finalize(maxes, neigh)
        np.testing.assert_array_equal(offs, out["fft"][0])  # served == plain
        out["fft_maxes"] = maxes.copy()
    return sc, begins, out


def test_scan_refined_split_bf16_double_equals_oracle(scan):
    sc, begins, out = scan
    eo, ec = O.get_offset(sc.pics[0].ravel(), sc.ref, sc.size, begins, (64, 64), (127, 127), 1, 9, O.POLICY_BRUTE,
                          np.float64)
    np.testing.assert_array_equal(out["mfma"][0], eo)
    np.testing.assert_array_equal(out["mfma"][1], ec)


def test_scan_fft_float_vs_oracle(scan):
    sc, begins, out = scan
    sub = np.arange(0, len(begins), 8)
    n_ok, err, match = _oracle_fft_check(sc.pics[0], sc.ref, sc.size, [begins[j] for j in sub], 64, 9,
                                         out["fft"][0][sub], out["fft_maxes"]["index"][sub])
    assert n_ok >= 0.5 * len(sub)
    print(f"scan FFT float: {n_ok}/{len(sub)} well-conditioned ROIs, max error {err:.2e} px, "
          f"integer peaks equal on {match:.1%}")
