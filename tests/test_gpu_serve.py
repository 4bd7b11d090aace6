"""Request server (continuous batching over resident plans) on the GPU.

Each request's offsets must be exactly those of the plain pipeline on the same
pattern (GpuOffset with batch 1), whether the request ran alone or mixed with
requests of other plans submitted concurrently from several threads, with and
without hipGraph replays, under a small HBM budget, and when one plan's idle slots
must be freed for another's.
"""
import threading

import numpy as np
import pytest

import emida_amd
from emida_amd import synth

pytestmark = pytest.mark.gpu

FFT, BRUTE, MFMA = emida_amd.CROSS_POLICY_FFT, emida_amd.CROSS_POLICY_BRUTE, emida_amd.CROSS_POLICY_BRUTE_MFMA

# (T, slice side, ROI grid k, policy, fitsize, matrix-core mode)
PLANS = [
    (np.float32, 32, 4, FFT, 5, None),
    (np.float32, 64, 3, FFT, 5, None),
    (np.float32, 32, 5, MFMA, 5, emida_amd.MFMA_BF16_SPLIT),
    (np.float64, 16, 3, BRUTE, 3, None),
    (np.float32, 16, 6, MFMA, 3, emida_amd.MFMA_F16),
]


@pytest.fixture(scope="module")
def workload():
    ps = synth.PatternSet("serve", 256, 12, t_range=2.0)
    plans = []
    for T, side, k, pol, fit, mode in PLANS:
        begins = synth.grid_begins(256, k, side // 2)
        cross = (2 * side - 1, 2 * side - 1)
        expected = []
        with emida_amd.GpuOffset(T, np.uint16, ps.size, begins, (side, side), cross, 1, pol, fit,
                                 mfma_mode=mode if mode is not None else emida_amd.MFMA_BF16_SPLIT) as g:
            g.allocate_memory(ps.ref)
            for i in range(ps.pics.shape[0]):
                expected.append(g.get_offset(ps.pics[i:i + 1]))
        plans.append(dict(T=T, begins=begins, side=side, cross=cross, pol=pol, fit=fit, mode=mode,
                          expected=expected))
    return ps, plans


def _server(ps, plans, **kw):
    srv = emida_amd.Server(**kw)
    ids = []
    for p in plans:
        ids.append(srv.add_plan(p["T"], ps.ref, p["begins"], (p["side"], p["side"]), p["cross"], p["pol"], p["fit"],
                                mfma_mode=p["mode"] if p["mode"] is not None else emida_amd.MFMA_BF16_SPLIT))
    return srv, ids


def _check(plans, results):
    for (pi, i), (offs, coefs) in results.items():
        eo, ec = plans[pi]["expected"][i]
        np.testing.assert_array_equal(offs, eo, err_msg=f"plan {pi} pattern {i}")
        np.testing.assert_array_equal(coefs, ec, err_msg=f"plan {pi} pattern {i}")


def _mix(srv, ids, ps, plans, rounds, threads=4, seed=0):
    rng = np.random.default_rng(seed)
    jobs = [(pi, i) for _ in range(rounds) for pi in range(len(plans)) for i in range(ps.pics.shape[0])]
    order = rng.permutation(len(jobs))
    chunks = [order[t::threads] for t in range(threads)]
    results, infos, lock = {}, [], threading.Lock()

    def worker(idx):
        mine = []
        for j in idx:
            pi, i = jobs[j]
            rid, offs, coefs = srv.submit(ids[pi], ps.pics[i])
            mine.append((rid, pi, i, offs, coefs))
        for rid, pi, i, offs, coefs in mine:
            info = srv.wait(rid)
            with lock:
                results[(pi, i, rid)] = (offs, coefs)
                infos.append(info)

    ts = [threading.Thread(target=worker, args=(c,)) for c in chunks]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert len(results) == len(jobs)
    return {(pi, i): v for (pi, i, _), v in results.items()}, results, infos


def test_alone_equals_plain_pipeline(workload):
    ps, plans = workload
    srv, ids = _server(ps, plans)
    with srv:
        results = {}
        for pi in range(len(plans)):
            for i in range(ps.pics.shape[0]):
                rid, offs, coefs = srv.submit(ids[pi], ps.pics[i])
                info = srv.wait(rid)
                assert info["batch"] == 1 and info["done_ns"] >= info["launch_ns"] >= info["submit_ns"]
                results[(pi, i)] = (offs, coefs)
        _check(plans, results)


@pytest.mark.parametrize("use_graphs", [True, False])
def test_mix_equals_alone(workload, use_graphs):
    ps, plans = workload
    # one batch slot per plan: requests for a plan that arrive while its batch runs queue
    # behind it and launch together
    srv, ids = _server(ps, plans, use_graphs=use_graphs, max_batch=8, streams=4, slots_per_plan=1)
    with srv:
        # alone first, then mixed rounds (graph replays with eager work between): at least
        # two, and up to four until some requests were batched together (whether they are
        # depends on how fast the submitting threads run against the device)
        alone = {}
        for pi in range(len(plans)):
            rid, offs, coefs = srv.submit(ids[pi], ps.pics[0])
            srv.wait(rid)
            alone[(pi, 0)] = (offs, coefs)
        _check(plans, alone)
        batched = False
        for r in range(4):
            latest, allres, infos = _mix(srv, ids, ps, plans, rounds=2, seed=r)
            for (pi, i, _), v in allres.items():
                _check(plans, {(pi, i): v})
            batched = batched or max(inf["batch"] for inf in infos) > 1
            if r >= 1 and batched:
                break
        assert batched  # requests were batched together
        st = srv.stats()
        assert st["completed"] == st["submitted"]
        assert st["launches"] < st["submitted"]
        if use_graphs:
            assert st["graph_launches"] == st["launches"] and st["graphs_captured"] >= len(plans)
        else:
            assert st["graph_launches"] == 0


def test_small_budget_still_completes(workload):
    # 256 x 256 patterns: a 4-pattern slot of these plans takes about 0.6 MiB, so a
    # 4 MiB budget admits a handful of slots and requests wait for free ones
    ps, plans = workload
    budget = 4 << 20
    srv, ids = _server(ps, plans[:2], hbm_budget=budget, max_batch=4, streams=2)
    with srv:
        _, allres, _ = _mix(srv, ids, ps, plans[:2], rounds=2, threads=2)
        for (pi, i, _), v in allres.items():
            _check(plans, {(pi, i): v})
        st = srv.stats()
        assert 0 < st["device_bytes_batches"] <= budget
        assert st["device_bytes_peak_inflight"] <= st["device_bytes_batches"]


def test_idle_slots_are_given_back_to_a_waiting_plan(workload):
    # slots_per_plan 0 lets one plan take the whole budget; a budget that holds one slot
    # of either plan but never both makes every switch of plan free the other plan's
    # idle slot (before, the second plan's request waited forever)
    ps, plans = workload
    two = plans[:2]
    sizes = []
    for p in two:
        with emida_amd.Server(hbm_budget=1 << 30, max_batch=1, streams=1, slots_per_plan=0) as probe:
            pid = probe.add_plan(p["T"], ps.ref, p["begins"], (p["side"], p["side"]), p["cross"], p["pol"], p["fit"])
            probe.wait(probe.submit(pid, ps.pics[0])[0])
            sizes.append(probe.stats()["device_bytes_batches"])
    budget = max(sizes)
    assert sum(sizes) > budget
    srv, ids = _server(ps, two, hbm_budget=budget, max_batch=1, streams=2, slots_per_plan=0)
    with srv:
        results = {}
        for _ in range(2):
            for pi in range(2):
                for i in range(3):
                    rid, offs, coefs = srv.submit(ids[pi], ps.pics[i])
                    srv.wait(rid)
                    results[(pi, i)] = (offs, coefs)
        _check(two, results)
        st = srv.stats()
        assert st["completed"] == st["submitted"] == 12
        assert st["slots"] == 1 and 0 < st["device_bytes_batches"] <= budget


def test_dropped_output_arrays_stay_alive(workload):
    # regression: a caller that drops the returned arrays must not leave the library
    # writing into freed memory; the server holds them until the request is waited for
    ps, plans = workload
    srv, ids = _server(ps, plans[:1])
    with srv:
        rids = [srv.submit(ids[0], ps.pics[i % ps.pics.shape[0]])[0] for i in range(32)]
        import gc
        gc.collect()
        for rid in rids:
            srv.wait(rid)


def test_errors():
    srv = emida_amd.Server(max_batch=2)
    with srv:
        ref = np.zeros((64, 64), np.uint16)
        with pytest.raises(emida_amd.EmidaError):
            srv.add_plan(np.float32, ref, [(60, 0)], (16, 16))  # ROI outside the pattern
        pid = srv.add_plan(np.float32, ref, [(0, 0)], (16, 16), policy=FFT, s=3)
        with pytest.raises(ValueError):
            srv.submit(pid, np.zeros((32, 32), np.uint16))
        with pytest.raises(emida_amd.EmidaError):
            srv.wait(12345)
