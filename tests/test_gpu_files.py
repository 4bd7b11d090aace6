"""The file pipeline and the command-line driver on the GPU.

  * emida_process_files (file_processor, src/process_files.hpp:98-241) returns, for
    every file, exactly what gpu_offset::get_offset returns for the same pixels,
    for each policy, ragged last batches and several loader threads; brute double
    results also equal the oracle's;
  * a file that cannot be read fails its own batch only, and the pipeline ends
    also when the last batch fails (the reference's GPU thread would wait forever);
  * the driver's standard output equals finalize_offsets' format
    (src/process_files.hpp:235-264) of those results, byte for byte;
  * the reference's own pattern (test/res/INITIAL_x0y0.tif) against itself, with
    the ROI file of the reference's example run (TEST_RUN/roi-cryst.txt): every
    offset is zero;
  * --crosspolicy mfma prints exactly the brute policy's output (exact refinement).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import emida_amd
from emida_amd import _capi as C
from emida_amd import files as F
from emida_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
CLI = os.path.join(ROOT, "emida_amd", "emida")
BRUTE, FFT, MFMA = emida_amd.CROSS_POLICY_BRUTE, emida_amd.CROSS_POLICY_FFT, emida_amd.CROSS_POLICY_BRUTE_MFMA

SIZE = 256
N_FILES = 7


@pytest.fixture(scope="module")
def scan(tmp_path_factory):
    d = tmp_path_factory.mktemp("scan")
    ps = synth.PatternSet("files", SIZE, N_FILES, t_range=3.0)
    init = str(d / "INITIAL.tif")
    F.write_tiff(init, ps.ref, rows_per_strip=8)
    paths = []
    for i in range(N_FILES):
        p = str(d / f"DEFORMED_x{60 * i}y0.tif")
        F.write_tiff(p, ps.pics[i], rows_per_strip=[1, 8, 256, 3][i % 4])
        paths.append(p)
    mids = [(48 + 80 * i, 48 + 80 * j) for j in range(3) for i in range(3)]
    return {"dir": d, "ps": ps, "initial": init, "paths": paths, "mids": mids}


def _direct(scan, T, policy, slice_size, cross, batch, fit=3):
    """Per-file results of get_offset over full batches (the last one padded with
    copies of the first pattern)."""
    ps = scan["ps"]
    begins = F.slice_begins(scan["mids"], slice_size)
    n_roi = len(begins)
    offs = np.zeros((N_FILES, n_roi, 2))
    coefs = np.zeros((N_FILES, n_roi, 6))
    with emida_amd.GpuOffset(T, np.uint16, (SIZE, SIZE), begins, slice_size, cross, batch, policy, fit) as g:
        g.allocate_memory(ps.ref)
        for b in range(0, N_FILES, batch):
            idx = [min(i, N_FILES - 1) if i < N_FILES else 0 for i in range(b, b + batch)]
            o, c = g.get_offset(ps.pics[idx])
            k = min(batch, N_FILES - b)
            offs[b:b + k] = o.reshape(batch, n_roi, 2)[:k]
            coefs[b:b + k] = c.reshape(batch, n_roi, 6)[:k]
    return offs, coefs


def _pipeline(scan, T, policy, slice_size, cross, batch, workers, paths=None, ordered=True, fit=3):
    begins = F.slice_begins(scan["mids"], slice_size)
    with emida_amd.GpuOffset(T, np.uint16, (SIZE, SIZE), begins, slice_size, cross, batch, policy, fit,
                             in_buffers_size=2 * workers) as g:
        g.allocate_memory(F.load_tiff(scan["initial"], (SIZE, SIZE)))
        return F.process_files(g, scan["paths"] if paths is None else paths, workers, ordered)


@pytest.mark.parametrize("T,policy,batch,workers", [
    (np.float64, BRUTE, 3, 1),
    (np.float64, BRUTE, 3, 3),
    (np.float32, FFT, 2, 2),
    (np.float32, FFT, 1, 4),
    (np.float64, MFMA, 4, 2),
])
def test_pipeline_equals_get_offset(scan, T, policy, batch, workers):
    print(f"pipeline T={np.dtype(T).name} policy={policy} batch={batch} workers={workers}", flush=True)
    res = _pipeline(scan, T, policy, (32, 32), (63, 63), batch, workers)
    assert (res.status == C.OK).all(), res.errors
    assert res.stats["batches"] == -(-N_FILES // batch) and res.stats["failed_batches"] == 0
    assert res.order == list(range(0, N_FILES, batch))  # ordered delivery
    offs, coefs = _direct(scan, T, policy, (32, 32), (63, 63), batch)
    np.testing.assert_array_equal(res.offsets, offs)
    np.testing.assert_array_equal(res.coefs, coefs)


def test_pipeline_brute_double_equals_oracle(scan):
    ps = scan["ps"]
    begins = F.slice_begins(scan["mids"], (32, 32))
    res = _pipeline(scan, np.float64, BRUTE, (32, 32), (63, 63), 3, 2)
    eo, ec = O.get_offset(ps.pics.ravel(), ps.ref, (SIZE, SIZE), begins, (32, 32), (63, 63), N_FILES, 3, BRUTE)
    np.testing.assert_array_equal(res.offsets.reshape(-1, 2), eo)
    np.testing.assert_array_equal(res.coefs.reshape(-1, 6), ec)


def test_unordered_delivery_has_same_results(scan):
    a = _pipeline(scan, np.float32, FFT, (32, 32), (63, 63), 1, 4, ordered=False)
    b = _pipeline(scan, np.float32, FFT, (32, 32), (63, 63), 1, 1, ordered=True)
    assert sorted(a.order) == list(range(N_FILES))
    np.testing.assert_array_equal(a.offsets, b.offsets)


@pytest.mark.parametrize("bad", [4, N_FILES - 1])
def test_unreadable_file_fails_its_batch_only(scan, bad):
    paths = list(scan["paths"])
    paths[bad] = str(scan["dir"] / "missing.tif")
    res = _pipeline(scan, np.float64, BRUTE, (32, 32), (63, 63), 3, 2, paths=paths)
    first = bad // 3 * 3
    assert res.stats["failed_batches"] == 1
    failed = np.zeros(N_FILES, bool)
    failed[first:first + 3] = True
    assert (res.status[failed] == C.ERR_ARG).all() and (res.status[~failed] == C.OK).all()
    assert "missing.tif" in res.errors[first]
    offs, _ = _direct(scan, np.float64, BRUTE, (32, 32), (63, 63), 3)
    np.testing.assert_array_equal(res.offsets[~failed], offs[~failed])
    assert np.isnan(res.offsets[failed]).all()


def _write_inputs(scan, radius=16):
    d = scan["dir"]
    pos = d / "slicepos.txt"
    pos.write_text(f"{radius}\n" + "".join(f"{x} {y}\n" for x, y in scan["mids"]))
    lst = d / "deformed.txt"
    lst.write_text("".join(f"{60.0 * i} {519.6152422706632 * (i % 2)} {p}\n" for i, p in enumerate(scan["paths"])))
    return str(pos), str(lst)


def _run_cli(*args):
    print("cli", " ".join(args[-8:]), flush=True)
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=100)


@pytest.mark.parametrize("policy,precision,batch,workers,coefs", [
    ("brute", "double", 3, 2, True),
    ("fft", "float", 2, 1, False),
])
def test_cli_stdout_equals_finalize_format(scan, policy, precision, batch, workers, coefs):
    pos, lst = _write_inputs(scan)
    args = ["-i", scan["initial"], "-d", lst, "-b", pos, "--crosspolicy", policy, "--precision", precision,
            "--batchsize", str(batch), "--loadworkers", str(workers)] + (["-q"] if coefs else [])
    r = _run_cli(*args)
    assert r.returncode == 0, r.stdout + r.stderr
    T = np.float64 if precision == "double" else np.float32
    pol = BRUTE if policy == "brute" else FFT
    res = _pipeline(scan, T, pol, (32, 32), (63, 63), batch, workers)
    work = F.read_deformed_list(lst)
    assert r.stdout == F.format_output(work, scan["mids"], res, write_coefs=coefs)
    assert r.stdout.count("\n") == N_FILES * (1 + len(scan["mids"]))


def test_reference_pattern_against_itself_gives_zero_offsets(tmp_path):
    init = os.path.join(GOLD, "INITIAL_x0y0.tif")
    lst = tmp_path / "def.txt"
    lst.write_text(f"0.0 0.0 {init}\n60.0 0.0 {init}\n")
    pos = os.path.join(GOLD, "roi-cryst.txt")
    for policy, precision in [("fft", "float"), ("brute", "double")]:
        r = _run_cli("-i", init, "-d", str(lst), "-b", pos, "--crosspolicy", policy, "--precision", precision,
                     "--batchsize", "2")
        assert r.returncode == 0, r.stdout + r.stderr
        lines = r.stdout.splitlines()
        assert len(lines) == 2 * 90 and lines[0] == f"0.000000 0.000000 89 {init}"
        vals = np.array([[float(v) for v in ln.split()[2:]] for ln in lines[1:90]])
        assert np.abs(vals).max() < 1e-3, (policy, np.abs(vals).max())


@pytest.mark.parametrize("precision", ["double", "float"])
def test_cli_mfma_policy_prints_the_brute_output(scan, precision):
    # the refined matrix-core policy gives the brute policy's results bit for bit, so
    # the driver's output is identical text
    pos, lst = _write_inputs(scan)
    common = ["-i", scan["initial"], "-d", lst, "-b", pos, "--precision", precision, "--batchsize", "3", "-q", "-f",
              "5"]
    a = _run_cli(*common, "--crosspolicy", "brute")
    b = _run_cli(*common, "--crosspolicy", "mfma")
    assert a.returncode == 0 and b.returncode == 0, a.stdout + a.stderr + b.stdout + b.stderr
    assert a.stdout == b.stdout and a.stdout.count("\n") == N_FILES * (1 + len(scan["mids"]))
