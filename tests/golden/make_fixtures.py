"""Writes the reference's own test fixtures as binary arrays under tests/golden/.

Run in a tree that has the reference mounted at /root/reference (read-only). The
inputs are data files that the reference's tests read:

  test/res/data_pic.txt, data_temp.txt   256x256 doubles (first line: 256)
  test/res/cross_corr/64_{A,B,res}.txt   64x64 ints, 127x127 int result
  test/res/hann/{5,128}_{A,res}.txt      window fixtures (numpy.hanning applied)

and copies, byte for byte, the data files of the file driver's tests and example run:

  test/res/INITIAL_x0y0.tif              873x873 uint16 pattern (load_tiff test)
  TEST_RUN/roi-cryst.txt                 slicepos file: radius 48, 89 ROI middles
  TEST_RUN/def.txt                       deformed list: 492 "<x> <y> <file>" lines

Text values are parsed with Python's correctly rounded float(), so the .npy files
hold exactly the doubles the reference's matrix::from_file reads
(src/matrix.hpp:20-32). SHA-256 of every output is written to fixtures.sha256.
"""
import hashlib
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def read_matrix(path, dtype):
    with open(path) as f:
        toks = f.read().split()
    n = int(toks[0])
    vals = toks[1:]
    if dtype == np.int32:
        arr = np.array([int(v) for v in vals], dtype=np.int32)
    else:
        arr = np.array([float(v) for v in vals], dtype=np.float64)
    if arr.size < n * n:
        raise ValueError(f"{path}: {arr.size} values for n={n}")
    return arr[: n * n].reshape(n, n)


def main():
    jobs = [
        ("test/res/data_pic.txt", "data_pic.npy", np.float64),
        ("test/res/data_temp.txt", "data_temp.npy", np.float64),
        ("test/res/cross_corr/64_A.txt", "cross_corr_64_A.npy", np.int32),
        ("test/res/cross_corr/64_B.txt", "cross_corr_64_B.npy", np.int32),
        ("test/res/cross_corr/64_res.txt", "cross_corr_64_res.npy", np.int32),
        ("test/res/hann/5_A.txt", "hann_5_A.npy", np.float64),
        ("test/res/hann/5_res.txt", "hann_5_res.npy", np.float64),
        ("test/res/hann/128_A.txt", "hann_128_A.npy", np.float64),
        ("test/res/hann/128_res.txt", "hann_128_res.npy", np.float64),
    ]
    lines = []
    for src, dst, dt in jobs:
        arr = read_matrix(os.path.join(REF, src), dt)
        path = os.path.join(OUT, dst)
        np.save(path, arr, allow_pickle=False)
        digest = hashlib.sha256(open(path, "rb").read()).hexdigest()
        lines.append(f"{digest}  {dst}  (from {src}, shape {arr.shape}, {arr.dtype})")
        print(lines[-1])
    copies = [
        ("test/res/INITIAL_x0y0.tif", "INITIAL_x0y0.tif"),
        ("TEST_RUN/roi-cryst.txt", "roi-cryst.txt"),
        ("TEST_RUN/def.txt", "def.txt"),
    ]
    for src, dst in copies:
        data = open(os.path.join(REF, src), "rb").read()
        with open(os.path.join(OUT, dst), "wb") as f:
            f.write(data)
        lines.append(f"{hashlib.sha256(data).hexdigest()}  {dst}  (copy of {src}, {len(data)} bytes)")
        print(lines[-1])
    with open(os.path.join(OUT, "fixtures.sha256"), "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
