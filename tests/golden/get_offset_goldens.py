"""End-to-end goldens of the reference's get_offset tests, as data
(the 49 batched offsets live in get_offset_batched.txt, shared with the C++ test).

Values are the expected offsets written in test/get_offset_test.cpp:25-26 (single
64x64 ROI at {0,0}, brute, cross 127x127, tolerance 2e-14) and :143-194 (49 ROIs of
64x64 at step 32 on the 256x256 fixture, BRUTE and FFT, batch 1 and 5, tolerance
7e-13). The reference produced them with python/test.py:74-113 (fitsize 3).
"""
BIGGER = (0.07583538046549165, -0.0982055210473689)
BIGGER_TOL = 2e-14
BATCHED_TOL = 7e-13
BATCHED = [tuple(float(v) for v in line.split())
           for line in open(__import__("os").path.join(__import__("os").path.dirname(__file__),
                                                       "get_offset_batched.txt"))
           if line.strip() and not line.startswith("#")]


def slice_begins(src=256, size=64, step=32):
    """get_slice_begins (src/slice_picture.hpp:56-66): row-major grid, x fastest."""
    out = []
    y = 0
    while y + size <= src:
        x = 0
        while x + size <= src:
            out.append((x, y))
            x += step
        y += step
    return out
