"""Synthetic EBSD-like patterns and ROI grids for tests and benchmarks.

Band model: a constant background of 2000 counts plus N Kikuchi-like bands
(direction U[0, pi), signed distance of the band centre line from the image centre
U over the image half-diagonal, Gaussian cross profile sigma U[6, 20] px, amplitude
U[500, 3000]) multiplied by a radial falloff exp(-r^2 / (2 * 200^2)) around the
image centre, then Poisson noise.  A deformed pattern resamples the noise-free
reference bilinearly under x' = (I + eps)(x - c) + c + t (c = image centre), then
draws fresh Poisson noise.  Content at reference position x therefore appears at
x' in the deformed pattern; the local displacement at a ROI centre p is
eps (p - c) + t and the offset the pipeline reports is its negative.
"""
from __future__ import annotations

import hashlib
import math

import numpy as np

try:
    from scipy import ndimage as _ndi
except Exception:  # pragma: no cover - scipy is present in this image
    _ndi = None


def seed_of(*parts) -> int:
    """Deterministic 63-bit seed from any printable parts (config name, index...)."""
    h = hashlib.sha256("/".join(str(p) for p in parts).encode()).digest()
    return int.from_bytes(h[:8], "little") & ((1 << 63) - 1)


def band_reference(size, n_bands=24, seed=0, background=2000.0, falloff=200.0):
    """Noise-free reference pattern (float64, shape (H, W))."""
    w, h = (size, size) if np.isscalar(size) else size
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    cx, cy = (w - 1) / 2.0, (h - 1) / 2.0
    x, y = xx - cx, yy - cy
    half_diag = 0.5 * math.hypot(w, h)
    bands = np.zeros((h, w))
    for _ in range(n_bands):
        th = rng.uniform(0, math.pi)
        rho = rng.uniform(-half_diag, half_diag)
        sigma = rng.uniform(6, 20)
        amp = rng.uniform(500, 3000)
        d = x * math.cos(th) + y * math.sin(th) - rho
        bands += amp * np.exp(-d * d / (2 * sigma * sigma))
    return background + bands * np.exp(-(x * x + y * y) / (2 * falloff * falloff))


def deform(clean, eps, t):
    """Bilinear resampling of `clean` under x' = (I + eps)(x - c) + c + t."""
    h, w = clean.shape
    a = np.eye(2) + np.asarray(eps, np.float64).reshape(2, 2)
    ainv = np.linalg.inv(a)
    cx, cy = (w - 1) / 2.0, (h - 1) / 2.0
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    u = xx - cx - t[0]
    v = yy - cy - t[1]
    sx = ainv[0, 0] * u + ainv[0, 1] * v + cx
    sy = ainv[1, 0] * u + ainv[1, 1] * v + cy
    if _ndi is not None:
        return _ndi.map_coordinates(clean, [sy, sx], order=1, mode="nearest")
    sx = np.clip(sx, 0, w - 1)
    sy = np.clip(sy, 0, h - 1)
    x0 = np.minimum(np.floor(sx).astype(int), w - 2)
    y0 = np.minimum(np.floor(sy).astype(int), h - 2)
    fx, fy = sx - x0, sy - y0
    return ((1 - fy) * ((1 - fx) * clean[y0, x0] + fx * clean[y0, x0 + 1]) +
            fy * ((1 - fx) * clean[y0 + 1, x0] + fx * clean[y0 + 1, x0 + 1]))


def poisson_u16(img, seed):
    rng = np.random.default_rng(seed)
    return np.clip(rng.poisson(np.maximum(img, 0)), 0, 65535).astype(np.uint16)


def grid_centres(extent, k, radius):
    """linspace(radius, extent - radius, k) rounded half-down."""
    c = np.linspace(radius, extent - radius, k)
    return np.ceil(c - 0.5).astype(int)


def grid_begins(size, k, radius):
    """ROI corners (x, y) of a k x k grid of 2r x 2r ROIs, row-major, x fastest."""
    w, h = (size, size) if np.isscalar(size) else size
    xs = grid_centres(w, k, radius) - radius
    ys = grid_centres(h, k, radius) - radius
    return [(int(x), int(y)) for y in ys for x in xs]


def displacement_at(points, eps, t, size):
    """eps (p - c) + t at ROI centres p ((n,2) array of x, y)."""
    w, h = (size, size) if np.isscalar(size) else size
    c = np.array([(w - 1) / 2.0, (h - 1) / 2.0])
    p = np.asarray(points, np.float64) - c
    return p @ np.asarray(eps, np.float64).reshape(2, 2).T + np.asarray(t, np.float64)


class PatternSet:
    """A reference pattern and deformed patterns with their deformation parameters."""

    def __init__(self, name, size, n_patterns, eps_sigma=1e-3, t_range=2.0, n_bands=24, seed=0, unique=None):
        self.name = name
        self.size = (size, size) if np.isscalar(size) else tuple(size)
        self.clean = band_reference(self.size, n_bands, seed_of(name, seed, "ref"))
        self.ref = poisson_u16(self.clean, seed_of(name, seed, "ref-noise"))
        self.eps, self.t = [], []
        unique = n_patterns if unique is None else min(unique, n_patterns)
        w, h = self.size
        base = np.empty((unique, h, w), np.uint16)
        for i in range(unique):
            rng = np.random.default_rng(seed_of(name, seed, "def", i))
            eps = rng.normal(0, eps_sigma, (2, 2))
            t = rng.uniform(-t_range, t_range, 2)
            base[i] = poisson_u16(deform(self.clean, eps, t), seed_of(name, seed, "def-noise", i))
            self.eps.append(eps)
            self.t.append(t)
        reps = -(-n_patterns // unique)
        self.pics = np.tile(base, (reps, 1, 1))[:n_patterns] if reps > 1 else base
        self.eps = [self.eps[i % unique] for i in range(n_patterns)]
        self.t = [self.t[i % unique] for i in range(n_patterns)]

    def expected_offsets(self, begins, slice_size):
        """-(displacement) at every ROI centre, (n_patterns * n_rois, 2)."""
        sx, sy = (slice_size, slice_size) if np.isscalar(slice_size) else slice_size
        centres = np.array([(x + sx / 2.0, y + sy / 2.0) for x, y in begins])
        out = [-displacement_at(centres, e, t, self.size) for e, t in zip(self.eps, self.t)]
        return np.concatenate(out)


def scan_raster(nx=100, ny=200, step=60.0):
    """Triangle raster of a scan: x = step i + (step / 2)(j mod 2), y = floor(step j sqrt(0.75))."""
    i, j = np.meshgrid(np.arange(nx), np.arange(ny))
    x = step * i + (step / 2) * (j % 2)
    y = np.floor(step * j * math.sqrt(0.75))
    return np.stack([x.ravel(), y.ravel()], axis=1), (ny, nx)


def smooth_strain_field(shape, amplitude=2e-3, corr_length=10.0, seed=0):
    """eps (2 x 2 per raster point) varying smoothly over the raster: white noise per
    component, Gaussian-filtered with sigma = corr_length points, rescaled to the given
    standard deviation.  Returns (ny * nx, 2, 2)."""
    rng = np.random.default_rng(seed)
    comps = []
    for _ in range(4):
        f = rng.standard_normal(shape)
        f = _ndi.gaussian_filter(f, corr_length, mode="wrap") if _ndi is not None else f
        comps.append(f / f.std() * amplitude)
    return np.stack(comps, axis=-1).reshape(-1, 2, 2)


_RENDER = {}


def _render_init(clean, eps_field, name, seed, t_range):
    _RENDER.update(clean=clean, eps_field=eps_field, name=name, seed=seed, t_range=t_range)


def _render_point(i):
    """One raster point of a scan (process-pool worker): its deformed pattern and t."""
    r = _RENDER
    rng = np.random.default_rng(seed_of(r["name"], r["seed"], "t", i))
    t = rng.uniform(-r["t_range"], r["t_range"], 2)
    pic = poisson_u16(deform(r["clean"], r["eps_field"][i], t), seed_of(r["name"], r["seed"], "def-noise", i))
    return pic, t


class ScanSet:
    """A scan: one reference and deformed patterns at raster points whose strain varies
    smoothly over the raster.  Only the first `unique` raster points are rendered;
    requests cycle over them.  render() draws any other raster points the same way."""

    def __init__(self, name, size, unique, t_range=2.0, n_bands=24, seed=0):
        self.name = name
        self.seed = seed
        self.t_range = t_range
        self.size = (size, size) if np.isscalar(size) else tuple(size)
        self.points, shape = scan_raster()
        self.eps_field = smooth_strain_field(shape, seed=seed_of(name, seed, "eps"))
        self.clean = band_reference(self.size, n_bands, seed_of(name, seed, "ref"))
        self.ref = poisson_u16(self.clean, seed_of(name, seed, "ref-noise"))
        self.pics, self.t = self.render(range(unique))
        self.eps = [self.eps_field[i] for i in range(unique)]

    def render(self, indices, workers=1):
        """Deformed patterns (len, h, w) and translations of the raster points `indices`;
        workers > 1 renders them in a process pool."""
        indices = list(indices)
        w, h = self.size
        pics = np.empty((len(indices), h, w), np.uint16)
        ts = []
        args = (self.clean, self.eps_field, self.name, self.seed, self.t_range)
        if workers > 1 and len(indices) > 1:
            import concurrent.futures as cf
            import multiprocessing as mp

            with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("fork"),
                                        initializer=_render_init, initargs=args) as ex:
                for k, (pic, t) in enumerate(ex.map(_render_point, indices, chunksize=4)):
                    pics[k] = pic
                    ts.append(t)
        else:
            _render_init(*args)
            for k, i in enumerate(indices):
                pics[k], t = _render_point(i)
                ts.append(t)
        return pics, ts


# The five workloads of the benchmark configurations (pattern side, patterns,
# ROI grid k x k, ROI radius, policy, precision, fitsize, batch).
CONFIGS = {
    "brute_f64_256": dict(size=256, patterns=16, k=5, radius=16, policy="brute", dtype="f64", fit=3, batch=16,
                          t_range=2.0),
    "fft_f32_873": dict(size=873, patterns=256, k=20, radius=32, policy="fft", dtype="f32", fit=5, batch=64,
                        t_range=3.0),
    "brute_873_r64": dict(size=873, patterns=128, k=10, radius=64, policy="brute", dtype="f64", fit=7, batch=128,
                          t_range=3.0),
    # the precision-form shape through the FFT policy in float (256-point transforms)
    "fft_f32_873_r64": dict(size=873, patterns=128, k=10, radius=64, policy="fft", dtype="f32", fit=7, batch=64,
                            t_range=3.0),
    # serving: Poisson arrivals of single patterns, each drawing a ROI radius and a
    # k x k grid; FFT in float, matrix-core (split bf16) brute for the smallest radius
    "serve_873": dict(size=873, rate=2000.0, duration=60.0, radii=(16, 32, 64), grids=(10, 15, 20), fit=5,
                      hbm_budget=64 << 30, t_range=3.0, policy="serve", dtype="f32"),
    # a scan streamed through the request loop: 20,000 patterns of 1024 x 1024 on a
    # triangle raster, 31 x 31 ROIs of 64 x 64, FFT float and split-bf16 matrix-core
    # brute, fitsize 9, error vs FP64 on a subset
    "scan_1024": dict(size=1024, patterns=20000, k=31, radius=32, fit=9, subset=500, t_range=2.0, policy="scan",
                      dtype="f32", hbm_budget=64 << 30),
    # the file driver end to end: 1024 TIFF files of 873 x 873 read by 1..16 loader
    # threads, uploaded and run through the FFT policy (400 ROIs of 64 x 64, float)
    "files_873": dict(size=873, patterns=1024, k=20, radius=32, fit=5, batch=8, t_range=3.0, policy="files",
                      dtype="f32", workers=(1, 2, 4, 8, 16)),
    # the deformation post-step over a scan's worth of ROI results: 20,000 patterns x
    # 400 ROIs (20 x 20 grid on 873 x 873) of offsets and fit coefficients in HBM
    "post_873": dict(size=873, patterns=20000, k=20, radius=32, policy="post", dtype="f64"),
}
