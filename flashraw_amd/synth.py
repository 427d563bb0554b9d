"""Seeded synthetic Bayer mosaics (flashraw_amd/lib/libfrsynth.so, host code).

The workloads: a 24 MP RGGB 14-bit frame (6048x4024, black 512, white 16383, gain 0.5 e-/DN,
read sigma 3 DN; 200 edges, 50 discs), its 45 MP BGGR variant with dense stripe patches
(8256x5504, 500 stripe patches), and the denoise frame with read sigma 12 DN.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libfrsynth.so")
_lib = None

BAYER = {"RGGB": (0, 1, 1, 2), "BGGR": (2, 1, 1, 0), "GRBG": (1, 0, 2, 1), "GBRG": (1, 2, 0, 1)}


def _L():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            raise ImportError(f"libfrsynth.so not built ({_PATH}); run __graft_entry__.build()")
        L = ctypes.CDLL(_PATH)
        L.frs_scene.restype = ctypes.c_int
        L.frs_scene.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_float,
                                ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int]
        L.frs_scene_clean.restype = ctypes.c_int
        L.frs_scene_clean.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                                      ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_uint32,
                                      ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int]
        _lib = L
    return _lib


def mosaic(width, height, cfa="RGGB", black=512.0, white=16383.0, gain=0.5, read_sigma=3.0, edges=200, discs=50,
           stripes=0, seed=0, threads=0) -> np.ndarray:
    """A (height, width) uint16 Bayer mosaic of the seeded scene."""
    out = np.empty((height, width), dtype=np.uint16)
    c = np.asarray(BAYER[cfa] if isinstance(cfa, str) else cfa, dtype=np.uint8)
    rc = _L().frs_scene(out.ctypes.data, width, height, c.ctypes.data, black, white, gain, read_sigma, edges, discs,
                        stripes, seed, threads)
    if rc != 0:
        raise ValueError("invalid scene parameters")
    return out


def clean_mosaic(width, height, cfa="RGGB", black=512.0, white=16383.0, gain=0.5, edges=200, discs=50, stripes=0,
                 seed=0, threads=0) -> np.ndarray:
    """The mosaic of the same seeded scene before noise and rounding: black + signal, f32 DN."""
    out = np.empty((height, width), dtype=np.float32)
    c = np.asarray(BAYER[cfa] if isinstance(cfa, str) else cfa, dtype=np.uint8)
    rc = _L().frs_scene_clean(out.ctypes.data, width, height, c.ctypes.data, black, white, gain, edges, discs,
                              stripes, seed, threads)
    if rc != 0:
        raise ValueError("invalid scene parameters")
    return out


def frame_24mp(seed=0, **kw):
    """6048x4024 RGGB, 14-bit, black 512, white 16383, gain 0.5 e-/DN, read sigma 3 DN."""
    return mosaic(6048, 4024, "RGGB", 512.0, 16383.0, 0.5, 3.0, 200, 50, 0, seed, **kw)


def frame_45mp_stripes(seed=1, **kw):
    """8256x5504 BGGR with 500 stripe patches (periods 2-6 px)."""
    return mosaic(8256, 5504, "BGGR", 512.0, 16383.0, 0.5, 3.0, 200, 50, 500, seed, **kw)


def frame_24mp_noisy(seed=2, **kw):
    """The 6048x4024 RGGB frame with read sigma 12 DN (denoise workload)."""
    return mosaic(6048, 4024, "RGGB", 512.0, 16383.0, 0.5, 12.0, 200, 50, 0, seed, **kw)


def frame_24mp_clean(seed=2, **kw):
    """The denoise workload's scene before noise (f32 DN, black included)."""
    return clean_mosaic(6048, 4024, "RGGB", 512.0, 16383.0, 0.5, 200, 50, 0, seed, **kw)


# Develop parameters of the workloads: WB (2.1, 1, 1.6) and a row-normalised camera->sRGB
# matrix with a diagonal near 1.6 and off-diagonals in [-0.5, -0.1].
WB = (2.1, 1.0, 1.6, 1.0)
MATRIX = ((1.62, -0.38, -0.24), (-0.21, 1.58, -0.37), (-0.12, -0.45, 1.57))


def edge_tables(seed=0):
    """Synthetic edge-aware filter tables for a Bayer pattern with the canonical colours
    (R, G, G, B), in the layout the edge-aware spec takes: g0 [4][49] and dir [4][3][4][49].

    They have the sparsity of trained Bayer tables, which sets the kernel's cost: the first green
    estimate is the sample itself at the green positions and a full 7x7 filter at red and blue;
    the directional filters are full 7x7 for the two missing colours and zero for the own colour.
    Every filter sums to one (a flat field stays flat) and leans along its direction, so the
    result is a plausible demosaic; the coefficients are not trained."""
    rng = np.random.default_rng(seed)
    canon = (0, 1, 1, 2)
    yy, xx = np.mgrid[-3:4, -3:4]
    along = [(0, 1), (1, 0), (1, 1), (1, -1)]  # horizontal, vertical, and the two diagonals

    def filt(d=None):
        if d is None:
            f = np.exp(-(xx ** 2 + yy ** 2) / 3.0)
        else:
            u = (xx * d[0] + yy * d[1]) / np.hypot(*d)     # distance across the direction
            v = (xx * d[1] - yy * d[0]) / np.hypot(*d)     # distance along it
            f = np.exp(-(u ** 2) / 0.6 - (v ** 2) / 4.0)
        f = f * rng.uniform(0.8, 1.2, f.shape)
        return (f / f.sum()).reshape(-1).astype(np.float32)

    g0 = np.zeros((4, 49), np.float32)
    dirt = np.zeros((4, 3, 4, 49), np.float32)
    for ph in range(4):
        if canon[ph] == 1:
            g0[ph, 24] = 1.0
        else:
            g0[ph] = filt()
        for c in range(3):
            if c == canon[ph]:
                continue
            for d in range(4):
                dirt[ph, c, d] = filt(along[d])
    return g0, dirt


_GOLDEN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "golden")


def trained_edge_tables(kind="bayer"):
    """The reference's trained edge-aware filter tables (crates/flashraw/src/edgeaware/bayer.bin,
    xtrans.bin, kept as test data under tests/golden/edge_{bayer,xtrans}.bin): little-endian f32,
    g0 [P*P][49] then dir [P*P][3][4][49] (load, edgeaware/mod.rs:131-135), P = 2 or 6. Returns
    (g0, dir) shaped that way."""
    p = 2 if kind == "bayer" else 6
    n = p * p
    f = np.fromfile(os.path.join(_GOLDEN, f"edge_{kind}.bin"), dtype="<f4")
    if f.size != n * 49 + n * 3 * 4 * 49:
        raise ValueError(f"edge_{kind}.bin has {f.size} floats")
    return f[: n * 49].reshape(n, 49).copy(), f[n * 49:].reshape(n, 3, 4, 49).copy()


# ---- the batched and served workloads ----

def batch_frames(n, seed=3):
    """The batched workload: n frames with sizes 12, 20, 24, 33, 45 or 61 MP at 3:2 or 4:3, a
    uniformly drawn Bayer phase, 12 or 14 bits, WB R and B in [1.2, 2.8] and the colour matrix
    jittered by up to 10 %, AHD to RGBA8 (sides rounded to even). Each entry carries the seed of
    its mosaic (black 64 * (bits - 10), white 2^bits - 1)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        mp = float(rng.choice([12, 20, 24, 33, 45, 61]))
        aspect = 1.5 if rng.random() < 0.5 else 4.0 / 3.0
        h = int(round((mp * 1e6 / aspect) ** 0.5 / 2)) * 2
        w = int(round(h * aspect / 2)) * 2
        out.append(dict(w=w, h=h, cfa=["RGGB", "BGGR", "GRBG", "GBRG"][int(rng.integers(0, 4))],
                        bits=12 if rng.random() < 0.5 else 14, wb=(float(rng.uniform(1.2, 2.8)), 1.0,
                                                                      float(rng.uniform(1.2, 2.8)), 1.0),
                        jitter=rng.uniform(0.9, 1.1, (3, 3)), seed=1000 + k))
    return out


def batch_mosaic(f):
    """The mosaic of a batch_frames entry."""
    white = (1 << f["bits"]) - 1
    return mosaic(f["w"], f["h"], f["cfa"], black=64.0 * (f["bits"] - 10), white=float(white), seed=f["seed"])


def jittered_matrix(jitter):
    """MATRIX scaled element-wise by `jitter` (3x3), as a tuple of rows."""
    return tuple(tuple(float(v) for v in row) for row in np.asarray(MATRIX) * jitter)


N_SERVE_MOSAICS = 32


def serve_mosaic_shapes(seed=5, n=N_SERVE_MOSAICS):
    """Sizes of the served request mosaics: log-uniform in 12-102 MP at 3:2 or 4:3 (sides
    multiples of 8), a Bayer phase and a bit depth each."""
    import math
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        mp = math.exp(float(rng.uniform(math.log(12.0), math.log(102.0))))
        aspect = 1.5 if rng.random() < 0.5 else 4.0 / 3.0
        h = max(16, int((mp * 1e6 / aspect) ** 0.5) // 8 * 8)
        w = max(16, int(h * aspect) // 8 * 8)
        out.append(dict(w=w, h=h, cfa=["RGGB", "BGGR", "GRBG", "GBRG"][int(rng.integers(0, 4))],
                        bits=12 if rng.random() < 0.5 else 14, seed=100 + k))
    return out


def serve_requests(rate, duration, seed, n_mosaics=N_SERVE_MOSAICS):
    """Poisson arrivals at `rate` req/s for `duration` s. Each request takes one of the pooled
    mosaics (so sizes follow the log-uniform 12-102 MP draw) with its own WB and matrix, 30 % AHD,
    20 % denoised first."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t > duration:
            return out
        out.append(dict(t=t, m=int(rng.integers(0, n_mosaics)),
                        wb=(float(rng.uniform(1.2, 2.8)), 1.0, float(rng.uniform(1.2, 2.8)), 1.0),
                        jitter=rng.uniform(0.9, 1.1, (3, 3)), ahd=bool(rng.random() < 0.3),
                        denoise=bool(rng.random() < 0.2)))

