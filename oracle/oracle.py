"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the oracle restates the reference's CPU develop path
(crates/flashraw/src/process.rs, crates/flashraw/src/ahd.rs) and is the checker the HIP path is
compared against, and the CPU baseline bench.py times. Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module; the product package never does.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

# CfaColor codes (crates/flashraw-core/src/types.rs:84-102)
R, G, B = 0, 1, 2
PATTERNS = {
    "RGGB": [[R, G], [G, B]],
    "BGGR": [[B, G], [G, R]],
    "GRBG": [[G, R], [B, G]],
    "GBRG": [[G, B], [R, G]],
}

DEMOSAIC = {"bilinear": 0, "ahd": 1, "edge": 2}
ENCODING = {"rgba8": 0, "srgb16": 1, "linear16": 2, "linear_f16": 3, "linear_f32": 4, "rgb8": 5}
_DTYPE = {0: np.uint8, 1: np.uint16, 2: np.uint16, 3: np.uint16, 4: np.float32, 5: np.uint8}


class FroFrame(ctypes.Structure):
    _fields_ = [
        ("u16", ctypes.c_void_p),
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("origin_x", ctypes.c_uint32),
        ("origin_y", ctypes.c_uint32),
        ("cfa_w", ctypes.c_uint32),
        ("cfa_h", ctypes.c_uint32),
        ("cfa", ctypes.c_uint8 * 64),
        ("black_rows", ctypes.c_uint32),
        ("black_cols", ctypes.c_uint32),
        ("black_samples", ctypes.c_uint32),
        ("black_values", ctypes.c_void_p),
        ("n_black_values", ctypes.c_uint32),
        ("row_deltas", ctypes.c_void_p),
        ("n_row_deltas", ctypes.c_uint32),
        ("col_deltas", ctypes.c_void_p),
        ("n_col_deltas", ctypes.c_uint32),
        ("white_values", ctypes.c_void_p),
        ("n_white", ctypes.c_uint32),
        ("f32", ctypes.c_void_p),
        ("cpp", ctypes.c_uint32),
    ]


class FroEdge(ctypes.Structure):
    _fields_ = [
        ("g0", ctypes.c_void_p),
        ("dir", ctypes.c_void_p),
        ("canon", ctypes.c_uint8 * 36),
        ("cx", ctypes.c_uint32),
        ("cy", ctypes.c_uint32),
        ("beta", ctypes.c_float),
        ("period", ctypes.c_uint32),
    ]


class FroTone(ctypes.Structure):
    _fields_ = [
        ("table", ctypes.c_void_p),
        ("n", ctypes.c_uint32),
        ("xmax", ctypes.c_float),
        ("scale", ctypes.c_float),
        ("m", ctypes.c_float * 9),
        ("white", ctypes.c_float * 3),
    ]


class FroParams(ctypes.Structure):
    _fields_ = [
        ("demosaic", ctypes.c_int32),
        ("crop_x", ctypes.c_uint32),
        ("crop_y", ctypes.c_uint32),
        ("crop_w", ctypes.c_uint32),
        ("crop_h", ctypes.c_uint32),
        ("wb", ctypes.c_float * 4),
        ("m", ctypes.c_float * 9),
        ("has_levels", ctypes.c_int32),
        ("levels_scale", ctypes.c_float),
        ("levels_offset", ctypes.c_float),
        ("orientation", ctypes.c_int32),
        ("encoding", ctypes.c_int32),
        ("edge", ctypes.POINTER(FroEdge)),
        ("half_size", ctypes.c_int32),
        ("tone", ctypes.POINTER(FroTone)),
        ("curve8", ctypes.c_void_p),
        ("curve8_len", ctypes.c_uint32),
        ("curve16", ctypes.c_void_p),
        ("curve16_len", ctypes.c_uint32),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        L.fro_develop.restype = ctypes.c_int
        L.fro_develop.argtypes = [ctypes.POINTER(FroFrame), ctypes.POINTER(FroParams), ctypes.c_void_p,
                                  ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32),
                                  ctypes.POINTER(ctypes.c_uint32)]
        L.fro_srgb8_table.argtypes = [ctypes.c_void_p]
        L.fro_srgb16_table.argtypes = [ctypes.c_void_p]
        L.fro_f32_to_f16.restype = ctypes.c_uint16
        L.fro_f32_to_f16.argtypes = [ctypes.c_float]
        L.fro_normalize.restype = ctypes.c_int
        L.fro_normalize.argtypes = [ctypes.POINTER(FroFrame), ctypes.c_void_p]
        L.fro_normalize_clamped.restype = ctypes.c_int
        L.fro_normalize_clamped.argtypes = [ctypes.POINTER(FroFrame), ctypes.c_void_p]
        L.fro_unpack.restype = ctypes.c_int
        L.fro_unpack.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                 ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p]
        L.fro_linearize.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint32]
        L.fro_edge_beta.restype = ctypes.c_float
        L.fro_edge_beta.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                                    ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float]
        L.fro_dct_denoise.restype = ctypes.c_int
        L.fro_dct_denoise.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                                      ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_uint32,
                                      ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        P, U, F = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float
        L.fro_nr_apply.restype = ctypes.c_int
        L.fro_nr_apply.argtypes = [P, U, U, F, F]
        L.fro_nr_apply_order.restype = ctypes.c_int
        L.fro_nr_apply_order.argtypes = [P, U, U, F, F, ctypes.c_int]
        L.fro_nr_box_mean_order.restype = ctypes.c_int
        L.fro_nr_box_mean_order.argtypes = [P, U, U, U, P, ctypes.c_int]
        L.fro_nr_level_count.restype = U
        L.fro_nr_level_count.argtypes = [U, U, U]
        L.fro_nr_estimate_step.restype = U
        L.fro_nr_estimate_step.argtypes = [U, U]
        L.fro_nr_stats.restype = U
        L.fro_nr_stats.argtypes = [P, P, U, U, P]
        L.fro_nr_model_from_stats.restype = ctypes.c_int
        L.fro_nr_model_from_stats.argtypes = [P, U, P, P, P]
        L.fro_nr_dct_block.argtypes = [P, U, U, U, P]
        L.fro_nr_inverse_block.argtypes = [P, P]
        L.fro_nr_box_mean.argtypes = [P, U, U, U, P]
        L.fro_nr_down.argtypes = [P, U, U, P]
        L.fro_nr_add_up.argtypes = [P, U, U, P]
        L.fro_nr_dct_pass.restype = ctypes.c_int
        L.fro_nr_dct_pass.argtypes = [P, P, P, U, U, U, P, P, F, F, U, P]
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data


def make_frame(mosaic, cfa, origin=(0, 0), black=(1, 1, 1, [0.0]), white=(16383,),
               row_deltas=None, col_deltas=None, cpp=1):
    """Build the frame descriptor. `mosaic` is (h, w) or (h, w * cpp) of uint16 or float32 samples
    (float32: raw values, RawImage PixelData::F32); `cfa` a pattern name, a 2-D array of CfaColor
    codes, or None (no CFA); `black` is (rows, cols, samples, values). The arrays the descriptor
    points to are returned alongside it and must stay alive while it is used."""
    arr = np.asarray(mosaic)
    f = FroFrame()
    if arr.dtype == np.float32:
        arr = np.ascontiguousarray(arr)
        f.f32 = arr.ctypes.data
    else:
        arr = np.ascontiguousarray(arr, dtype=np.uint16)
        f.u16 = arr.ctypes.data
    f.cpp = cpp
    f.height = arr.shape[0]
    f.width = arr.shape[1] // cpp
    f.origin_x, f.origin_y = origin
    if cfa is not None:
        if isinstance(cfa, str):
            cfa = PATTERNS[cfa]
        cfa = np.asarray(cfa, dtype=np.uint8)
        f.cfa_h, f.cfa_w = cfa.shape
        for i, v in enumerate(cfa.reshape(-1)):
            f.cfa[i] = int(v)
    rows, cols, samples, values = black
    bv = np.ascontiguousarray(values, dtype=np.float32)
    f.black_rows, f.black_cols, f.black_samples = rows, cols, samples
    f.black_values, f.n_black_values = _ptr(bv), bv.size
    keep = [arr, bv]
    if row_deltas is not None:
        rd = np.ascontiguousarray(row_deltas, dtype=np.float32)
        f.row_deltas, f.n_row_deltas = _ptr(rd), rd.size
        keep.append(rd)
    if col_deltas is not None:
        cd = np.ascontiguousarray(col_deltas, dtype=np.float32)
        f.col_deltas, f.n_col_deltas = _ptr(cd), cd.size
        keep.append(cd)
    wv = np.ascontiguousarray(white, dtype=np.uint32)
    f.white_values, f.n_white = _ptr(wv), wv.size
    keep.append(wv)
    return f, keep


def develop(mosaic, cfa, *, origin=(0, 0), black=(1, 1, 1, [0.0]), white=(16383,), wb=(1, 1, 1, 1),
            matrix=((1, 0, 0), (0, 1, 0), (0, 0, 1)), demosaic="bilinear", encoding="rgba8",
            orientation=1, crop=None, levels=None, row_deltas=None, col_deltas=None, threads=0, edge=None,
            cpp=1, half_size=False, tone=None, curve=None):
    """Fused develop (process.rs:1475-1591) of a frame of `cpp` components (u16 or f32 samples,
    `cfa` None for none). Returns an (out_h, out_w, channels) array of the encoding's sample
    type. `edge` (demosaic "edge") is a dict with g0 (4 x 49), dir (4 x 3 x 4 x 49) float32
    tables, cx, cy (canonical position of sample (0, 0) in the RGGB-ordered canonical pattern)
    and beta. `half_size` is Demosaic::HalfSize. `tone` is a dict (table: (n, 2) float32 pairs,
    xmax, scale, to_prophoto 3x3, white 3) for the display outputs; `curve` a pair (lut8 uint8,
    lut16 float32) of output-curve tables."""
    f, keep = make_frame(mosaic, cfa, origin, black, white, row_deltas, col_deltas, cpp)
    h, w = f.height, f.width
    p = FroParams()
    p.demosaic = DEMOSAIC[demosaic]
    cx, cy, cw, ch = crop if crop is not None else (0, 0, w, h)
    p.crop_x, p.crop_y, p.crop_w, p.crop_h = cx, cy, cw, ch
    for i in range(4):
        p.wb[i] = float(wb[i] if i < len(wb) else 1.0)
    mm = np.asarray(matrix, dtype=np.float32).reshape(-1)
    for i in range(9):
        p.m[i] = float(mm[i])
    if levels is not None:
        p.has_levels, p.levels_scale, p.levels_offset = 1, levels[0], levels[1]
    p.orientation = orientation
    p.encoding = ENCODING[encoding]
    p.half_size = int(bool(half_size))
    if tone is not None:
        tt = np.ascontiguousarray(tone["table"], dtype=np.float32).reshape(-1)
        to = FroTone()
        to.table, to.n = tt.ctypes.data, tt.size // 2
        to.xmax, to.scale = float(tone["xmax"]), float(tone["scale"])
        for i, v in enumerate(np.asarray(tone["to_prophoto"], np.float32).reshape(-1)):
            to.m[i] = float(v)
        for i, v in enumerate(np.asarray(tone["white"], np.float32).reshape(-1)):
            to.white[i] = float(v)
        p.tone = ctypes.pointer(to)
        keep += [tt, to]
    if curve is not None:
        c8 = np.ascontiguousarray(curve[0], dtype=np.uint8)
        c16 = np.ascontiguousarray(curve[1], dtype=np.float32)
        p.curve8, p.curve8_len = c8.ctypes.data, c8.size
        p.curve16, p.curve16_len = c16.ctypes.data, c16.size
        keep += [c8, c16]
    if edge is not None:
        g0 = np.ascontiguousarray(edge["g0"], dtype=np.float32).reshape(-1)
        dr = np.ascontiguousarray(edge["dir"], dtype=np.float32).reshape(-1)
        period = int(edge.get("period", 2))
        if g0.size != period * period * 49 or dr.size != period * period * 3 * 4 * 49:
            raise ValueError("edge-aware tables have the wrong size")
        ed = FroEdge()
        ed.g0, ed.dir = g0.ctypes.data, dr.ctypes.data
        ed.period = period
        for i, c in enumerate(edge.get("canon", (0, 1, 1, 2))):
            ed.canon[i] = c
        ed.cx, ed.cy, ed.beta = edge["cx"], edge["cy"], float(edge.get("beta", 0.0))
        p.edge = ctypes.pointer(ed)
        keep += [g0, dr, ed]
    chans = 4 if p.encoding == 0 else 3
    cw = min(cw, w - min(cx, w))
    ch = min(ch, h - min(cy, h))
    out = np.zeros(max(cw * ch * chans, 1), dtype=_DTYPE[p.encoding])
    ow, oh = ctypes.c_uint32(), ctypes.c_uint32()
    rc = lib().fro_develop(ctypes.byref(f), ctypes.byref(p), out.ctypes.data, out.nbytes, threads,
                           ctypes.byref(ow), ctypes.byref(oh))
    del keep
    if rc != 0:
        raise ValueError("oracle develop rejected the request")
    return out[: oh.value * ow.value * chans].reshape(oh.value, ow.value, chans)


def srgb8_table():
    t = np.zeros(4097, dtype=np.uint8)
    lib().fro_srgb8_table(t.ctypes.data)
    return t


def srgb16_table():
    t = np.zeros(16385, dtype=np.float32)
    lib().fro_srgb16_table(t.ctypes.data)
    return t


def f32_to_f16(x):
    return lib().fro_f32_to_f16(float(x))


def dct_denoise(mosaic, black, sig_a, sig_b, thr=2.7, step=4, precision=64, threads=0):
    """CFA-plane DCT block denoise; returns (f32 mosaic in DN above black, u16 mosaic)."""
    mosaic = np.ascontiguousarray(mosaic, dtype=np.uint16)
    h, w = mosaic.shape
    bl = np.ascontiguousarray(black, dtype=np.float32)
    of = np.zeros((h, w), dtype=np.float32)
    ou = np.zeros((h, w), dtype=np.uint16)
    rc = lib().fro_dct_denoise(mosaic.ctypes.data, w, h, bl.ctypes.data, sig_a, sig_b, thr, step, precision,
                               of.ctypes.data, ou.ctypes.data, threads)
    if rc != 0:
        raise ValueError("oracle denoise rejected the request")
    return of, ou


def normalize(mosaic, cfa, origin=(0, 0), black=(1, 1, 1, [0.0]), white=(16383,), row_deltas=None,
              col_deltas=None, cpp=1, clamped=False):
    """The normalised frame as f32 (h, w * cpp): normalize_row of every row (process.rs:881-945),
    or with `clamped` process::normalize (process.rs:302-320, each value max(v, 0))."""
    f, keep = make_frame(mosaic, cfa, origin, black, white, row_deltas, col_deltas, cpp)
    out = np.zeros(np.asarray(mosaic).shape, dtype=np.float32)
    fn = lib().fro_normalize_clamped if clamped else lib().fro_normalize
    if fn(ctypes.byref(f), out.ctypes.data) != 0:
        raise ValueError("oracle rejected the frame")
    return out


PACK_ORDER = {"msb_bytes": 0, "lsb_bytes": 1, "msb16": 2, "msb32": 3}


def unpack(src, w, h, bits, order, stride):
    """unpack_row_scalar over h rows of w samples (crates/flashraw-codecs/src/packed.rs:137-180)."""
    src = np.ascontiguousarray(src, dtype=np.uint8)
    out = np.zeros((h, w), dtype=np.uint16)
    o = PACK_ORDER[order] if isinstance(order, str) else int(order)
    if lib().fro_unpack(src.ctypes.data, src.size, w, h, bits, o, stride, out.ctypes.data) != 0:
        raise ValueError("oracle rejected the unpack")
    return out


def linearize(v, lut):
    """v -> lut[min(v, len - 1)] (run.rs:38-39) on a copy of the uint16 array v."""
    out = np.ascontiguousarray(v, dtype=np.uint16).copy()
    lut = np.ascontiguousarray(lut, dtype=np.uint16)
    lib().fro_linearize(out.ctypes.data, out.size, lut.ctypes.data, lut.size)
    return out


def edge_beta(norm, cx, cy, g_gain, canon=(0, 1, 1, 2)):
    """Plan::set_noise_from (edgeaware/mod.rs:369-409) over a normalised frame; `canon` the
    canonical pattern (4 or 36 colours)."""
    norm = np.ascontiguousarray(norm, dtype=np.float32)
    c = np.ascontiguousarray(canon, dtype=np.uint8)
    h, w = norm.shape
    period = 6 if c.size == 36 else 2
    return lib().fro_edge_beta(norm.ctypes.data, w, h, c.ctypes.data, period, cx, cy, g_gain)


# ---- luma / chroma noise reduction (crates/flashraw/src/denoise/mod.rs, nr.c) ----

def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def nr_apply(rgb, luma, chroma, box_direct=False):
    """denoise::apply on a copy of linear RGB (h, w, 3) float32. box_direct: the guided filter's
    box means as direct window sums in the reference GPU kernels' order (denoise.wgsl:191-218)
    instead of the CPU path's running sums (mod.rs:280-335)."""
    out = _f32(rgb).copy()
    h, w = out.shape[:2]
    if lib().fro_nr_apply_order(out.ctypes.data, w, h, luma, chroma, int(box_direct)) != 0:
        raise ValueError("oracle rejected the noise reduction")
    return out


def nr_level_count(w, h, mx):
    return lib().fro_nr_level_count(w, h, mx)


def nr_estimate_step(w, h):
    return lib().fro_nr_estimate_step(w, h)


def nr_stats(p, guide):
    """The measured blocks of plane p binned by guide: (n, 64) float32 (entry 0 the guide mean)."""
    p, guide = _f32(p), _f32(guide)
    h, w = p.shape
    n = lib().fro_nr_stats(p.ctypes.data, guide.ctypes.data, w, h, None)
    out = np.zeros((max(n, 1), 64), np.float32)
    if n:
        lib().fro_nr_stats(p.ctypes.data, guide.ctypes.data, w, h, out.ctypes.data)
    return out[:n]


def nr_model_from_stats(stats):
    """Model::from_stats: (means (bins,), sigmas (bins, 64)) or None."""
    st = _f32(stats).reshape(-1, 64)
    bins = ctypes.c_uint32(0)
    means = np.zeros(12, np.float32)
    sig = np.zeros((12, 64), np.float32)
    if lib().fro_nr_model_from_stats(st.ctypes.data, st.shape[0], ctypes.byref(bins), means.ctypes.data,
                                     sig.ctypes.data) != 0:
        return None
    return means[:bins.value].copy(), sig[:bins.value].copy()


def nr_dct_block(block):
    """Dct::block of an 8x8 array: (8, 8) with [l, k] = horizontal l, vertical k."""
    b = _f32(block)
    out = np.zeros(64, np.float32)
    lib().fro_nr_dct_block(b.ctypes.data, 8, 0, 0, out.ctypes.data)
    return out.reshape(8, 8)


def nr_inverse_block(coef):
    c = _f32(coef)
    out = np.zeros(64, np.float32)
    lib().fro_nr_inverse_block(c.ctypes.data, out.ctypes.data)
    return out.reshape(8, 8)


def nr_box_mean(p, r, direct=False):
    p = _f32(p)
    h, w = p.shape
    out = np.zeros_like(p)
    lib().fro_nr_box_mean_order(p.ctypes.data, w, h, r, out.ctypes.data, int(direct))
    return out


def nr_down(p):
    p = _f32(p)
    h, w = p.shape
    out = np.zeros(((h + 1) // 2, (w + 1) // 2), np.float32)
    lib().fro_nr_down(p.ctypes.data, w, h, out.ctypes.data)
    return out


def nr_add_up(d, coarse):
    out = _f32(d).copy()
    h, w = out.shape
    c = _f32(coarse)
    assert c.shape == ((h + 1) // 2, (w + 1) // 2)
    lib().fro_nr_add_up(out.ctypes.data, w, h, c.ctypes.data)
    return out


def nr_dct_pass(src, guide, pilot, means, sigmas, gain=1.0, hard=2.7, step=2):
    src, guide = _f32(src), _f32(guide)
    h, w = src.shape
    pil = None if pilot is None else _f32(pilot)
    m, s = _f32(means), _f32(sigmas).reshape(-1, 64)
    out = np.zeros_like(src)
    if lib().fro_nr_dct_pass(src.ctypes.data, guide.ctypes.data, None if pil is None else pil.ctypes.data, w, h,
                             m.size, m.ctypes.data, s.ctypes.data, gain, hard, step, out.ctypes.data) != 0:
        raise ValueError("oracle rejected the pass")
    return out
