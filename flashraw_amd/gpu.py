"""Host-side mirror of the reference's `flashraw-gpu` develop interface over the C ABI.

Names, argument meaning and error behaviour follow crates/flashraw-gpu/src/run.rs
(`GpuFrame`, `GpuOp`, `DevelopSpec`, `EdgeSpec`, `DevelopOutput`, `GpuTimings`, `GpuRun`),
crates/flashraw-gpu/src/context.rs (`context()`, `available()`, `override_limits`) and
crates/flashraw-gpu/src/image.rs (`GpuImage`). Every call goes to libflashraw_hip.so; there is no
CPU fallback, and the module fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes
import enum
import os
import threading
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libflashraw_hip.so")


# ---- errors (crates/flashraw-core/src/error.rs kinds; numbering of flashraw.h:44-114) ----

class FlashRawError(Exception):
    kind = "Other"

    def __init__(self, code: int, message: str):
        super().__init__(f"{self.kind}: {message}")
        self.code = code
        self.message = message


class GpuUnavailable(FlashRawError):
    kind = "GpuUnavailable"


class OutOfMemory(FlashRawError):
    kind = "OutOfMemory"


class LimitExceeded(FlashRawError):
    kind = "LimitExceeded"


class NotAvailable(FlashRawError):
    kind = "NotAvailable"


class InvalidArgument(FlashRawError):
    kind = "InvalidArgument"


class Internal(FlashRawError):
    kind = "Internal"


class Cancelled(FlashRawError):
    kind = "Cancelled"


_ERRORS = {6: GpuUnavailable, 7: OutOfMemory, 8: LimitExceeded, 9: NotAvailable, 10: InvalidArgument,
           12: Internal, 14: Cancelled}


# ---- C structures (include/flashraw_hip.h) ----

class _Rect(ctypes.Structure):
    _fields_ = [("x", ctypes.c_uint32), ("y", ctypes.c_uint32), ("width", ctypes.c_uint32),
                ("height", ctypes.c_uint32)]


class _Frame(ctypes.Structure):
    _fields_ = [
        ("size", ctypes.c_uint32), ("input", ctypes.c_int32), ("data", ctypes.c_void_p),
        ("data_len", ctypes.c_uint64), ("data_on_device", ctypes.c_int32),
        ("raster_width", ctypes.c_uint32), ("raster_height", ctypes.c_uint32), ("cpp", ctypes.c_uint32),
        ("window", _Rect), ("origin_x", ctypes.c_uint32), ("origin_y", ctypes.c_uint32),
        ("linearization", ctypes.c_void_p), ("linearization_len", ctypes.c_uint32),
        ("bits", ctypes.c_uint32), ("order", ctypes.c_int32), ("row_stride", ctypes.c_uint64),
    ]


class _FrameInfo(ctypes.Structure):
    _fields_ = [
        ("size", ctypes.c_uint32), ("cfa_width", ctypes.c_uint32), ("cfa_height", ctypes.c_uint32),
        ("cfa", ctypes.c_uint8 * 64), ("black_rows", ctypes.c_uint32), ("black_cols", ctypes.c_uint32),
        ("black_samples", ctypes.c_uint32), ("black_values", ctypes.c_void_p),
        ("black_values_len", ctypes.c_uint32), ("black_row_deltas", ctypes.c_void_p),
        ("black_row_deltas_len", ctypes.c_uint32), ("black_col_deltas", ctypes.c_void_p),
        ("black_col_deltas_len", ctypes.c_uint32), ("white_values", ctypes.c_void_p),
        ("white_values_len", ctypes.c_uint32),
    ]


class _EdgeSpec(ctypes.Structure):
    _fields_ = [
        ("size", ctypes.c_uint32), ("period", ctypes.c_uint32), ("cx", ctypes.c_uint32), ("cy", ctypes.c_uint32),
        ("beta", ctypes.c_float), ("halo", ctypes.c_uint32), ("canon", ctypes.c_void_p),
        ("canon_len", ctypes.c_uint32), ("g0", ctypes.c_void_p), ("g0_len", ctypes.c_uint32),
        ("dir", ctypes.c_void_p), ("dir_len", ctypes.c_uint32), ("ahd", ctypes.c_int32),
    ]


class _Curve(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("lut8", ctypes.c_void_p), ("lut8_len", ctypes.c_uint32),
                ("lut16", ctypes.c_void_p), ("lut16_len", ctypes.c_uint32)]


class _Tone(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("table", ctypes.c_void_p), ("table_len", ctypes.c_uint32),
                ("xmax", ctypes.c_float), ("scale", ctypes.c_float), ("to_prophoto", ctypes.c_float * 9),
                ("white", ctypes.c_float * 3)]


class _DevelopSpec(ctypes.Structure):
    _fields_ = [
        ("size", ctypes.c_uint32), ("crop", _Rect), ("half_size", ctypes.c_int32), ("wb", ctypes.c_float * 4),
        ("cam_to_rgb", ctypes.c_float * 9), ("orientation", ctypes.c_int32), ("output", ctypes.c_int32),
        ("output2", ctypes.c_int32), ("has_levels", ctypes.c_int32), ("levels", ctypes.c_float * 2),
        ("edge", ctypes.POINTER(_EdgeSpec)), ("curve", ctypes.POINTER(_Curve)), ("tone", ctypes.POINTER(_Tone)),
        ("cancel", ctypes.c_void_p),
    ]


class _Timings(ctypes.Structure):
    _fields_ = [("upload_ms", ctypes.c_double), ("compute_ms", ctypes.c_double), ("kernel_ms", ctypes.c_double),
                ("has_kernel_ms", ctypes.c_int32), ("bands", ctypes.c_uint32)]


class _NrPlane(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_uint32), ("gain", ctypes.c_float), ("steps", ctypes.c_uint32 * 2),
                ("hard_only0", ctypes.c_int32), ("keep", ctypes.c_float)]


class _NrSpec(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("has_luma", ctypes.c_int32), ("has_chroma", ctypes.c_int32),
                ("luma", _NrPlane), ("chroma", _NrPlane), ("chroma_radius", ctypes.c_uint32),
                ("hard", ctypes.c_float), ("max_blocks", ctypes.c_uint32), ("output", ctypes.c_int32),
                ("orientation", ctypes.c_int32), ("tone", ctypes.POINTER(_Tone)), ("curve", ctypes.POINTER(_Curve)),
                ("cancel", ctypes.c_void_p)]


_lib = None
_lib_lock = threading.Lock()
_POLL_FN = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p)


class CancelToken:
    """CancelToken (crates/flashraw-core/src/cancel.rs:26-76) over frh_cancel_token: a shared
    flag that stops a develop or noise reduction between its stages (the call raises Cancelled).
    `poll`, when given, is asked on every check while the token is not cancelled (with_poll)."""

    def __init__(self, poll=None):
        L = lib()
        if poll is None:
            self._cb = None
            self._h = L.frh_cancel_token_new()
        else:
            self._cb = _POLL_FN(lambda _user: 1 if poll() else 0)
            self._h = L.frh_cancel_token_with_poll(self._cb, None)
        if not self._h:
            raise OutOfMemory(7, "cancel token")

    def cancel(self) -> None:
        lib().frh_cancel_token_cancel(self._h)

    def is_cancelled(self) -> bool:
        return bool(lib().frh_cancel_token_is_cancelled(self._h))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.frh_cancel_token_free(self._h)
            self._h = None


def lib():
    """The loaded C ABI library (raises when it has not been built)."""
    global _lib
    with _lib_lock:
        if _lib is None:
            # One HIP runtime per process: torch ships its own libamdhip64 (soname libamdhip64.so.7),
            # so it is loaded first and this library binds to that copy.
            try:
                import torch  # noqa: F401
            except ImportError:
                pass
            if not os.path.exists(_LIB_PATH):
                raise ImportError(f"libflashraw_hip.so not built ({_LIB_PATH}); run __graft_entry__.build()")
            L = ctypes.CDLL(_LIB_PATH)
            vp, i32, u32, u64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint64
            L.frh_available.restype = i32
            L.frh_context_create.argtypes = [i32, ctypes.POINTER(vp)]
            L.frh_context_create.restype = i32
            L.frh_context_destroy.argtypes = [vp]
            L.frh_context_info.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
            L.frh_context_info.restype = i32
            L.frh_set_limits.argtypes = [vp, u64]
            L.frh_set_limits.restype = i32
            L.frh_last_error_message.restype = ctypes.c_char_p
            L.frh_run_develop.argtypes = [vp, ctypes.POINTER(_Frame), ctypes.POINTER(_FrameInfo),
                                          ctypes.POINTER(_DevelopSpec), ctypes.POINTER(vp), ctypes.POINTER(vp),
                                          ctypes.POINTER(_Timings)]
            L.frh_run_develop.restype = i32
            L.frh_run_unpack.argtypes = [vp, ctypes.POINTER(_Frame), ctypes.POINTER(vp), ctypes.POINTER(_Timings)]
            L.frh_run_unpack.restype = i32
            L.frh_run_normalize.argtypes = [vp, ctypes.POINTER(_Frame), ctypes.POINTER(_FrameInfo), ctypes.POINTER(vp),
                                            ctypes.POINTER(_Timings)]
            L.frh_run_normalize.restype = i32
            L.frh_unpack_into.argtypes = [vp, ctypes.POINTER(_Frame), vp, u64, vp]
            L.frh_unpack_into.restype = i32
            L.frh_normalize_into.argtypes = [vp, ctypes.POINTER(_Frame), ctypes.POINTER(_FrameInfo), vp, u64, vp]
            L.frh_normalize_into.restype = i32
            L.frh_develop_into.argtypes = [vp, ctypes.POINTER(_Frame), ctypes.POINTER(_FrameInfo),
                                           ctypes.POINTER(_DevelopSpec), vp, u64, vp, u64, vp]
            L.frh_develop_into.restype = i32
            L.frh_develop_output_size.argtypes = [ctypes.POINTER(_Frame), ctypes.POINTER(_DevelopSpec), i32,
                                                  ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(u32),
                                                  ctypes.POINTER(u64)]
            L.frh_develop_output_size.restype = i32
            for name, rt in [("frh_image_width", u32), ("frh_image_height", u32), ("frh_image_channels", u32),
                             ("frh_image_format", i32), ("frh_image_device_ptr", vp), ("frh_image_bytes", u64)]:
                getattr(L, name).argtypes = [vp]
                getattr(L, name).restype = rt
            L.frh_image_download.argtypes = [vp, vp, u64]
            L.frh_image_download.restype = i32
            L.frh_image_free.argtypes = [vp]
            L.frh_srgb8_table.argtypes = [vp, ctypes.c_size_t]
            L.frh_srgb8_table.restype = i32
            L.frh_check_sqrt.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
            L.frh_check_sqrt.restype = i32
            L.frh_plan_create.argtypes = [vp, ctypes.POINTER(_Frame), ctypes.POINTER(_FrameInfo),
                                          ctypes.POINTER(_DevelopSpec), ctypes.POINTER(vp)]
            L.frh_plan_create.restype = i32
            L.frh_plan_launch.argtypes = [vp, vp, vp, vp, vp]
            L.frh_plan_launch.restype = i32
            L.frh_plan_autotune.argtypes = [vp, vp, vp, vp, vp, i32, ctypes.POINTER(u32)]
            L.frh_plan_autotune.restype = i32
            L.frh_plan_tile_shape.argtypes = [vp, ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(u32)]
            L.frh_plan_tile_shape.restype = i32
            L.frh_plan_output_bytes.argtypes = [vp, i32]
            L.frh_plan_output_bytes.restype = u64
            L.frh_plan_destroy.argtypes = [vp]
            L.frh_context_stream.argtypes = [vp]
            L.frh_context_stream.restype = vp
            L.frh_denoise_cfa.argtypes = [vp, ctypes.POINTER(_DnSpec), vp, vp, vp, vp]
            L.frh_denoise_cfa.restype = i32
            L.frh_nr_spec_for.argtypes = [u32, u32, ctypes.c_float, ctypes.c_float, i32, ctypes.POINTER(_NrSpec)]
            L.frh_nr_spec_for.restype = i32
            L.frh_run_denoise.argtypes = [vp, vp, ctypes.POINTER(_NrSpec), ctypes.POINTER(vp), ctypes.POINTER(_Timings)]
            L.frh_run_denoise.restype = i32
            L.frh_denoise_into.argtypes = [vp, vp, u32, u32, ctypes.POINTER(_NrSpec), vp, u64, vp]
            L.frh_denoise_into.restype = i32
            L.frh_nr_model_from_stats.argtypes = [vp, u32, ctypes.POINTER(u32), vp, vp]
            L.frh_nr_model_from_stats.restype = i32
            L.frh_nr_model_fit.argtypes = [vp, vp, u32, ctypes.POINTER(u32), vp, vp]
            L.frh_nr_model_fit.restype = i32
            L.frh_batch_create.argtypes = [vp, u32, ctypes.POINTER(_Frame), ctypes.POINTER(_FrameInfo),
                                           ctypes.POINTER(_DevelopSpec), ctypes.POINTER(vp), ctypes.POINTER(vp)]
            L.frh_batch_create.restype = i32
            L.frh_batch_launch.argtypes = [vp, vp]
            L.frh_batch_launch.restype = i32
            L.frh_batch_capture.argtypes = [vp, vp]
            L.frh_batch_capture.restype = i32
            L.frh_batch_autotune.argtypes = [vp, vp, i32, ctypes.POINTER(u32), ctypes.POINTER(u32)]
            L.frh_batch_autotune.restype = i32
            L.frh_batch_set_strip_rows.argtypes = [vp, u32, u32]
            L.frh_batch_set_strip_rows.restype = i32
            L.frh_batch_items.argtypes = [vp]
            L.frh_batch_items.restype = u64
            L.frh_batch_destroy.argtypes = [vp]
            L.frh_server_create.argtypes = [vp, ctypes.POINTER(_ServerConfig), ctypes.POINTER(vp)]
            L.frh_server_create.restype = i32
            L.frh_server_submit.argtypes = [vp, ctypes.POINTER(_Frame), ctypes.POINTER(_FrameInfo),
                                            ctypes.POINTER(_DevelopSpec), ctypes.POINTER(_DnSpec), vp, u64,
                                            ctypes.POINTER(u64)]
            L.frh_server_submit.restype = i32
            L.frh_server_wait.argtypes = [vp, u64, i32, ctypes.POINTER(_JobResult), ctypes.POINTER(i32)]
            L.frh_server_wait.restype = i32
            L.frh_server_get_stats.argtypes = [vp, ctypes.POINTER(_ServerStats)]
            L.frh_server_get_stats.restype = i32
            L.frh_server_destroy.argtypes = [vp]
            L.frh_host_alloc.argtypes = [u64]
            L.frh_host_alloc.restype = vp
            L.frh_host_free.argtypes = [vp]
            L.frh_cancel_token_new.restype = vp
            L.frh_cancel_token_with_poll.argtypes = [_POLL_FN, vp]
            L.frh_cancel_token_with_poll.restype = vp
            L.frh_cancel_token_cancel.argtypes = [vp]
            L.frh_cancel_token_is_cancelled.argtypes = [vp]
            L.frh_cancel_token_is_cancelled.restype = i32
            L.frh_cancel_token_free.argtypes = [vp]
            L.frh_context_table_cache.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(u64)]
            L.frh_context_table_cache.restype = i32
            L.frh_plan_kernel.argtypes = [vp]
            L.frh_plan_kernel.restype = i32
            L.frh_batch_develop_noalloc.argtypes = [vp, u32, ctypes.POINTER(_Frame), ctypes.POINTER(_FrameInfo),
                                                    ctypes.POINTER(_DevelopSpec), ctypes.POINTER(vp), u32, u32, vp]
            L.frh_batch_develop_noalloc.restype = i32
            _lib = L
        return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().frh_last_error_message().decode(errors="replace")
        raise _ERRORS.get(rc, FlashRawError)(rc, msg)


# ---- core types (crates/flashraw-core/src/types.rs) ----

class CfaColor(enum.IntEnum):
    Red = 0
    Green = 1
    Blue = 2
    Cyan = 3
    Magenta = 4
    Yellow = 5
    White = 6
    Emerald = 7


@dataclass(frozen=True)
class CfaPattern:
    """CfaPattern (types.rs:152-252): colours row-major, at most 8x8."""
    width: int
    height: int
    colors: tuple

    def color_at(self, x: int, y: int) -> int:
        return self.colors[(y % self.height) * self.width + (x % self.width)]

    def is_bayer(self) -> bool:
        if (self.width, self.height) != (2, 2):
            return False
        return sorted(self.colors) == [0, 1, 1, 2]

    @staticmethod
    def bayer(name: str) -> "CfaPattern":
        m = {"R": 0, "G": 1, "B": 2}
        return CfaPattern(2, 2, tuple(m[c] for c in name))

    @staticmethod
    def parse(rows: str) -> "CfaPattern":
        """Rows of colour letters separated by '/', e.g. the X-Trans
        "GGRGGB/GGBGGR/BRGRBG/GGBGGR/GGRGGB/RBGBRG" (CfaPattern::parse, types.rs)."""
        m = {"R": 0, "G": 1, "B": 2, "C": 3, "M": 4, "Y": 5, "W": 6, "E": 7}
        rr = rows.split("/")
        return CfaPattern(len(rr[0]), len(rr), tuple(m[c] for r in rr for c in r))


CfaPattern.RGGB = CfaPattern.bayer("RGGB")
CfaPattern.BGGR = CfaPattern.bayer("BGGR")
CfaPattern.GRBG = CfaPattern.bayer("GRBG")
CfaPattern.GBRG = CfaPattern.bayer("GBRG")


@dataclass
class BlackLevel:
    """BlackLevel (types.rs:357-436)."""
    rows: int = 1
    cols: int = 1
    samples: int = 1
    values: Sequence[float] = (0.0,)
    row_deltas: Sequence[float] = ()
    col_deltas: Sequence[float] = ()

    @staticmethod
    def uniform(v: float) -> "BlackLevel":
        return BlackLevel(1, 1, 1, (float(v),))

    @staticmethod
    def pattern2x2(v: Sequence[float]) -> "BlackLevel":
        return BlackLevel(2, 2, 1, tuple(float(x) for x in v))


@dataclass
class WhiteLevel:
    """WhiteLevel (types.rs:438-467)."""
    values: Sequence[int] = (65535,)

    @staticmethod
    def uniform(v: int) -> "WhiteLevel":
        return WhiteLevel((int(v),))


@dataclass
class FrameInfo:
    """The FrameInfo fields the develop path reads (crates/flashraw-core/src/decoder.rs:25-74)."""
    width: int
    height: int
    cfa: Optional[CfaPattern] = None
    black_level: BlackLevel = field(default_factory=BlackLevel)
    white_level: WhiteLevel = field(default_factory=WhiteLevel)

    def _c(self):
        s = _FrameInfo()
        s.size = ctypes.sizeof(_FrameInfo)
        keep = []
        if self.cfa is not None:
            s.cfa_width, s.cfa_height = self.cfa.width, self.cfa.height
            for i, c in enumerate(self.cfa.colors):
                s.cfa[i] = int(c)
        b = self.black_level
        s.black_rows, s.black_cols, s.black_samples = b.rows, b.cols, b.samples
        for name, vals in (("black_values", b.values), ("black_row_deltas", b.row_deltas),
                           ("black_col_deltas", b.col_deltas)):
            if len(vals):
                a = np.ascontiguousarray(vals, dtype=np.float32)
                keep.append(a)
                setattr(s, name, a.ctypes.data)
                setattr(s, name + "_len", a.size)
        wv = np.ascontiguousarray(self.white_level.values, dtype=np.uint32)
        keep.append(wv)
        s.white_values, s.white_values_len = wv.ctypes.data, wv.size
        return s, keep


@dataclass
class Rect:
    x: int
    y: int
    width: int
    height: int

    @staticmethod
    def full(w: int, h: int) -> "Rect":
        return Rect(0, 0, w, h)


# ---- run.rs types ----

class DevelopOutput(enum.IntEnum):
    """DevelopOutput (run.rs:96-111); 16-bit encodings hold 3 samples per pixel here."""
    Rgba8 = 0
    LinearRgbF32 = 1
    Rgb16 = 2
    LinearRgb16 = 3
    LinearRgbF16 = 4


class Orientation(enum.IntEnum):
    """Orientation (types.rs:616-640), EXIF numbering."""
    Normal = 1
    FlipH = 2
    Rotate180 = 3
    FlipV = 4
    Transpose = 5
    Rotate90 = 6
    Transverse = 7
    Rotate270 = 8

    def swaps_dimensions(self) -> bool:
        return self.value >= 5


@dataclass
class EdgeSpec:
    """EdgeSpec (run.rs:141-160). The AHD form (Bayer) uses the pattern and the halo only; the
    edge-aware form, Bayer (period 2) or X-Trans (period 6), also `canon` ([P*P]), `g0`
    ([P*P][49]), `dir` ([P*P][3][4][49]) and `beta`, with a halo of 8."""
    period: int = 2
    cx: int = 0
    cy: int = 0
    beta: float = 0.0
    halo: int = 6
    canon: Sequence[int] = ()
    g0: Sequence[float] = ()
    dir: Sequence[float] = ()
    ahd: bool = False


class NoiseReduction(enum.IntEnum):
    """NoiseReduction (crates/flashraw/src/options.rs:341-353): the smoothing weight of the
    edge-aware demosaic is the frame's estimate (Auto), zero (Off) or at least 0.25 / 0.5 / 0.8."""
    Auto = 0
    Off = 1
    Low = 2
    Medium = 3
    High = 4


_LEVEL_FLOOR = {NoiseReduction.Auto: None, NoiseReduction.Off: 0.0, NoiseReduction.Low: 0.25,
                NoiseReduction.Medium: 0.5, NoiseReduction.High: 0.8}  # options.rs:70-78

EDGE_CANON = (0, 1, 1, 2)  # canonical Bayer colours (R, G, G, B), edgeaware/mod.rs:163-166
# canonical X-Trans colours GGRGGB / GGBGGR / BRGRBG / GGBGGR / GGRGGB / RBGBRG (mod.rs:168-185)
EDGE_CANON_XTRANS = tuple({"R": 0, "G": 1, "B": 2}[c] for c in "GGRGGBGGBGGRBRGRBGGGBGGRGGRGGBRBGBRG")
EDGE_HALO = 8              # edgeaware::HALO


def edge_plan(info: FrameInfo, origin=(0, 0), with_pattern=False) -> Optional[tuple]:
    """Plan::new (edgeaware/mod.rs:207-232): the canonical column and row (cx, cy) of frame sample
    (0, 0) when the pattern is a shifted copy of the canonical Bayer (2x2) or X-Trans (6x6) one,
    else None. With `with_pattern`, (cx, cy, canonical colours, period)."""
    if info.cfa is None:
        return None
    cols = tuple(info.cfa.colors)
    for canon, p in ((EDGE_CANON, 2), (EDGE_CANON_XTRANS, 6)):
        if info.cfa.width != p or info.cfa.height != p:
            continue
        for sy in range(p):
            for sx in range(p):
                if all(cols[y * p + x] == canon[((y + sy) % p) * p + (x + sx) % p]
                       for y in range(p) for x in range(p)):
                    cx, cy = (sx + origin[0]) % p, (sy + origin[1]) % p
                    return (cx, cy, canon, p) if with_pattern else (cx, cy)
    return None


def estimate_beta(width: int, height: int, g_gain: float, fetch, cx: int, cy: int, canon=EDGE_CANON) -> float:
    """Plan::set_noise_from (edgeaware/mod.rs:369-409): the smoothing weight from the noise of
    48 rows spread over the frame. `fetch(y)` returns row y of the normalised mosaic (f32,
    `width` samples); `canon` is the canonical pattern (4 or 36 colours). Green pairs one or two
    samples apart (x < width - 2) with a mean in (0.03, 0.6) give |a - b| / sqrt(mean); the lower
    quintile q gives sigma = q / (sqrt(2) * 0.2533) and beta = 0.5 * clamp((sigma^2 - 2e-4) /
    6e-4, 0, 1), all in f32."""
    f32 = np.float32
    if height < 16 or width < 16:
        return 0.0
    canon = np.asarray(canon, dtype=np.int64)
    P = 6 if canon.size == 36 else 2
    rows = 48
    vals = []
    xs = np.arange(width)
    for n in range(rows):
        y = 4 + n * (height - 9) // (rows - 1)
        row = np.asarray(fetch(y), dtype=f32).reshape(-1)[:width]
        green = canon[((y + cy) % P) * P + (xs + cx) % P] == 1
        for d in (1, 2):
            x = xs[:width - 2]
            sel = x[green[x] & green[x + d]]
            a = row[sel] * f32(g_gain)
            b = row[sel + d] * f32(g_gain)
            mid = f32(0.5) * (a + b)
            ok = (mid > f32(0.03)) & (mid < f32(0.6))
            vals.append(np.abs(a[ok] - b[ok]) / np.sqrt(mid[ok]))
    v = np.concatenate(vals) if vals else np.zeros(0, f32)
    if v.size < 1000:
        return 0.0
    k = v.size // 5
    q = np.partition(v, k)[k]
    sigma = f32(q) / (f32(np.sqrt(2.0)) * f32(0.2533))
    nu = sigma * sigma
    t = (nu - f32(2e-4)) / f32(6e-4)
    return float(f32(0.5) * f32(min(max(t, f32(0.0)), f32(1.0))))


def edge_spec(info: FrameInfo, origin, size, wb, fetch, tables, level=NoiseReduction.Auto) -> Optional[EdgeSpec]:
    """flashraw::gpu::edge_spec (crates/flashraw/src/gpu.rs:115-128): the edge-aware spec of a
    Bayer or X-Trans frame of `size` = (width, height) samples whose sample (0, 0) sits at sensor
    `origin`, with the smoothing weight estimated from the frame (`fetch(y)` = normalised row y,
    green gain wb[1]) and then raised to the floor of `level` (Plan::apply_level,
    mod.rs:413-419). `tables` = (g0 [P*P][49], dir [P*P][3][4][49]) are the filter tables of the
    canonical pattern (R, G, G, B for Bayer; the X-Trans arrangement of EDGE_CANON_XTRANS).
    None for other patterns."""
    plan = edge_plan(info, origin, with_pattern=True)
    if plan is None:
        return None
    cx, cy, canon, period = plan
    beta = estimate_beta(int(size[0]), int(size[1]), float(wb[1]), fetch, cx, cy, canon)
    floor = _LEVEL_FLOOR[NoiseReduction(level)]
    if floor is not None:
        beta = 0.0 if floor == 0.0 else max(beta, floor)
    g0, dirt = tables
    return EdgeSpec(period=period, cx=cx, cy=cy, beta=float(beta), halo=EDGE_HALO, canon=canon,
                    g0=np.asarray(g0, np.float32), dir=np.asarray(dirt, np.float32), ahd=False)


def ahd_spec(info: FrameInfo, origin=(0, 0)) -> Optional[EdgeSpec]:
    """flashraw::gpu::ahd_spec (crates/flashraw/src/gpu.rs:133-138): the AHD spec of a Bayer frame
    whose sample (0, 0) sits at sensor `origin` (HALO 6, ahd.rs:33), or None for other patterns."""
    if info.cfa is None or not info.cfa.is_bayer():
        return None
    canon = tuple(info.cfa.colors)
    return EdgeSpec(period=2, cx=origin[0] % 2, cy=origin[1] % 2, beta=0.0, halo=6, canon=canon, ahd=True)


@dataclass
class GpuCurve:
    """GpuCurve (run.rs:181-188): 8-bit codes at len(lut8) - 1 segments of [0, 1] (RGBA8) and
    16-bit codes as f32 at 16385 points (sRGB16)."""
    lut8: Sequence[int] = ()
    lut16: Sequence[float] = ()


@dataclass
class GpuTone:
    """GpuTone (run.rs:190-205): the standard tone of the display outputs. `table` holds
    (value, step) pairs of the curve at evenly spaced sqrt(x / xmax); `scale` is
    len(table) / sqrt(xmax); `to_prophoto` maps linear output RGB to linear ProPhoto and `white`
    is ProPhoto white in the output space."""
    table: Sequence = ()
    xmax: float = 1.0
    scale: float = 1024.0
    to_prophoto: Sequence[Sequence[float]] = ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0))
    white: Sequence[float] = (1.0, 1.0, 1.0)


def _tone_curve_c(tone, curve, keep):
    """Pointers to the C forms of a GpuTone and a GpuCurve (or NULL); `keep` holds their arrays."""
    tp, cp = ctypes.POINTER(_Tone)(), ctypes.POINTER(_Curve)()
    if curve is not None:
        c = _Curve()
        c.size = ctypes.sizeof(_Curve)
        l8 = np.ascontiguousarray(curve.lut8, dtype=np.uint8)
        l16 = np.ascontiguousarray(curve.lut16, dtype=np.float32)
        keep += [l8, l16, c]
        c.lut8, c.lut8_len = (l8.ctypes.data if l8.size else None), l8.size
        c.lut16, c.lut16_len = (l16.ctypes.data if l16.size else None), l16.size
        cp = ctypes.pointer(c)
    if tone is not None:
        t = _Tone()
        t.size = ctypes.sizeof(_Tone)
        tab = np.ascontiguousarray(tone.table, dtype=np.float32).reshape(-1)
        keep += [tab, t]
        t.table, t.table_len = (tab.ctypes.data if tab.size else None), tab.size // 2
        t.xmax, t.scale = float(tone.xmax), float(tone.scale)
        for i, v in enumerate(np.asarray(tone.to_prophoto, np.float32).reshape(-1)):
            t.to_prophoto[i] = float(v)
        for i, v in enumerate(np.asarray(tone.white, np.float32).reshape(-1)):
            t.white[i] = float(v)
        tp = ctypes.pointer(t)
    return tp, cp


@dataclass
class DevelopSpec:
    """DevelopSpec (run.rs:115-135). `output2` asks the same pass for a second encoding."""
    crop: Rect
    wb: Sequence[float] = (1.0, 1.0, 1.0, 1.0)
    cam_to_rgb: Sequence[Sequence[float]] = ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0))
    orientation: Orientation = Orientation.Normal
    output: DevelopOutput = DevelopOutput.Rgba8
    levels: Optional[Sequence[float]] = None
    half_size: bool = False
    curve: object = None
    tone: object = None
    edge: Optional[EdgeSpec] = None
    output2: Optional[DevelopOutput] = None
    cancel: Optional[CancelToken] = None

    def _c(self):
        s = _DevelopSpec()
        s.size = ctypes.sizeof(_DevelopSpec)
        s.crop = _Rect(self.crop.x, self.crop.y, self.crop.width, self.crop.height)
        s.half_size = int(self.half_size)
        for i in range(4):
            s.wb[i] = float(self.wb[i]) if i < len(self.wb) else 1.0
        m = np.asarray(self.cam_to_rgb, dtype=np.float32).reshape(-1)
        for i in range(9):
            s.cam_to_rgb[i] = float(m[i])
        s.orientation = int(self.orientation)
        s.output = int(self.output)
        s.output2 = -1 if self.output2 is None else int(self.output2)
        if self.levels is not None:
            s.has_levels = 1
            s.levels[0], s.levels[1] = float(self.levels[0]), float(self.levels[1])
        keep = []
        s.tone, s.curve = _tone_curve_c(self.tone, self.curve, keep)
        if self.cancel is not None:
            s.cancel = self.cancel._h
            keep.append(self.cancel)
        if self.edge is not None:
            e = _EdgeSpec()
            e.size = ctypes.sizeof(_EdgeSpec)
            e.period, e.cx, e.cy, e.beta, e.halo = self.edge.period, self.edge.cx, self.edge.cy, \
                self.edge.beta, self.edge.halo
            e.ahd = int(self.edge.ahd)
            if len(self.edge.canon):
                c = np.ascontiguousarray(self.edge.canon, dtype=np.uint8)
                keep.append(c)
                e.canon, e.canon_len = c.ctypes.data, c.size
            if len(self.edge.g0):
                g0 = np.ascontiguousarray(self.edge.g0, dtype=np.float32).reshape(-1)
                keep.append(g0)
                e.g0, e.g0_len = g0.ctypes.data, g0.size
            if len(self.edge.dir):
                dr = np.ascontiguousarray(self.edge.dir, dtype=np.float32).reshape(-1)
                keep.append(dr)
                e.dir, e.dir_len = dr.ctypes.data, dr.size
            keep.append(e)
            s.edge = ctypes.pointer(e)
        return s, keep


class DnPrecision(enum.IntEnum):
    """frh_dn_precision: MFMA with f16 or bf16 operands (f32 accumulation), or scalar f32."""
    F16 = 0
    BF16 = 1
    F32 = 2
    F16X3 = 3


class _DnSpec(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("stride", ctypes.c_uint32), ("black", ctypes.c_float * 4), ("sig_a", ctypes.c_float),
                ("sig_b", ctypes.c_float), ("threshold", ctypes.c_float), ("step", ctypes.c_uint32),
                ("precision", ctypes.c_int32), ("seg_rows", ctypes.c_uint32)]


@dataclass
class CfaDenoiseSpec:
    """The CFA-plane DCT block denoiser's parameters (frh_dn_spec): 8x8 blocks at stride `step`
    on each 2x2 phase plane, AC coefficients below threshold * sigma zeroed,
    sigma^2 = sig_a * max(block mean, 0) + sig_b (DN), uniform overlap-add."""
    width: int
    height: int
    black: Sequence[float] = (0.0, 0.0, 0.0, 0.0)
    sig_a: float = 0.0
    sig_b: float = 1.0
    threshold: float = 2.7
    step: int = 4
    precision: DnPrecision = DnPrecision.F16
    stride: int = 0
    seg_rows: int = 0

    def _c(self):
        s = _DnSpec()
        s.size = ctypes.sizeof(_DnSpec)
        s.width, s.height, s.stride = self.width, self.height, self.stride
        for i in range(4):
            s.black[i] = float(self.black[i])
        s.sig_a, s.sig_b, s.threshold = self.sig_a, self.sig_b, self.threshold
        s.step, s.precision, s.seg_rows = self.step, int(self.precision), self.seg_rows
        return s


class _ServerConfig(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("device_budget_bytes", ctypes.c_uint64),
                ("max_batch_frames", ctypes.c_uint32), ("max_inflight_batches", ctypes.c_uint32)]


class _JobResult(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("status", ctypes.c_int32), ("queue_ms", ctypes.c_double),
                ("upload_ms", ctypes.c_double), ("batch_wait_ms", ctypes.c_double), ("latency_ms", ctypes.c_double),
                ("batch_frames", ctypes.c_uint32)]


class _ServerStats(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32)] + [(n, ctypes.c_uint64) for n in (
        "jobs_submitted", "jobs_done", "jobs_failed", "batches", "batched_frames", "peak_device_bytes",
        "device_budget_bytes")]


class GpuOp:
    """GpuOp (run.rs:219-225): Develop, Unpack (the window as u16 samples, packed payloads unpacked,
    integer samples linearised) or Normalize (max((v - black) * scale, 0) as f32)."""

    @dataclass
    class Develop:
        spec: DevelopSpec

    class Unpack:
        pass

    class Normalize:
        pass


class PackOrder(enum.IntEnum):
    """PackOrder (crates/flashraw-core/src/decoder.rs:203-212)."""
    MsbBytes = 0
    LsbBytes = 1
    Msb16 = 2
    Msb32 = 3


def _device_ptr(t):
    """Device pointer of a torch tensor (or an int address)."""
    if isinstance(t, int):
        return t
    return t.data_ptr()


@dataclass
class GpuFrame:
    """GpuFrame (run.rs:27-93). `data` is a numpy array (host) or a torch tensor / int device
    address (`on_device=True`, resident input): u16 or f32 samples (`input` "u16" / "f32"), or a
    packed payload of bytes (`input` "packed", with `bits`, `order` and `row_stride`)."""
    data: object
    raster_width: int
    raster_height: int
    cpp: int = 1
    window: Optional[Rect] = None
    origin: tuple = (0, 0)
    on_device: bool = False
    linearization: object = None
    input: str = "u16"
    bits: int = 0
    order: PackOrder = PackOrder.MsbBytes
    row_stride: int = 0
    data_len: Optional[int] = None

    @staticmethod
    def from_u16(data, width: int, height: int, cpp: int = 1) -> "GpuFrame":
        on_dev = not isinstance(data, np.ndarray)
        return GpuFrame(data, width, height, cpp, Rect.full(width, height), (0, 0), on_dev)

    @staticmethod
    def from_f32(data, width: int, height: int, cpp: int = 1) -> "GpuFrame":
        on_dev = not isinstance(data, np.ndarray)
        return GpuFrame(data, width, height, cpp, Rect.full(width, height), (0, 0), on_dev, input="f32")

    @staticmethod
    def from_packed(data, width: int, height: int, bits: int, order=PackOrder.MsbBytes, row_stride=None,
                    data_len=None) -> "GpuFrame":
        """GpuFrame::from_packed (run.rs:43-53): a single-component payload of `bits`-bit samples;
        rows `row_stride` bytes apart (tight by default). `data_len` (bytes) is needed for device
        payloads given as an address."""
        on_dev = not isinstance(data, np.ndarray)
        stride = row_stride if row_stride is not None else (width * bits + 7) // 8
        return GpuFrame(data, width, height, 1, Rect.full(width, height), (0, 0), on_dev, input="packed", bits=bits,
                        order=PackOrder(order), row_stride=stride, data_len=data_len)

    def with_window(self, r: Rect) -> "GpuFrame":
        w = self.window or Rect.full(self.raster_width, self.raster_height)
        self.origin = (self.origin[0] + r.x - w.x, self.origin[1] + r.y - w.y)
        self.window = r
        return self

    def with_origin(self, origin) -> "GpuFrame":
        self.origin = tuple(origin)
        return self

    def with_linearization(self, lut) -> "GpuFrame":
        """GpuFrame::with_linearization (run.rs:84-87): integer samples v -> lut[min(v, len - 1)];
        None or an empty table for none."""
        self.linearization = None if lut is None or len(lut) == 0 else np.ascontiguousarray(lut, dtype=np.uint16)
        return self

    def _c(self):
        f = _Frame()
        f.size = ctypes.sizeof(_Frame)
        f.input = {"u16": 0, "f32": 1, "packed": 2}[self.input]
        keep = []
        if self.input == "packed":
            dtype = np.uint8
            n = self.data_len
        else:
            dtype = np.float32 if self.input == "f32" else np.uint16
            n = self.raster_width * self.raster_height * self.cpp
        if self.on_device:
            f.data = _device_ptr(self.data)
            f.data_on_device = 1
            if n is None:
                n = self.data.numel() * self.data.element_size()
            f.data_len = n
        else:
            a = np.ascontiguousarray(self.data)
            if a.dtype != dtype:
                a = a.view(dtype) if a.dtype.itemsize == np.dtype(dtype).itemsize else a.astype(dtype)
            keep.append(a)
            f.data = a.ctypes.data
            f.data_len = a.size
        f.raster_width, f.raster_height, f.cpp = self.raster_width, self.raster_height, self.cpp
        w = self.window or Rect.full(self.raster_width, self.raster_height)
        f.window = _Rect(w.x, w.y, w.width, w.height)
        f.origin_x, f.origin_y = self.origin
        f.bits, f.order, f.row_stride = self.bits, int(self.order), self.row_stride
        if self.linearization is not None:
            lut = np.ascontiguousarray(self.linearization, dtype=np.uint16)
            keep.append(lut)
            f.linearization, f.linearization_len = lut.ctypes.data, lut.size
        return f, keep


@dataclass
class GpuTimings:
    """GpuTimings (run.rs:229-239)."""
    upload_ms: float
    compute_ms: float
    kernel_ms: Optional[float]
    bands: int


@dataclass
class PlaneSpec:
    """PlaneSpec (crates/flashraw-gpu/src/denoise.rs:44-57): how one plane of the noise
    reduction is denoised."""
    levels: int
    gain: float = 1.0
    steps: Sequence[int] = (2, 2)
    hard_only0: bool = False
    keep: float = 0.0


@dataclass
class DenoiseSpec:
    """DenoiseSpec (denoise.rs:59-80) of GpuContext.denoise: the luma and colour planes (None
    leaves a plane unchanged), the guided filter's radius at half resolution, the hard threshold,
    the noise estimate's block budget and the output encoding. The noise models are measured and
    computed by the library (the reference's `models` callback, flashraw::denoise::gpu_models)."""
    luma: Optional[PlaneSpec] = None
    chroma: Optional[PlaneSpec] = None
    chroma_radius: int = 1
    hard: float = 2.7
    max_blocks: int = 50000
    output: "DevelopOutput" = None
    orientation: "Orientation" = None
    tone: Optional[GpuTone] = None
    curve: Optional[GpuCurve] = None
    cancel: Optional[CancelToken] = None

    def _c(self):
        s = _NrSpec()
        s.size = ctypes.sizeof(_NrSpec)
        for name, p in (("luma", self.luma), ("chroma", self.chroma)):
            setattr(s, "has_" + name, int(p is not None))
            if p is not None:
                c = getattr(s, name)
                c.levels, c.gain, c.hard_only0, c.keep = int(p.levels), float(p.gain), int(p.hard_only0), float(p.keep)
                c.steps[0], c.steps[1] = int(p.steps[0]), int(p.steps[1])
        s.chroma_radius, s.hard, s.max_blocks = int(self.chroma_radius), float(self.hard), int(self.max_blocks)
        s.output = int(DevelopOutput.LinearRgbF32 if self.output is None else self.output)
        s.orientation = int(Orientation.Normal if self.orientation is None else self.orientation)
        keep = []
        s.tone, s.curve = _tone_curve_c(self.tone, self.curve, keep)
        if self.cancel is not None:
            s.cancel = self.cancel._h
            keep.append(self.cancel)
        return s, keep


def denoise_spec(width: int, height: int, luma: float, chroma: float, output=None) -> DenoiseSpec:
    """flashraw::denoise::gpu_spec (crates/flashraw/src/denoise/mod.rs:155-193): the spec of
    strengths `luma` and `chroma` (clamped to [0, 1]; 0 leaves that part alone) for a
    width x height image, from the library (frh_nr_spec_for)."""
    s = _NrSpec()
    out = DevelopOutput.LinearRgbF32 if output is None else output
    check(lib().frh_nr_spec_for(width, height, luma, chroma, int(out), ctypes.byref(s)))

    def plane(c):
        return PlaneSpec(c.levels, c.gain, (c.steps[0], c.steps[1]), bool(c.hard_only0), c.keep)

    return DenoiseSpec(plane(s.luma) if s.has_luma else None, plane(s.chroma) if s.has_chroma else None,
                       s.chroma_radius, s.hard, s.max_blocks, DevelopOutput(s.output), Orientation(s.orientation))


def nr_model_fit(ctx: "GpuContext", stats) -> Optional[tuple]:
    """nr_model_from_stats fitted on the context's device, as GpuContext.denoise fits its levels
    (frh_nr_model_fit)."""
    st = np.ascontiguousarray(stats, dtype=np.float32).reshape(-1, 64)
    bins = ctypes.c_uint32(0)
    means = np.zeros(12, np.float32)
    sig = np.zeros((12, 64), np.float32)
    rc = lib().frh_nr_model_fit(ctx._h, st.ctypes.data, st.shape[0], ctypes.byref(bins), means.ctypes.data,
                                sig.ctypes.data)
    if rc == 9:
        return None
    check(rc)
    return means[:bins.value].copy(), sig[:bins.value].copy()


def nr_model_from_stats(stats) -> Optional[tuple]:
    """Model::from_stats (mod.rs:803-848) in the library: (means, sigmas) of the measured blocks
    (n, 64) (entry 0 the luma mean, then the squared coefficients), or None below 32 blocks."""
    st = np.ascontiguousarray(stats, dtype=np.float32).reshape(-1, 64)
    bins = ctypes.c_uint32(0)
    means = np.zeros(12, np.float32)
    sig = np.zeros((12, 64), np.float32)
    rc = lib().frh_nr_model_from_stats(st.ctypes.data, st.shape[0], ctypes.byref(bins), means.ctypes.data,
                                       sig.ctypes.data)
    if rc == 9:
        return None
    check(rc)
    return means[:bins.value].copy(), sig[:bins.value].copy()


_FMT_DTYPE = {0: np.uint8, 1: np.float32, 2: np.uint16}


class GpuImage:
    """A device-resident develop result (image.rs:52-306); freed with the object."""

    def __init__(self, handle: int):
        self._h = handle

    def width(self) -> int:
        return lib().frh_image_width(self._h)

    def height(self) -> int:
        return lib().frh_image_height(self._h)

    def channels(self) -> int:
        return lib().frh_image_channels(self._h)

    def device_ptr(self) -> int:
        return lib().frh_image_device_ptr(self._h)

    def download(self) -> np.ndarray:
        """GpuImage::download (image.rs:171-214): (height, width, channels) host array."""
        L = lib()
        dt = _FMT_DTYPE[L.frh_image_format(self._h)]
        out = np.empty((self.height(), self.width(), self.channels()), dtype=dt)
        check(L.frh_image_download(self._h, out.ctypes.data, out.nbytes))
        return out

    def download_into(self, out: np.ndarray) -> None:
        """Copy the pixels into an existing host array (pinned memory copies fastest)."""
        n = lib().frh_image_bytes(self._h)
        if out.nbytes < n:
            raise InvalidArgument(10, "output array too small")
        check(lib().frh_image_download(self._h, out.ctypes.data, n))

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().frh_image_free(self._h)
            except Exception:
                pass
            self._h = None


@dataclass
class GpuRun:
    """GpuRun (run.rs:243-246); `image2` holds DevelopSpec.output2 when asked for."""
    image: GpuImage
    timings: GpuTimings
    image2: Optional[GpuImage] = None


class GpuContext:
    """GpuContext (context.rs:116-150): one HIP device, its stream, tables and buffer pool."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        check(lib().frh_context_create(device, ctypes.byref(h)))
        self._h = h.value

    @property
    def handle(self) -> int:
        return self._h

    def info(self) -> str:
        buf = ctypes.create_string_buffer(256)
        check(lib().frh_context_info(self._h, buf, 256))
        return buf.value.decode()

    def check_sqrt(self) -> int:
        """How many floats in [2^-96, 2^128) (and +-0) the kernels' square root rounds differently
        from IEEE sqrt on this device (frh_check_sqrt); 0 where the develop is exact."""
        n = ctypes.c_uint64()
        check(lib().frh_check_sqrt(self._h, ctypes.byref(n)))
        return int(n.value)

    def table_cache(self):
        """(entries, device bytes) of the context's per-frame table cache (frh_context_table_cache)."""
        n, b = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().frh_context_table_cache(self._h, ctypes.byref(n), ctypes.byref(b)))
        return int(n.value), int(b.value)

    def override_limits(self, max_binding_bytes: int) -> None:
        """context.rs:704: cap the bytes one band binds (forces banded runs); 0 resets."""
        check(lib().frh_set_limits(self._h, max_binding_bytes))

    def run(self, frame: GpuFrame, info: FrameInfo, op) -> GpuRun:
        """GpuContext::run (run.rs:426); blocking (run_blocking, run.rs:1076)."""
        if op is GpuOp.Unpack or isinstance(op, GpuOp.Unpack) or op is GpuOp.Normalize or \
                isinstance(op, GpuOp.Normalize):
            normalize = op is GpuOp.Normalize or isinstance(op, GpuOp.Normalize)
            f, k1 = frame._c()
            out = ctypes.c_void_p()
            t = _Timings()
            if normalize:
                i, k2 = info._c()
                check(lib().frh_run_normalize(self._h, ctypes.byref(f), ctypes.byref(i), ctypes.byref(out),
                                              ctypes.byref(t)))
            else:
                k2 = None
                check(lib().frh_run_unpack(self._h, ctypes.byref(f), ctypes.byref(out), ctypes.byref(t)))
            del k1, k2
            timings = GpuTimings(t.upload_ms, t.compute_ms, t.kernel_ms if t.has_kernel_ms else None, t.bands)
            return GpuRun(GpuImage(out.value), timings, None)
        if not isinstance(op, GpuOp.Develop):
            raise InvalidArgument(10, "unknown GpuOp")
        f, k1 = frame._c()
        i, k2 = info._c()
        s, k3 = op.spec._c()
        out, out2 = ctypes.c_void_p(), ctypes.c_void_p()
        t = _Timings()
        check(lib().frh_run_develop(self._h, ctypes.byref(f), ctypes.byref(i), ctypes.byref(s), ctypes.byref(out),
                                    ctypes.byref(out2), ctypes.byref(t)))
        del k1, k2, k3
        timings = GpuTimings(t.upload_ms, t.compute_ms, t.kernel_ms if t.has_kernel_ms else None, t.bands)
        return GpuRun(GpuImage(out.value), timings, GpuImage(out2.value) if out2.value else None)

    run_blocking = run

    def stream_handle(self, stream=None) -> int:
        """hipStream_t of `stream` (a torch stream or an address); None is the context's stream."""
        if stream is None:
            return lib().frh_context_stream(self._h) or 0
        return stream if isinstance(stream, int) else stream.cuda_stream

    def prepare(self, frame: GpuFrame, info: FrameInfo, spec: DevelopSpec) -> "Plan":
        """A validated develop to launch repeatedly on resident buffers (frh_plan_create)."""
        return Plan(self, frame, info, spec)

    def develop_into(self, frame: GpuFrame, info: FrameInfo, spec: DevelopSpec, dst, dst2=None, stream=None):
        """Resident develop into caller-owned device buffers (torch tensors or addresses) on
        `stream` (a torch stream, a raw hipStream_t address, or None for the context stream);
        returns without synchronising."""
        f, k1 = frame._c()
        i, k2 = info._c()
        s, k3 = spec._c()
        sp = self.stream_handle(stream)
        n1 = dst.numel() * dst.element_size() if hasattr(dst, "numel") else (1 << 62)
        n2 = 0 if dst2 is None else (dst2.numel() * dst2.element_size() if hasattr(dst2, "numel") else (1 << 62))
        check(lib().frh_develop_into(self._h, ctypes.byref(f), ctypes.byref(i), ctypes.byref(s),
                                     _device_ptr(dst), n1, None if dst2 is None else _device_ptr(dst2), n2, sp))
        del k1, k2, k3

    def unpack_into(self, frame: GpuFrame, dst, stream=None):
        """GpuOp::Unpack of a device-resident frame into the caller's device buffer (u16 samples)."""
        f, k1 = frame._c()
        n = dst.numel() * dst.element_size() if hasattr(dst, "numel") else (1 << 62)
        check(lib().frh_unpack_into(self._h, ctypes.byref(f), _device_ptr(dst), n, self.stream_handle(stream)))
        del k1

    def normalize_into(self, frame: GpuFrame, info: FrameInfo, dst, stream=None):
        """GpuOp::Normalize of a device-resident frame into the caller's device buffer (f32)."""
        f, k1 = frame._c()
        i, k2 = info._c()
        n = dst.numel() * dst.element_size() if hasattr(dst, "numel") else (1 << 62)
        check(lib().frh_normalize_into(self._h, ctypes.byref(f), ctypes.byref(i), _device_ptr(dst), n,
                                       self.stream_handle(stream)))
        del k1, k2

    def batch(self, jobs) -> "Batch":
        """A batched develop of `jobs`, a sequence of (frame, info, spec, dst): device-resident
        frames developed into the device buffers `dst` by one persistent launch per demosaic
        kind (frh_batch_create)."""
        return Batch(self, jobs)

    def batch_develop_noalloc(self, jobs, ahd_rows: int = 0, bilinear_rows: int = 0, stream=None) -> None:
        """The batched develop of `jobs` ((frame, info, spec, dst) as for batch()) in one call that
        allocates no device memory (frh_batch_develop_noalloc); asynchronous on `stream`."""
        n = len(jobs)
        frames, infos, specs = (_Frame * n)(), (_FrameInfo * n)(), (_DevelopSpec * n)()
        dsts = (ctypes.c_void_p * n)()
        keep = []
        for k, (frame, info, spec, dst) in enumerate(jobs):
            f, k1 = frame._c()
            i, k2 = info._c()
            s, k3 = spec._c()
            frames[k], infos[k], specs[k] = f, i, s
            dsts[k] = _device_ptr(dst)
            keep += [k1, k2, k3]
        check(lib().frh_batch_develop_noalloc(self._h, n, frames, infos, specs, dsts, ahd_rows, bilinear_rows,
                                              self.stream_handle(stream)))
        del keep

    def server(self, device_budget_bytes: int = 0, max_batch_frames: int = 0, max_inflight_batches: int = 0) -> "Server":
        """A serving loop on this context (frh_server_create)."""
        return Server(self, device_budget_bytes, max_batch_frames, max_inflight_batches)

    def denoise(self, image: "GpuImage", spec: DenoiseSpec) -> GpuRun:
        """GpuContext::denoise (crates/flashraw-gpu/src/denoise.rs:817): the luma / chroma noise
        reduction of a linear f32 RGB image (a develop to DevelopOutput.LinearRgbF32), encoded to
        spec.output; blocking (denoise_blocking, denoise.rs:1180)."""
        s, keep = spec._c()
        out = ctypes.c_void_p()
        t = _Timings()
        check(lib().frh_run_denoise(self._h, image._h, ctypes.byref(s), ctypes.byref(out), ctypes.byref(t)))
        del keep
        timings = GpuTimings(t.upload_ms, t.compute_ms, t.kernel_ms if t.has_kernel_ms else None, t.bands)
        return GpuRun(GpuImage(out.value), timings, None)

    denoise_blocking = denoise

    def denoise_into(self, rgb, width: int, height: int, spec: DenoiseSpec, dst, stream=None) -> None:
        """The noise reduction of device-resident linear RGB (width * height * 3 floats) into the
        caller's device buffer `dst` (torch tensors or addresses) on `stream`."""
        s, keep = spec._c()
        n = dst.numel() * dst.element_size() if hasattr(dst, "numel") else (1 << 62)
        check(lib().frh_denoise_into(self._h, _device_ptr(rgb), width, height, ctypes.byref(s), _device_ptr(dst), n,
                                     self.stream_handle(stream)))
        del keep

    def denoise_cfa(self, spec: CfaDenoiseSpec, src, dst=None, dst_f32=None, stream=None) -> None:
        """Denoise the device u16 mosaic `src` into `dst` (u16) and/or `dst_f32` (f32, DN above
        black), torch tensors or device addresses, on `stream` (frh_denoise_cfa)."""
        s = spec._c()
        check(lib().frh_denoise_cfa(self._h, ctypes.byref(s), _device_ptr(src),
                                    None if dst is None else _device_ptr(dst),
                                    None if dst_f32 is None else _device_ptr(dst_f32), self.stream_handle(stream)))

    def output_size(self, frame: GpuFrame, spec: DevelopSpec, output: DevelopOutput):
        """(width, height, channels, bytes) of `output` for this frame and spec."""
        f, k1 = frame._c()
        s, k3 = spec._c()
        w, h, c, b = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
        check(lib().frh_develop_output_size(ctypes.byref(f), ctypes.byref(s), int(output), ctypes.byref(w),
                                            ctypes.byref(h), ctypes.byref(c), ctypes.byref(b)))
        return w.value, h.value, c.value, b.value

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().frh_context_destroy(self._h)
            except Exception:
                pass
            self._h = None


def pinned_empty(shape, dtype) -> np.ndarray:
    """A numpy array in pinned host memory (frh_host_alloc), freed with the array."""
    dt = np.dtype(dtype)
    n = int(np.prod(shape)) * dt.itemsize
    ptr = lib().frh_host_alloc(max(n, 1))
    if not ptr:
        raise OutOfMemory(7, "pinned host allocation failed")
    buf = (ctypes.c_uint8 * max(n, 1)).from_address(ptr)
    arr = np.frombuffer(buf, dtype=np.uint8, count=n).view(dt).reshape(shape)
    import weakref
    weakref.finalize(buf, lib().frh_host_free, ptr)
    return arr


@dataclass
class JobResult:
    status: int
    queue_ms: float
    upload_ms: float
    batch_wait_ms: float
    latency_ms: float
    batch_frames: int


class Server:
    """frh_server: a request loop with device-memory admission and continuous batching."""

    def __init__(self, ctx: "GpuContext", device_budget_bytes: int = 0, max_batch_frames: int = 0,
                 max_inflight_batches: int = 0):
        self.ctx = ctx
        c = _ServerConfig(ctypes.sizeof(_ServerConfig), device_budget_bytes, max_batch_frames, max_inflight_batches)
        h = ctypes.c_void_p()
        check(lib().frh_server_create(ctx._h, ctypes.byref(c), ctypes.byref(h)))
        self._h = h.value
        self._keep = {}

    def submit(self, frame: GpuFrame, info: FrameInfo, spec: DevelopSpec, out: np.ndarray,
               denoise: Optional[CfaDenoiseSpec] = None) -> int:
        """Queue a job; `frame` holds host data, `out` receives the output (both must stay alive
        until the job is collected by wait())."""
        f, k1 = frame._c()
        i, k2 = info._c()
        sp, k3 = spec._c()
        d = denoise._c() if denoise is not None else None
        t = ctypes.c_uint64()
        check(lib().frh_server_submit(self._h, ctypes.byref(f), ctypes.byref(i), ctypes.byref(sp),
                                      None if d is None else ctypes.byref(d), out.ctypes.data, out.nbytes,
                                      ctypes.byref(t)))
        self._keep[t.value] = (frame, out)
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1) -> Optional[JobResult]:
        r = _JobResult()
        r.size = ctypes.sizeof(_JobResult)
        done = ctypes.c_int32()
        check(lib().frh_server_wait(self._h, ticket, timeout_ms, ctypes.byref(r), ctypes.byref(done)))
        if not done.value:
            return None
        self._keep.pop(ticket, None)
        return JobResult(r.status, r.queue_ms, r.upload_ms, r.batch_wait_ms, r.latency_ms, r.batch_frames)

    def stats(self) -> dict:
        st = _ServerStats()
        st.size = ctypes.sizeof(_ServerStats)
        check(lib().frh_server_get_stats(self._h, ctypes.byref(st)))
        return {n: int(getattr(st, n)) for n, _ in _ServerStats._fields_[1:]}

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().frh_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """frh_batch: launch, capture into a HIP graph, autotune the strip heights."""

    def __init__(self, ctx: "GpuContext", jobs):
        self.ctx = ctx
        n = len(jobs)
        frames = (_Frame * n)()
        infos = (_FrameInfo * n)()
        specs = (_DevelopSpec * n)()
        dsts = (ctypes.c_void_p * n)()
        self._keep = []
        for k, (frame, info, spec, dst) in enumerate(jobs):
            f, k1 = frame._c()
            i, k2 = info._c()
            s, k3 = spec._c()
            frames[k], infos[k], specs[k] = f, i, s
            dsts[k] = _device_ptr(dst)
            self._keep += [k1, k2, k3, dst, frame]
        h = ctypes.c_void_p()
        check(lib().frh_batch_create(ctx._h, n, frames, infos, specs, dsts, ctypes.byref(h)))
        self._h = h.value

    def launch(self, stream=None) -> None:
        check(lib().frh_batch_launch(self._h, self.ctx.stream_handle(stream)))

    def capture(self, stream) -> None:
        """Capture the launch sequence on `stream` (a torch stream) into a HIP graph."""
        check(lib().frh_batch_capture(self._h, self.ctx.stream_handle(stream)))

    def autotune(self, stream=None, iters: int = 2):
        a, b = ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().frh_batch_autotune(self._h, self.ctx.stream_handle(stream), iters, ctypes.byref(a),
                                       ctypes.byref(b)))
        return a.value, b.value

    def set_strip_rows(self, ahd: int = 0, bilinear: int = 0) -> None:
        check(lib().frh_batch_set_strip_rows(self._h, ahd, bilinear))

    @property
    def items(self) -> int:
        return int(lib().frh_batch_items(self._h))

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().frh_batch_destroy(self._h)
            except Exception:
                pass
            self._h = None


class Plan:
    """A prepared develop (frh_plan): one ctypes call per launch."""

    def __init__(self, ctx: GpuContext, frame: GpuFrame, info: FrameInfo, spec: DevelopSpec):
        f, k1 = frame._c()
        i, k2 = info._c()
        s, k3 = spec._c()
        h = ctypes.c_void_p()
        check(lib().frh_plan_create(ctx.handle, ctypes.byref(f), ctypes.byref(i), ctypes.byref(s), ctypes.byref(h)))
        del k1, k2, k3
        self._h = h.value
        self.ctx = ctx
        self.output_bytes = (lib().frh_plan_output_bytes(self._h, 0), lib().frh_plan_output_bytes(self._h, 1))
        self._launch = lib().frh_plan_launch

    def launch(self, src, dst, dst2=None, stream=None) -> None:
        """Develop device input `src` into `dst` (and `dst2`) on `stream`, asynchronously."""
        rc = self._launch(self._h, _device_ptr(src), _device_ptr(dst), None if dst2 is None else _device_ptr(dst2),
                          self.ctx.stream_handle(stream))
        if rc:
            check(rc)

    def autotune(self, src, dst, dst2=None, stream=None, iters: int = 3) -> int:
        """Time the candidate tile shapes (strip height x waves per block) on these resident buffers
        and keep the fastest (frh_plan_autotune); returns its strip height (tile_shape has both)."""
        best = ctypes.c_uint32()
        check(lib().frh_plan_autotune(self._h, _device_ptr(src), _device_ptr(dst),
                                      None if dst2 is None else _device_ptr(dst2), self.ctx.stream_handle(stream),
                                      iters, ctypes.byref(best)))
        return best.value

    @property
    def tile_shape(self) -> tuple:
        """(strip height, waves per block, frame columns per lane) of the plan's row-walker launch,
        (0, 0, 0) for the kernels without one (frh_plan_tile_shape)."""
        rows, wpb, cpl = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().frh_plan_tile_shape(self._h, ctypes.byref(rows), ctypes.byref(wpb), ctypes.byref(cpl)))
        return rows.value, wpb.value, cpl.value

    KERNELS = ("bilinear_walk", "bilinear_tiled", "ahd_walk", "ahd_tiled", "edge_tame", "edge_general", "generic")

    @property
    def kernel(self) -> str:
        """The kernel family the plan's launches take (frh_plan_kernel)."""
        k = lib().frh_plan_kernel(self._h)
        if k < 0 or k >= len(self.KERNELS):
            check(k)
        return self.KERNELS[k]

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().frh_plan_destroy(self._h)
            except Exception:
                pass
            self._h = None


_ctx = None
_ctx_lock = threading.Lock()


def available() -> bool:
    """flashraw_gpu::available (context.rs:775-842)."""
    return bool(lib().frh_available())


def context() -> GpuContext:
    """flashraw_gpu::context(): the process-wide context on device 0."""
    global _ctx
    with _ctx_lock:
        if _ctx is None:
            _ctx = GpuContext(0)
        return _ctx


def srgb8_table() -> np.ndarray:
    """The 4097-entry sRGB table every develop encodes with."""
    t = np.zeros(4097, dtype=np.uint8)
    check(lib().frh_srgb8_table(t.ctypes.data, t.size))
    return t
