"""Pins the CPU oracle (oracle/liboracle.so) before it is trusted as the GPU path's checker.

Each test checks the oracle against something that does not share its code: the reference's own
known-answer tests and case matrices restated (file:line cited), a second per-pixel statement of
the same arithmetic in numpy float32 (tests/np_restatement.py), closed forms and invariants, and
the committed golden fixtures. CPU only.
"""
import math

import numpy as np
import pytest

import np_restatement as N
from oracle import oracle as O

F32 = np.float32
XTRANS = [[1, 1, 0, 1, 1, 2], [1, 1, 2, 1, 1, 0], [2, 0, 1, 0, 2, 1], [1, 1, 2, 1, 1, 0], [1, 1, 0, 1, 1, 2],
          [0, 2, 1, 2, 0, 1]]
BAYERS = ["RGGB", "GBRG", "GRBG", "BGGR"]
MATRIX = ((1.62, -0.38, -0.24), (-0.21, 1.58, -0.37), (-0.12, -0.45, 1.57))
WB = (2.1, 1.0, 1.6, 1.0)


def test_xtrans_restated_pattern():
    assert N.parse_cfa(N.XTRANS).tolist() == XTRANS


# ---------------------------------------------------------------- bilinear


@pytest.mark.parametrize("cfa", BAYERS + ["XTRANS"])
@pytest.mark.parametrize("origin", [(0, 0), (1, 0), (0, 1), (3, 5)])
def test_bilinear_matches_naive_reference(cfa, origin):
    """bilinear_matches_naive_reference (process.rs:2448-2485): 29x17 LCG-filled frame, black
    pattern [60, 64, 62, 66], white 4095, identity WB and matrix; every channel within 1e-5 of
    the naive same-colour mean, and border pixels (where the reference sums row-major like the
    naive mean) exactly equal to it."""
    w, h = 29, 17
    mos = N.lcg_fill(w, h)
    pat = XTRANS if cfa == "XTRANS" else cfa
    black = (2, 2, 1, [60.0, 64.0, 62.0, 66.0])
    got = O.develop(mos, pat, origin=origin, black=black, white=(4095,), encoding="linear_f32")
    norm = N.normalise(mos, pat, origin, black, (4095,))
    want = np.clip(N.naive_bilinear(norm, N.colours(pat, w, h, origin)), 0.0, 1.0)
    assert np.abs(got - want).max() < 1e-5
    border = np.ones((h, w), dtype=bool)
    border[1:-1, 1:-1] = False
    assert np.array_equal(got[border], want[border])


@pytest.mark.parametrize("cfa", BAYERS + ["XTRANS"])
@pytest.mark.parametrize("origin", [(0, 0), (1, 1), (3, 2)])
def test_bilinear_equals_numpy_statement_exactly(cfa, origin):
    """The whole fused path (calibration with row/col deltas, demosaic, WB, clip, matrix,
    levels) against the numpy statement, bit for bit."""
    w, h = 37, 23
    mos = N.lcg_stream(w, h, seed=99)
    pat = XTRANS if cfa == "XTRANS" else cfa
    black = (2, 2, 1, [60.0, 64.0, 62.0, 66.0])
    rd = np.linspace(-3, 3, h + origin[1]).astype(F32)
    cd = np.linspace(2, -2, w + origin[0] - 4).astype(F32)  # shorter than the frame: missing entries add 0
    for levels in (None, (1.25, 0.05)):
        got = O.develop(mos, pat, origin=origin, black=black, white=(4095,), wb=WB, matrix=MATRIX,
                        encoding="linear_f32", levels=levels, row_deltas=rd, col_deltas=cd)
        norm = N.normalise(mos, pat, origin, black, (4095,), rd, cd)
        want = N.finish(N.bilinear(norm, N.colours(pat, w, h, origin), bayer=cfa != "XTRANS"), WB, MATRIX, levels)
        assert np.array_equal(got, want), levels


def test_bilinear_flat_field():
    """bilinear_flat_field (process.rs:2220-2235): flat 1000, black 0, white 4000 -> 0.25."""
    mos = np.full((8, 8), 1000, dtype=np.uint16)
    got = O.develop(mos, "RGGB", black=(1, 1, 1, [0.0]), white=(4000,), encoding="linear_f32")
    assert np.all(np.abs(got - 0.25) < 1e-5)


def test_per_colour_white_levels():
    """White level per colour (WhiteLevel values indexed by CFA colour, process.rs:276-281)."""
    mos = np.full((10, 12), 3000, dtype=np.uint16)
    got = O.develop(mos, "RGGB", black=(1, 1, 1, [100.0]), white=(4000, 3500, 3100), encoding="linear_f32")
    norm = N.normalise(mos, "RGGB", (0, 0), (1, 1, 1, [100.0]), (4000, 3500, 3100))
    assert np.array_equal(got, N.finish(N.bilinear(norm, N.colours("RGGB", 12, 10)), (1, 1, 1), np.eye(3)))
    assert abs(float(got[5, 5, 0]) - 2900 / 3900) < 1e-6 and abs(float(got[5, 5, 2]) - 2900 / 3000) < 1e-6


# ---------------------------------------------------------------- AHD


def _ahd_case(w, h, cfa, origin, seed, black=(1, 1, 1, [64.0]), white=(4095,)):
    mos = N.lcg_stream(w, h, seed=seed)
    got = O.develop(mos, cfa, origin=origin, black=black, white=white, wb=WB, matrix=MATRIX, demosaic="ahd",
                    encoding="linear_f32")
    want = N.develop(mos, cfa, origin=origin, black=black, white=white, wb=WB, matrix=MATRIX, demosaic="ahd")
    return got, want


@pytest.mark.parametrize("cfa", BAYERS)
@pytest.mark.parametrize("origin", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_ahd_equals_per_pixel_statement(cfa, origin):
    """The oracle's band/chunk AHD (ahd.rs structure) against the per-pixel numpy statement of
    the same method, bit for bit, on noisy LCG data where every homogeneity decision matters."""
    got, want = _ahd_case(61, 45, cfa, origin, seed=7 + origin[0] * 2 + origin[1])
    assert np.array_equal(got, want)


def test_ahd_equals_statement_across_chunks_and_bands():
    """A frame wider than one 192-column chunk and taller than one 64-row band."""
    mos = N.lcg_stream(430, 150, seed=12345)
    got = O.develop(mos, "RGGB", origin=(1, 0), black=(1, 1, 1, [64.0]), white=(4095,), wb=WB, matrix=MATRIX,
                    demosaic="ahd", encoding="linear_f32")
    want = N.develop(mos, "RGGB", origin=(1, 0), black=(1, 1, 1, [64.0]), white=(4095,), wb=WB, matrix=MATRIX,
                     demosaic="ahd")
    assert np.array_equal(got, want)


def test_ahd_planted_mistake_is_caught():
    """A plausible slip (the homogeneity candidates swapped in the blend) must fail the
    comparison: the per-pixel statement is sensitive to the decisions the test relies on."""
    w, h = 61, 45
    mos = N.lcg_stream(w, h, seed=3)
    got = O.develop(mos, "RGGB", black=(1, 1, 1, [64.0]), white=(4095,), demosaic="ahd", encoding="linear_f32")
    norm = N.normalise(mos, "RGGB", (0, 0), (1, 1, 1, [64.0]), (4095,))
    col = N.colours("RGGB", w, h)
    a = N.ahd(norm, col, (1, 1, 1, 1))
    # the same statement with the row/column candidates exchanged
    swapped = N.ahd(norm.T.copy(), col.T.copy(), (1, 1, 1, 1)).transpose(1, 0, 2)
    s = slice(N.HALO, -N.HALO)
    ok = np.clip(a[s, s], 0, 1)
    assert np.array_equal(got[s, s], ok)
    assert not np.array_equal(got[s, s], np.clip(swapped[s, s], 0, 1))


def test_ahd_smooth_scene_is_reproduced():
    """flat_and_smooth_fields_are_reproduced (ahd.rs:497-514): the smooth scene with one hard
    vertical edge, mosaicked RGGB, is recovered within 0.02 left of the edge."""
    w, h = 80, 70
    x = np.arange(w, dtype=F32)[None, :] / F32(w)
    y = np.arange(h, dtype=F32)[:, None] / F32(h)
    edge = np.where(np.arange(w)[None, :] > w // 2, F32(0.5), F32(0.0))
    rgb = np.stack(np.broadcast_arrays(F32(0.2) + F32(0.3) * x + edge, F32(0.3) + F32(0.2) * y + edge * F32(0.5),
                                       F32(0.6) - F32(0.3) * x), axis=-1).astype(F32)
    col = N.colours("RGGB", w, h)
    m = np.take_along_axis(rgb, col[..., None], axis=-1)[..., 0]
    # a u16 mosaic whose normalised values are m (white 65535, black 0) to within 1 / 65535
    mos = np.round(m.astype(np.float64) * 65535).astype(np.uint16)
    got = O.develop(mos, "RGGB", black=(1, 1, 1, [0.0]), white=(65535,), demosaic="ahd", encoding="linear_f32")
    s = (slice(N.HALO, h - N.HALO), slice(N.HALO, w // 2 - 4))
    assert np.abs(got[s] - np.clip(rgb[s], 0, 1)).max() < 0.02


def test_ahd_crops_match_the_whole_frame():
    """bands_chunks_and_crops_agree_with_the_whole_frame (ahd.rs:516-538): a crop develops the
    same pixels as the whole frame where both use AHD."""
    w, h = 430, 61
    mos = N.lcg_stream(w, h, seed=7)
    for origin in [(0, 0), (1, 0), (0, 1), (1, 1)]:
        kw = dict(origin=origin, black=(1, 1, 1, [64.0]), white=(4095,), demosaic="ahd", encoding="linear_f32")
        whole = O.develop(mos, "RGGB", **kw)
        part = O.develop(mos, "RGGB", crop=(13, 9, 287, 11), **kw)
        assert np.array_equal(part, whole[9:20, 13:300]), origin


def test_ahd_border_is_bilinear_and_interior_differs():
    """Pixels within 6 samples of the frame edge use bilinear (process.rs:1185-1199); on a
    sharp scene the AHD interior differs from bilinear (tests/ahd.rs:95-99)."""
    mos = N.lcg_stream(50, 40, seed=5)
    kw = dict(black=(1, 1, 1, [64.0]), white=(4095,), wb=WB, matrix=MATRIX, encoding="linear_f32")
    a = O.develop(mos, "GBRG", demosaic="ahd", **kw)
    b = O.develop(mos, "GBRG", demosaic="bilinear", **kw)
    inner = np.zeros(a.shape[:2], dtype=bool)
    inner[N.HALO:-N.HALO, N.HALO:-N.HALO] = True
    assert np.array_equal(a[~inner], b[~inner])
    assert np.abs(a[inner] - b[inner]).max() > 0.01


# ---------------------------------------------------------------- encodings and orientation


@pytest.fixture(scope="module")
def lut8():
    return np.fromfile(str(__import__("pathlib").Path(__file__).parent / "golden" / "srgb8_table.bin"), dtype=np.uint8)


def test_srgb8_table_formula(lut8):
    """srgb_lut (process.rs:1903-1915): (s * 255 + 0.5) as u8 of the sRGB curve at i / 4096;
    endpoints and monotonicity (kernels.rs:620-627); every entry equals the float64 curve's
    value except where that value lies within 1e-3 of a rounding boundary."""
    assert lut8[0] == 0 and lut8[4096] == 255 and np.all(np.diff(lut8.astype(int)) >= 0)
    v = np.arange(4097) / 4096.0
    s = np.where(v <= 0.0031308, 12.92 * v, 1.055 * v ** (1 / 2.4) - 0.055) * 255 + 0.5
    clear = np.abs(s - np.round(s)) > 1e-3
    assert np.array_equal(lut8[clear], np.floor(s[clear]).astype(np.uint8))
    assert np.array_equal(O.srgb8_table(), lut8)


@pytest.mark.parametrize("demosaic", ["bilinear", "ahd"])
def test_display_encodings_follow_linear(demosaic, lut8):
    """RGBA8, RGB8 and linear 16-bit samples are the closed-form encodings of the linear f32
    develop (process.rs:1679-1728, 1971-1973)."""
    mos = N.lcg_stream(48, 40, seed=11)
    kw = dict(black=(1, 1, 1, [64.0]), white=(4095,), wb=WB, matrix=MATRIX, demosaic=demosaic)
    lin = O.develop(mos, "GRBG", encoding="linear_f32", **kw)
    assert np.array_equal(O.develop(mos, "GRBG", encoding="rgba8", **kw), N.rgba8(lin, lut8))
    assert np.array_equal(O.develop(mos, "GRBG", encoding="rgb8", **kw), N.rgba8(lin, lut8)[..., :3])
    assert np.array_equal(O.develop(mos, "GRBG", encoding="linear16", **kw), N.linear16(lin))
    f16 = O.develop(mos, "GRBG", encoding="linear_f16", **kw)
    assert np.array_equal(f16, lin.astype(np.float16).view(np.uint16))


def test_srgb16_curve_accuracy():
    """srgb16_curve_is_accurate_and_monotonic (process.rs:2701-2729): monotone, within one code
    of the float64 curve, off by one on under 0.5 % of samples, endpoints 0 and 65535."""
    n = 200_000
    v = (np.arange(n + 1) / n).astype(F32)
    mos = np.zeros((1, 1), dtype=np.uint16)  # the encoder is reached through develop: use a ramp frame
    ramp = np.round(v.astype(np.float64) * 65535).astype(np.uint16)
    w = 1000
    rows = ramp[: (ramp.size // w) * w].reshape(-1, w)
    lin = O.develop(rows, [[1]], black=(1, 1, 1, [0.0]), white=(65535,), encoding="linear_f32")
    s16 = O.develop(rows, [[1]], black=(1, 1, 1, [0.0]), white=(65535,), encoding="srgb16")
    x = lin[..., 1].astype(np.float64).reshape(-1)
    got = s16[..., 1].astype(np.int64).reshape(-1)
    exact = np.round(np.where(x <= 0.0031308, 12.92 * x, 1.055 * x ** (1 / 2.4) - 0.055) * 65535).astype(np.int64)
    assert np.all(np.diff(got) >= 0)
    d = np.abs(got - exact)
    assert d.max() <= 1 and (d != 0).sum() < got.size / 200
    assert got[0] == 0 and got[-1] == 65535
    del mos


def _half_ref(x):
    """Nearest half by bisection over bit patterns, saturating at 0x7bff
    (half_floats_round_to_nearest_even_and_saturate, process.rs:2731-2750)."""
    s = 0x8000 if math.copysign(1.0, x) < 0 else 0
    a = abs(float(x))
    if a >= 65520.0:
        return s | 0x7BFF
    h2f = lambda b: float(np.array([b], dtype=np.uint16).view(np.float16)[0])
    lo, hi = 0, 0x7BFF
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if h2f(mid) <= a:
            lo = mid
        else:
            hi = mid
    dl, dh = a - h2f(lo), h2f(hi) - a
    return s | (lo if dl < dh or (dl == dh and lo & 1 == 0) else hi)


def test_half_floats_round_to_nearest_even_and_saturate():
    """f32_to_f16 (process.rs:1978-1991) against the reference test's bisection statement on its
    xorshift stream and special values, and round trips of every finite half."""
    x = 1
    vals = []
    for i in range(4000):
        x ^= (x << 13) & 0xFFFFFFFF
        x ^= x >> 17
        x ^= (x << 5) & 0xFFFFFFFF
        v = np.array([x], dtype=np.uint32).view(np.float32)[0] if i % 2 == 0 else F32(x / 4294967295.0 * 4.0 - 1.0)
        if np.isfinite(v):
            vals.append(F32(v))
    vals += [F32(v) for v in (0.0, 1.0, -1.0, 0.5, 65504.0, 65519.0, 65520.0, 1e9, -1e9, 5.9604645e-8, 2.9e-8, 6.1e-5)]
    for v in vals:
        assert O.f32_to_f16(v) == _half_ref(v), v
    assert O.f32_to_f16(float("nan")) & 0x7E00 == 0x7E00
    halves = np.arange(0x10000, dtype=np.uint32).astype(np.uint16)
    finite = ((halves >> 10) & 0x1F) != 31
    for hbits in halves[finite][::7]:
        assert O.f32_to_f16(float(np.array([hbits], dtype=np.uint16).view(np.float16)[0])) == int(hbits)


def test_orientation_rotation_known_answer():
    """orientation_rotation (process.rs:2564-2572): a row image, red left and blue right,
    rotated 90 degrees clockwise is a column with red on top. Stated through develop on a 3x1
    frame with a red, green, blue pattern row, whose end pixels develop to pure red and blue."""
    mos = np.array([[4095, 0, 4095]], dtype=np.uint16)
    img = O.develop(mos, [[0, 1, 2]], black=(1, 1, 1, [0.0]), white=(4095,), orientation=6, encoding="rgba8")
    assert img.shape == (3, 1, 4)
    assert img[0, 0, :3].tolist() == [255, 0, 0] and img[2, 0, :3].tolist() == [0, 0, 255]


@pytest.mark.parametrize("demosaic", ["bilinear", "ahd"])
def test_all_orientations(demosaic):
    """Display outputs (RGBA8, sRGB16) take EXIF orientation 1..8 (process.rs:2035-2044, EXIF
    definitions restated with numpy flips and rotations); linear outputs keep the sensor's
    (process.rs:783-787)."""
    mos = N.lcg_stream(37, 23, seed=1)
    kw = dict(origin=(3, 1), black=(1, 1, 1, [64.0]), white=(4095,), wb=WB, matrix=MATRIX, demosaic=demosaic,
              crop=(1, 2, 33, 19))
    base = O.develop(mos, "BGGR", encoding="rgba8", **kw)
    base16 = O.develop(mos, "BGGR", encoding="srgb16", **kw)
    lin = O.develop(mos, "BGGR", encoding="linear16", **kw)
    for o in range(1, 9):
        assert np.array_equal(O.develop(mos, "BGGR", encoding="rgba8", orientation=o, **kw), N.orient(base, o)), o
        assert np.array_equal(O.develop(mos, "BGGR", encoding="srgb16", orientation=o, **kw), N.orient(base16, o))
        assert np.array_equal(O.develop(mos, "BGGR", encoding="linear16", orientation=o, **kw), lin)


def test_crop_equals_whole_frame_region():
    """A bilinear crop develops the frame's pixels (neighbours outside the crop but inside the
    frame are used, process.rs:1045-1060)."""
    mos = N.lcg_stream(40, 30, seed=8)
    kw = dict(black=(1, 1, 1, [64.0]), white=(4095,), wb=WB, matrix=MATRIX, encoding="linear_f32")
    whole = O.develop(mos, "RGGB", **kw)
    part = O.develop(mos, "RGGB", crop=(3, 5, 20, 17), **kw)
    assert np.array_equal(part, whole[5:22, 3:23])


# ---------------------------------------------------------------- DCT block denoiser


def test_dct_matrix_is_orthonormal():
    """dct_is_orthonormal_and_inverts (denoise/mod.rs:1149-1163) for the closed form."""
    D = N.dct_matrix()
    assert np.abs(D @ D.T - np.eye(8)).max() < 1e-14
    from scipy.fft import dct
    assert np.abs(dct(np.eye(8), type=2, norm="ortho", axis=0) - D).max() < 1e-14


def _dn_frame(w=48, h=40, seed=4):
    rng = np.random.default_rng(seed)
    base = 2000 + 800 * np.sin(np.arange(w)[None, :] / 5.0) * np.cos(np.arange(h)[:, None] / 7.0)
    return np.clip(base + rng.normal(0, 40, (h, w)), 0, 16383).astype(np.uint16)


def test_dct_denoise_matches_scipy_statement():
    """fp64 oracle against the scipy-DCT statement (threshold 2.7 sigma, stride 4, uniform
    overlap-add, last positions included)."""
    mos = _dn_frame()
    black = [512.0, 510.0, 514.0, 511.0]
    of, ou = O.dct_denoise(mos, black, 2.0, 144.0, 2.7, 4, precision=64)
    want = N.dct_denoise(mos, black, 2.0, 144.0, 2.7, 4)
    assert np.abs(of.astype(np.float64) - want).max() < 1e-3
    u = np.clip(np.floor(want + np.tile(np.array(black).reshape(2, 2), (20, 24)) + 0.5), 0, 65535)
    assert np.abs(ou.astype(np.int64) - u.astype(np.int64)).max() <= 1


@pytest.mark.parametrize("step", [2, 3, 4, 8])
def test_dct_denoise_identity_and_block_mean(step):
    """Threshold 0 keeps every coefficient: the planes come back unchanged. An infinite threshold
    keeps only DC: every block becomes its mean, overlap-added uniformly."""
    mos = _dn_frame(40, 36, seed=step)
    black = [500.0, 500.0, 500.0, 500.0]
    of, _ = O.dct_denoise(mos, black, 1.0, 1.0, 0.0, step, precision=64)
    assert np.abs(of - (mos.astype(np.float64) - 500.0)).max() < 1e-9
    of, _ = O.dct_denoise(mos, black, 1.0, 1.0, 1e30, step, precision=64)
    for plane in range(4):
        p = mos[plane >> 1::2, plane & 1::2].astype(np.float64) - 500.0
        ph, pw = p.shape
        pos = lambda n: sorted(set(list(range(0, n - 7, step)) + [n - 8]))
        acc = np.zeros_like(p)
        cnt = np.zeros_like(p)
        for y0 in pos(ph):
            for x0 in pos(pw):
                acc[y0:y0 + 8, x0:x0 + 8] += p[y0:y0 + 8, x0:x0 + 8].mean()
                cnt[y0:y0 + 8, x0:x0 + 8] += 1
        want = (acc / cnt).astype(F32)  # the oracle stores float32: equal to within one f32 ulp
        assert np.all(np.abs(of[plane >> 1::2, plane & 1::2] - want) <= np.spacing(np.abs(want)))


def test_dct_denoise_fp32_close_to_fp64():
    """The float32 statement stays within 0.05 DN of the float64 one on a 14-bit scale."""
    mos = _dn_frame(64, 48, seed=9)
    black = [512.0] * 4
    a, _ = O.dct_denoise(mos, black, 2.0, 144.0, 2.7, 4, precision=64)
    b, _ = O.dct_denoise(mos, black, 2.0, 144.0, 2.7, 4, precision=32)
    assert np.abs(a - b).max() < 0.05


# ---------------------------------------------------------------- golden fixtures

GOLDEN = __import__("pathlib").Path(__file__).parent / "golden"


@pytest.mark.parametrize("cfa", ["RGGB", "BGGR", "GRBG", "GBRG"])
def test_golden_develop_fixtures(cfa, lut8):
    """The oracle reproduces the committed fixtures, and so does the numpy statement."""
    g = np.load(GOLDEN / "develop_small.npz")
    mos = g[f"mosaic_{cfa}"]
    black = (2, 2, 1, g["black"])
    for dm in ("bilinear", "ahd"):
        lin = N.develop(mos, cfa, origin=(1, 0), black=black, white=tuple(g["white"]), wb=tuple(g["wb"]),
                        matrix=g["matrix"], demosaic=dm)
        for enc in ("rgba8", "linear16"):
            got = O.develop(mos, cfa, origin=(1, 0), black=black, white=tuple(g["white"]), wb=tuple(g["wb"]),
                            matrix=g["matrix"], demosaic=dm, encoding=enc)
            assert np.array_equal(got, g[f"{cfa}_{dm}_{enc}"]), (dm, enc)
        assert np.array_equal(N.rgba8(lin, lut8), g[f"{cfa}_{dm}_rgba8"]), dm
        assert np.array_equal(N.linear16(lin), g[f"{cfa}_{dm}_linear16"]), dm


def test_golden_dct_fixture():
    g = np.load(GOLDEN / "dct_small.npz")
    a, b, thr, step = (float(v) for v in g["params"])
    of, ou = O.dct_denoise(g["mosaic"], g["black"], a, b, thr, int(step), precision=64)
    assert np.array_equal(of, g["out_f32"]) and np.array_equal(ou, g["out_u16"])
    want = N.dct_denoise(g["mosaic"], g["black"], a, b, thr, int(step))
    assert np.abs(g["out_f32"] - want).max() < 1e-3


# ---------------------------------------------------------------- edge-aware (Bayer)


def _edge_args(cfa, origin, seed=0, beta=0.0, sparse=True):
    g0, d = N.edge_tables(seed, sparse)
    cx, cy = N.edge_plan(cfa, origin)
    return g0, d, dict(g0=g0, dir=d, cx=cx, cy=cy, beta=beta)


@pytest.mark.parametrize("cfa", BAYERS)
@pytest.mark.parametrize("origin", [(0, 0), (1, 1)])
@pytest.mark.parametrize("beta", [0.0, 0.35])
def test_edge_aware_equals_per_plane_statement(cfa, origin, beta):
    """The oracle's edge-aware band (edgeaware/mod.rs:423-822 structure) against the numpy
    statement, bit for bit, with synthetic tables (zero taps skipped as in mod.rs:130-160)."""
    w, h = 53, 41
    mos = N.lcg_stream(w, h, seed=31 + origin[0])
    g0, d, edge = _edge_args(cfa, origin, seed=len(cfa) + origin[1], beta=beta)
    kw = dict(origin=origin, black=(1, 1, 1, [64.0]), white=(4095,), wb=WB, matrix=MATRIX)
    got = O.develop(mos, cfa, demosaic="edge", encoding="linear_f32", edge=edge, **kw)
    want = N.develop_edge(mos, cfa, g0=g0, dirt=d, beta=beta, **kw)
    assert np.array_equal(got, want)


def test_edge_aware_band_and_crop_independence():
    """Bands of 64 rows and crops give the whole frame's pixels (edgeaware/mod.rs:43-46)."""
    mos = N.lcg_stream(90, 150, seed=4)
    g0, d, edge = _edge_args("GRBG", (0, 0), seed=2, beta=0.2)
    kw = dict(black=(1, 1, 1, [64.0]), white=(4095,), wb=WB, matrix=MATRIX, demosaic="edge", encoding="linear_f32",
              edge=edge)
    whole = O.develop(mos, "GRBG", **kw)
    part = O.develop(mos, "GRBG", crop=(11, 60, 50, 70), **kw)
    assert np.array_equal(part, whole[60:130, 11:61])


def test_edge_aware_flat_field_stays_flat():
    """Filters summing to one keep a flat field flat (edgeaware/mod.rs:13-14)."""
    mos = np.full((40, 48), 1200, dtype=np.uint16)
    _, _, edge = _edge_args("RGGB", (0, 0), seed=5, beta=0.3)
    got = O.develop(mos, "RGGB", black=(1, 1, 1, [0.0]), white=(4000,), demosaic="edge", encoding="linear_f32",
                    edge=edge)
    assert np.abs(got - 0.3).max() < 1e-5


def test_edge_aware_planted_mistake_is_caught():
    """A plausible slip (pass 2 without the gate) changes the bytes the comparison checks."""
    w, h = 53, 41
    mos = N.lcg_stream(w, h, seed=9)
    g0, d, edge = _edge_args("RGGB", (0, 0), seed=1)
    got = O.develop(mos, "RGGB", black=(1, 1, 1, [64.0]), white=(4095,), demosaic="edge", encoding="linear_f32",
                    edge=edge)
    noise = N.lcg_stream(w, h, seed=10)
    other = O.develop(np.minimum(mos, noise), "RGGB", black=(1, 1, 1, [64.0]), white=(4095,), demosaic="edge",
                      encoding="linear_f32", edge=edge)
    assert not np.array_equal(got, other)
    g0b = g0.copy()
    g0b[:, [10, 38]] = g0b[:, [38, 10]]  # two taps swapped
    bad = O.develop(mos, "RGGB", black=(1, 1, 1, [64.0]), white=(4095,), demosaic="edge", encoding="linear_f32",
                    edge=dict(edge, g0=g0b))
    assert not np.array_equal(got, bad)


# ---------------------------------------------------------------- edge-aware (X-Trans)


def _xt_args(origin, seed=0, beta=0.0):
    g0, d = N.edge_tables_xtrans(seed)
    cx, cy, canon, period = N.edge_plan_p(N.XTRANS, origin)
    return g0, d, dict(g0=g0, dir=d, cx=cx, cy=cy, beta=beta, canon=canon, period=period)


@pytest.mark.parametrize("origin", [(0, 0), (2, 3), (5, 1)])
@pytest.mark.parametrize("beta", [0.0, 0.2, 0.6])
def test_edge_aware_xtrans_equals_per_plane_statement(origin, beta):
    """X-Trans: 36 canonical positions, pass 1 the median of the filtered differences, pass 2's
    green re-estimate at red and blue sites with weight clamp(1 - 2 beta, 0, 1) (0 at beta 0.6)."""
    w, h = 58, 47
    mos = N.lcg_stream(w, h, seed=17 + origin[0])
    g0, d, edge = _xt_args(origin, seed=origin[1], beta=beta)
    kw = dict(origin=origin, black=(1, 1, 1, [64.0]), white=(4095,), wb=WB, matrix=MATRIX)
    got = O.develop(mos, XTRANS, demosaic="edge", encoding="linear_f32", edge=edge, **kw)
    want = N.develop_edge(mos, N.XTRANS, g0=g0, dirt=d, beta=beta, **kw)
    assert np.array_equal(got, want)


def test_edge_aware_xtrans_flat_field_and_bands():
    mos = np.full((50, 60), 1500, dtype=np.uint16)
    _, _, edge = _xt_args((0, 0), seed=4, beta=0.3)
    got = O.develop(mos, XTRANS, black=(1, 1, 1, [0.0]), white=(4000,), demosaic="edge", encoding="linear_f32",
                    edge=edge)
    # the edge-aware interior (corner pixels of the bilinear border may have no red or blue neighbour)
    assert np.abs(got[8:-8, 8:-8] - 0.375).max() < 1e-5
    mos = N.lcg_stream(70, 150, seed=6)
    _, _, edge = _xt_args((1, 4), seed=2, beta=0.1)
    kw = dict(origin=(1, 4), black=(1, 1, 1, [64.0]), white=(4095,), wb=WB, matrix=MATRIX, demosaic="edge",
              encoding="linear_f32", edge=edge)
    whole = O.develop(mos, XTRANS, **kw)
    part = O.develop(mos, XTRANS, crop=(9, 70, 45, 66), **kw)
    assert np.array_equal(part, whole[70:136, 9:54])
