"""Oracle pinning for the display outputs' standard tone (ToneMap, crates/flashraw/src/tone.rs:
81-207) and output curves (CurveTables, crates/flashraw/src/color.rs:343-367; Lut8 and srgb16,
process.rs:1679-1728, 1942-1950): the oracle's toned RGBA8 / sRGB16 bytes equal a numpy statement
applied to its linear output, neutral pixels get the curve on every channel, the middle channel
keeps its place between the toned extremes (the hue-preserving form), and a planted mistake
(extremes swapped) is caught."""
import numpy as np
import pytest

import np_restatement as N
from oracle import oracle as O

F32 = np.float32
MATRIX = ((1.5, -0.3, -0.2), (-0.2, 1.4, -0.2), (0.05, -0.4, 1.35))
WB = (2.1, 1.0, 1.6, 1.0)


def _lin(mos, **kw):
    return O.develop(mos, "RGGB", black=(1, 1, 1, [64.0]), white=(4095,), wb=WB, matrix=MATRIX,
                     encoding="linear_f32", **kw)


@pytest.mark.parametrize("demosaic", ["bilinear", "ahd"])
def test_tone_and_curve_match_the_numpy_statement(demosaic):
    mos = N.lcg_stream(61, 43, seed=3)
    tone = N.synthetic_tone()
    lut8, lut16 = N.synthetic_curve()
    kw = dict(black=(1, 1, 1, [64.0]), white=(4095,), wb=WB, matrix=MATRIX, demosaic=demosaic)
    lin = O.develop(mos, "RGGB", encoding="linear_f32", **kw)
    toned = N.tone_apply(lin, tone["table"], tone["xmax"], tone["scale"], tone["to_prophoto"], tone["white"])
    lut = O.srgb8_table()
    got = O.develop(mos, "RGGB", encoding="rgba8", tone=tone, **kw)
    want = N.rgba8(toned, lut)
    assert np.array_equal(got, want)
    got = O.develop(mos, "RGGB", encoding="rgba8", tone=tone, curve=(lut8, lut16), **kw)
    assert np.array_equal(got[..., :3], N.curve8(toned, lut8))
    got = O.develop(mos, "RGGB", encoding="srgb16", tone=tone, curve=(lut8, lut16), **kw)
    assert np.array_equal(got, N.curve16(toned, lut16))
    got = O.develop(mos, "RGGB", encoding="srgb16", curve=(lut8, lut16), **kw)
    assert np.array_equal(got, N.curve16(lin, lut16))
    # linear outputs ignore tone and curve
    assert np.array_equal(O.develop(mos, "RGGB", encoding="linear16", tone=tone, curve=(lut8, lut16), **kw),
                          O.develop(mos, "RGGB", encoding="linear16", **kw))


def test_neutral_pixels_get_the_curve_on_every_channel():
    tone = N.synthetic_tone(ev=0.0)
    tone["to_prophoto"] = np.eye(3, dtype=F32)
    tone["white"] = np.ones(3, F32)
    x = np.linspace(0.0, 1.2, 97, dtype=F32)
    lin = np.stack([x, x, x], -1)[None]
    out = N.tone_apply(lin, tone["table"], tone["xmax"], tone["scale"], tone["to_prophoto"], tone["white"])
    u = np.sqrt(np.minimum(x.astype(np.float64), 1.0)) * 1024
    i = np.minimum(u.astype(int), 1023)
    tab = tone["table"].astype(np.float64)
    want = tab[i, 0] + tab[i, 1] * (u - i)
    assert np.abs(out[0, :, 0] - want).max() < 1e-6
    assert np.array_equal(out[0, :, 0], out[0, :, 1]) and np.array_equal(out[0, :, 1], out[0, :, 2])


def test_middle_channel_keeps_its_place():
    tone = N.synthetic_tone(ev=0.3)
    tone["to_prophoto"] = np.eye(3, dtype=F32)
    tone["white"] = np.ones(3, F32)
    rng = np.random.default_rng(2)
    p = np.sort(rng.random((1, 500, 3)).astype(F32), axis=-1)[..., ::-1].copy()  # r >= g >= b
    out = N.tone_apply(p, tone["table"], tone["xmax"], tone["scale"], tone["to_prophoto"], tone["white"])
    r, g, b = (p[0, :, k].astype(np.float64) for k in range(3))
    fr, fb = out[0, :, 0].astype(np.float64), out[0, :, 2].astype(np.float64)
    ok = r - b > 1e-3
    mid = fb + (g - b) * (fr - fb) / np.where(ok, r - b, 1)
    assert np.abs(out[0, ok, 1] - mid[ok]).max() < 1e-5


def test_planted_mistake_is_caught():
    mos = N.lcg_stream(40, 30, seed=8)
    tone = N.synthetic_tone()
    lin = _lin(mos)
    got = O.develop(mos, "RGGB", black=(1, 1, 1, [64.0]), white=(4095,), wb=WB, matrix=MATRIX, encoding="srgb16",
                    tone=tone)
    lut16 = O.srgb16_table()
    bad = dict(tone, to_prophoto=np.asarray(tone["to_prophoto"]).T.copy())  # the matrix transposed
    worse = N.tone_apply(lin, bad["table"], bad["xmax"], bad["scale"], bad["to_prophoto"], bad["white"])
    assert not np.array_equal(got, N.curve16(worse, lut16))
    toned = N.tone_apply(lin, tone["table"], tone["xmax"], tone["scale"], tone["to_prophoto"], tone["white"])
    assert np.array_equal(got, N.curve16(toned, lut16))
