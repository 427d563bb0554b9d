"""GPU input side and develop layouts beyond u16 Bayer mosaics, through the C ABI, against the
oracle, bit for bit:

- GpuOp::Unpack over the reference's 768-case matrix (crates/flashraw-gpu/tests/equivalence.rs:
  92-125: every depth 1..16, every PackOrder, four sizes, tight / padded / truncated payloads),
  windows with a linearisation curve, and u16 input through the same op (equivalence.rs:127-163);
- GpuOp::Normalize of Bayer, X-Trans, RGB and monochrome frames with the reference's
  calibrated_info (equivalence.rs:166-215), and of packed, linearised and f32 input;
- the develop of Bayer, X-Trans, RGB (3 components) and monochrome frames at full and half size in
  every orientation (equivalence.rs:217-330), packed and f32 input through bilinear, AHD and
  edge-aware, resident input, and bands.
"""
import dataclasses

import numpy as np
import pytest

import flashraw_amd as F
import np_restatement as N
from flashraw_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ORDERS = [F.PackOrder.MsbBytes, F.PackOrder.LsbBytes, F.PackOrder.Msb16, F.PackOrder.Msb32]
XTRANS = "GGRGGB/GGBGGR/BRGRBG/GGBGGR/GGRGGB/RBGBRG"
MATRIX = ((1.5, -0.3, -0.2), (-0.2, 1.4, -0.2), (0.05, -0.4, 1.35))
WB = (2.1, 1.0, 1.6, 1.0)
ENC = {
    F.DevelopOutput.Rgba8: "rgba8",
    F.DevelopOutput.LinearRgbF32: "linear_f32",
    F.DevelopOutput.Rgb16: "srgb16",
    F.DevelopOutput.LinearRgb16: "linear16",
    F.DevelopOutput.LinearRgbF16: "linear_f16",
}


def pseudo(n, seed):
    """equivalence.rs:39-50: xorshift bytes."""
    x = (seed | 1) & 0xFFFFFFFFFFFFFFFF
    out = bytearray(n)
    for i in range(n):
        x ^= (x << 13) & 0xFFFFFFFFFFFFFFFF
        x ^= x >> 7
        x ^= (x << 17) & 0xFFFFFFFFFFFFFFFF
        out[i] = x & 0xFF
    return np.frombuffer(bytes(out), dtype=np.uint8)


def _unpacked(run):
    return run.image.download().reshape(run.image.height(), -1)


def test_unpack_bit_exact_768_cases(ctx):
    cases = 0
    for bits in range(1, 17):
        for order in ORDERS:
            for w, h in ((1, 3), (7, 5), (33, 9), (257, 17)):
                rb = (w * bits + 7) // 8
                for stride, cut in ((rb, 0), (rb + 3, 0), (rb + 1, 7)):
                    src = pseudo(max(stride * h - cut, 0), bits * 31 + w * 7 + stride)
                    want = O.unpack(src, w, h, bits, int(order), stride)
                    frame = F.GpuFrame.from_packed(src, w, h, bits, order, stride)
                    got = _unpacked(ctx.run(frame, F.FrameInfo(w, h), F.GpuOp.Unpack))
                    assert np.array_equal(got, want), (bits, order, w, h, stride, cut)
                    cases += 1
    assert cases == 768


def test_unpack_window_lut_and_u16(ctx):
    w, h, bits = 301, 47, 14
    stride = (w * bits + 7) // 8 + 5
    src = pseudo(stride * h, 99)
    full = O.unpack(src, w, h, bits, 0, stride)
    lut = np.array([(v * 3 // 2 + 7) & 0xFFFF for v in range(12000)], np.uint16)
    win = F.Rect(13, 9, 200, 31)
    frame = F.GpuFrame.from_packed(src, w, h, bits, F.PackOrder.MsbBytes, stride).with_window(win).with_linearization(lut)
    got = _unpacked(ctx.run(frame, F.FrameInfo(w, h), F.GpuOp.Unpack))
    want = O.linearize(full[win.y:win.y + win.height, win.x:win.x + win.width], lut)
    assert np.array_equal(got, want)
    # u16 input through the same op, odd width
    data = ((np.arange(w * h, dtype=np.uint64) * 2654435761) & 0xFFFFFFFF) >> 18
    data = data.astype(np.uint16).reshape(h, w)
    got = _unpacked(ctx.run(F.GpuFrame.from_u16(data, w, h), F.FrameInfo(w, h), F.GpuOp.Unpack))
    assert np.array_equal(got, data)
    got = _unpacked(ctx.run(F.GpuFrame.from_u16(data, w, h).with_window(win).with_linearization(lut),
                            F.FrameInfo(w, h), F.GpuOp.Unpack))
    assert np.array_equal(got, O.linearize(data[win.y:win.y + win.height, win.x:win.x + win.width], lut))


def _calibrated(w, h, cfa, cpp):
    """equivalence.rs:166-177."""
    black = F.BlackLevel(2, 2, 1, (510.0, 512.0, 509.0, 515.0), tuple((y % 5) * 0.5 for y in range(h + 8)),
                         tuple((x % 3) * -0.25 for x in range(w + 8)))
    return F.FrameInfo(w, h, cfa, black, F.WhiteLevel((16000, 16383, 15800)))


def _oracle_kw(info):
    b = info.black_level
    return dict(black=(b.rows, b.cols, b.samples, list(b.values)), white=tuple(info.white_level.values),
                row_deltas=list(b.row_deltas) or None, col_deltas=list(b.col_deltas) or None)


def _synthetic(w, h, cpp, seed):
    """equivalence.rs:179-189."""
    i = np.arange(w * h * cpp, dtype=np.uint64)
    x = (i // cpp) % w
    y = (i // cpp) // w
    n = ((i * np.uint64(2654435761) + np.uint64(seed)) & np.uint64(0xFFFFFFFF)) >> np.uint64(22)
    return np.minimum(300 + x * 40 + y * 25 + n, 16383).astype(np.uint16).reshape(h, w * cpp)


def _pat(cfa):
    return None if cfa is None else (F.CfaPattern.parse(cfa) if "/" in cfa else F.CfaPattern.bayer(cfa))


def _opat(cfa):
    return None if cfa is None else N.parse_cfa(cfa)


@pytest.mark.parametrize("cfa,cpp", [("RGGB", 1), (None, 3), (None, 1), (XTRANS, 1), (None, 4)])
def test_normalize_layouts(ctx, cfa, cpp):
    w, h = 123, 77
    info = _calibrated(w, h, _pat(cfa), cpp)
    arr = _synthetic(w, h, cpp, 5)
    origin = (3, 1)
    run = ctx.run(F.GpuFrame.from_u16(arr, w, h, cpp).with_origin(origin), info, F.GpuOp.Normalize)
    got = run.image.download().reshape(h, w * cpp)
    want = O.normalize(arr, _opat(cfa), origin=origin, cpp=cpp, clamped=True, **_oracle_kw(info))
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_normalize_packed_linearised_and_f32(ctx):
    w, h, bits = 211, 60, 12
    stride = (w * bits + 7) // 8 + 2
    src = pseudo(stride * h - 9, 5)
    lut = np.array([(v * 3 // 2 + 7) & 0xFFFF for v in range(4096)], np.uint16)
    info = _calibrated(w, h, F.CfaPattern.RGGB, 1)
    info.white_level = F.WhiteLevel((16383,))
    win = F.Rect(13, 9, 181, 41)
    for order in ORDERS:
        frame = F.GpuFrame.from_packed(src, w, h, bits, order, stride).with_window(win).with_linearization(lut)
        got = ctx.run(frame, info, F.GpuOp.Normalize).image.download().reshape(win.height, win.width)
        mos = O.linearize(O.unpack(src, w, h, bits, int(order), stride), lut)[win.y:win.y + win.height,
                                                                              win.x:win.x + win.width]
        want = O.normalize(mos, "RGGB", origin=(win.x, win.y), clamped=True, **_oracle_kw(info))
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), order
    rng = np.random.default_rng(4)
    arr = (rng.random((h, w * 3)) * 20000.0 - 1000.0).astype(np.float32)
    info3 = _calibrated(w, h, None, 3)
    got = ctx.run(F.GpuFrame.from_f32(arr, w, h, 3).with_linearization(lut), info3,
                  F.GpuOp.Normalize).image.download().reshape(h, w * 3)
    want = O.normalize(arr, None, cpp=3, clamped=True, **_oracle_kw(info3))
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def _compare_develop(ctx, frame, info, spec, want_fn, what):
    run = ctx.run(frame, info, F.GpuOp.Develop(spec))
    got = run.image.download()
    want = want_fn(ENC[spec.output])
    assert got.shape == want.shape, (what, got.shape, want.shape)
    if spec.output == F.DevelopOutput.LinearRgbF32:
        bad = np.count_nonzero(got.view(np.uint32) != want.view(np.uint32))
    else:
        bad = np.count_nonzero(got.astype(np.int64) != want.astype(np.int64))
    assert bad == 0, f"{what}: {bad} samples differ"
    return run


@pytest.mark.parametrize("case", ["bayer", "xtrans", "rgb", "mono"])
@pytest.mark.parametrize("half", [False, True])
def test_develop_layouts_orientations(ctx, case, half):
    """equivalence.rs:217-330: 157 x 93 frames, crop (5, 4, w - 9, h - 7), origin (1, 3), all eight
    orientations, at full and half size."""
    cfa, cpp = {"bayer": ("RGGB", 1), "xtrans": (XTRANS, 1), "rgb": (None, 3), "mono": (None, 1)}[case]
    w, h = 157, 93
    info = _calibrated(w, h, _pat(cfa), cpp)
    arr = _synthetic(w, h, cpp, 11)
    origin = (1, 3)
    crop = (5, 4, w - 9, h - 7)
    frame = F.GpuFrame.from_u16(arr, w, h, cpp).with_origin(origin)
    for orientation in range(1, 9):
        for out in (F.DevelopOutput.Rgba8, F.DevelopOutput.LinearRgbF32, F.DevelopOutput.Rgb16):
            if out == F.DevelopOutput.LinearRgbF32 and orientation != 1:
                continue
            spec = F.DevelopSpec(F.Rect(*crop), wb=WB, cam_to_rgb=MATRIX, orientation=F.Orientation(orientation),
                                 output=out, half_size=half)

            def want(enc):
                return O.develop(arr, _opat(cfa), cpp=cpp, origin=origin, wb=WB, matrix=MATRIX, crop=crop,
                                 orientation=orientation, half_size=half, encoding=enc, **_oracle_kw(info))
            _compare_develop(ctx, frame, info, spec, want, f"{case} half={half} orientation={orientation} {out.name}")


@pytest.mark.parametrize("case", ["xtrans", "rgb", "mono2", "bayer_half"])
def test_develop_generic_every_encoding_and_levels(ctx, case):
    cfa, cpp, half = {"xtrans": (XTRANS, 1, False), "rgb": (None, 3, False), "mono2": (None, 2, False),
                      "bayer_half": ("GBRG", 1, True)}[case]
    w, h = 96, 71
    info = _calibrated(w, h, _pat(cfa), cpp)
    arr = _synthetic(w, h, cpp, 3)
    frame = F.GpuFrame.from_u16(arr, w, h, cpp).with_origin((2, 1))
    for out in ENC:
        spec = F.DevelopSpec(F.Rect.full(w, h), wb=WB, cam_to_rgb=MATRIX, output=out, half_size=half,
                             levels=(1.3, 0.01))

        def want(enc):
            return O.develop(arr, _opat(cfa), cpp=cpp, origin=(2, 1), wb=WB, matrix=MATRIX, half_size=half,
                             levels=(1.3, 0.01), encoding=enc, **_oracle_kw(info))
        _compare_develop(ctx, frame, info, spec, want, f"{case} {out.name}")


@pytest.mark.parametrize("demosaic", ["bilinear", "ahd", "edge"])
def test_develop_packed_and_linearised(ctx, demosaic):
    """Packed payloads of every order, windowed, through the linearisation curve, developed like the
    unpacked, linearised mosaic."""
    w, h, bits = 211, 130, 12
    stride = (w * bits + 7) // 8 + 2
    scene = synth.mosaic(w, h, "RGGB", black=256.0, white=4095.0, seed=7, edges=30, discs=8)
    lut = np.array([min(v * 3, 65535) for v in range(4096)], np.uint16)
    win = F.Rect(7, 5, 190, 118)
    tables = synth.edge_tables(1)
    for order in ORDERS:
        word = {F.PackOrder.Msb16: 2, F.PackOrder.Msb32: 4}.get(order, 1)
        # the scene packed where the format allows it (whole words per row), else random payload bytes
        src = N.pack(scene, bits, int(order), stride) if ((w * bits + 7) // 8) % word == 0 else pseudo(stride * h, 3)
        mos = O.linearize(O.unpack(src, w, h, bits, int(order), stride), lut)[win.y:win.y + win.height,
                                                                              win.x:win.x + win.width]
        info = F.FrameInfo(win.width, win.height, F.CfaPattern.RGGB, F.BlackLevel.uniform(768.0),
                           F.WhiteLevel.uniform(12285))
        origin = (win.x, win.y)
        edge = None
        oedge = None
        if demosaic == "ahd":
            edge = F.ahd_spec(info, origin)
        elif demosaic == "edge":
            cx, cy = F.edge_plan(info, origin)
            edge = F.EdgeSpec(period=2, cx=cx, cy=cy, beta=0.2, halo=8, canon=F.EDGE_CANON, g0=tables[0],
                              dir=tables[1])
            oedge = dict(g0=tables[0], dir=tables[1], cx=cx, cy=cy, beta=0.2)
        frame = F.GpuFrame.from_packed(src, w, h, bits, order, stride).with_window(win).with_linearization(lut)
        for out in (F.DevelopOutput.Rgba8, F.DevelopOutput.LinearRgbF32):
            spec = F.DevelopSpec(F.Rect.full(win.width, win.height), wb=WB, cam_to_rgb=MATRIX, output=out, edge=edge)

            def want(enc):
                return O.develop(mos, "RGGB", origin=origin, black=(1, 1, 1, [768.0]), white=(12285,), wb=WB,
                                 matrix=MATRIX, demosaic=demosaic, edge=oedge, encoding=enc)
            _compare_develop(ctx, frame, info, spec, want, f"{demosaic} {order.name} {out.name}")


@pytest.mark.parametrize("demosaic", ["bilinear", "ahd", "edge"])
def test_develop_f32_input(ctx, demosaic):
    """f32 raw values take the LDS-tiled kernels (the register walkers read u16 rows)."""
    w, h = 150, 101
    scene = synth.mosaic(w, h, "BGGR", black=512.0, white=16383.0, seed=4, edges=30, discs=8).astype(np.float32)
    scene += np.float32(0.25)  # values between integers
    info = _calibrated(w, h, F.CfaPattern.bayer("BGGR"), 1)
    tables = synth.edge_tables(2)
    edge, oedge = None, None
    if demosaic == "ahd":
        edge = F.ahd_spec(info, (1, 0))
    elif demosaic == "edge":
        cx, cy = F.edge_plan(info, (1, 0))
        edge = F.EdgeSpec(period=2, cx=cx, cy=cy, beta=0.1, halo=8, canon=F.EDGE_CANON, g0=tables[0], dir=tables[1])
        oedge = dict(g0=tables[0], dir=tables[1], cx=cx, cy=cy, beta=0.1)
    frame = F.GpuFrame.from_f32(scene, w, h).with_origin((1, 0))
    for out in (F.DevelopOutput.Rgba8, F.DevelopOutput.LinearRgbF32):
        spec = F.DevelopSpec(F.Rect(3, 2, 140, 95), wb=WB, cam_to_rgb=MATRIX, output=out, edge=edge)

        def want(enc):
            return O.develop(scene, "BGGR", origin=(1, 0), wb=WB, matrix=MATRIX, demosaic=demosaic, edge=oedge,
                             crop=(3, 2, 140, 95), encoding=enc, **_oracle_kw(info))
        _compare_develop(ctx, frame, info, spec, want, f"f32 {demosaic} {out.name}")
    # X-Trans and RGB f32 frames through the generic kernel
    for cfa, cpp in ((XTRANS, 1), (None, 3)):
        arr = (_synthetic(w, h, cpp, 9).astype(np.float32) * np.float32(1.5))
        inf = _calibrated(w, h, _pat(cfa), cpp)
        spec = F.DevelopSpec(F.Rect.full(w, h), wb=WB, cam_to_rgb=MATRIX, output=F.DevelopOutput.LinearRgbF32)

        def want(enc):
            return O.develop(arr, _opat(cfa), cpp=cpp, wb=WB, matrix=MATRIX, encoding=enc, **_oracle_kw(inf))
        _compare_develop(ctx, F.GpuFrame.from_f32(arr, w, h, cpp), inf, spec, want, f"f32 {cfa} {cpp}")


def test_resident_packed_develop_and_ops(ctx):
    import torch
    w, h, bits = 256, 64, 14
    stride = (w * bits + 7) // 8
    scene = synth.mosaic(w, h, "GRBG", black=512.0, white=16383.0, seed=2)
    src = N.pack(scene, bits, 0, stride)
    d_src = torch.from_numpy(src.copy()).to("cuda")
    info = F.FrameInfo(w, h, F.CfaPattern.bayer("GRBG"), F.BlackLevel.uniform(512.0), F.WhiteLevel.uniform(16383))
    frame = F.GpuFrame.from_packed(d_src, w, h, bits, F.PackOrder.MsbBytes, stride, data_len=src.size)
    d_u16 = torch.empty((h, w), dtype=torch.int16, device="cuda")
    ctx.unpack_into(frame, d_u16)
    d_norm = torch.empty((h, w), dtype=torch.float32, device="cuda")
    ctx.normalize_into(frame, info, d_norm)
    spec = F.DevelopSpec(F.Rect.full(w, h), wb=WB, cam_to_rgb=MATRIX, edge=F.ahd_spec(info))
    dst = torch.empty((h, w, 4), dtype=torch.uint8, device="cuda")
    ctx.develop_into(frame, info, spec, dst)
    torch.cuda.synchronize()
    assert np.array_equal(d_u16.cpu().numpy().view(np.uint16), scene)
    want_n = O.normalize(scene, "GRBG", black=(1, 1, 1, [512.0]), white=(16383,), clamped=True)
    assert np.array_equal(d_norm.cpu().numpy(), want_n)
    want = O.develop(scene, "GRBG", black=(1, 1, 1, [512.0]), white=(16383,), wb=WB, matrix=MATRIX, demosaic="ahd")
    assert np.array_equal(dst.cpu().numpy(), want)


def test_generic_banded_equals_whole(ctx):
    w, h = 300, 257
    info = _calibrated(w, h, F.CfaPattern.parse(XTRANS), 1)
    arr = _synthetic(w, h, 1, 6)
    frame = F.GpuFrame.from_u16(arr, w, h)
    for half in (False, True):
        spec = F.DevelopSpec(F.Rect.full(w, h), wb=WB, cam_to_rgb=MATRIX, half_size=half)
        whole = ctx.run(frame, info, F.GpuOp.Develop(spec))
        ctx.override_limits(40 * 600)
        try:
            banded = ctx.run(frame, info, F.GpuOp.Develop(spec))
        finally:
            ctx.override_limits(0)
        assert banded.timings.bands > 1
        assert np.array_equal(whole.image.download(), banded.image.download())


def test_input_errors(ctx):
    import torch
    w, h = 64, 48
    info = F.FrameInfo(w, h, F.CfaPattern.RGGB, F.BlackLevel.uniform(512.0), F.WhiteLevel.uniform(16383))
    f32 = np.zeros((h, w), np.float32)
    with pytest.raises(F.InvalidArgument):
        ctx.run(F.GpuFrame.from_f32(f32, w, h), info, F.GpuOp.Unpack)
    src = pseudo(w * h * 2, 1)
    with pytest.raises(F.InvalidArgument):
        ctx.run(F.GpuFrame.from_packed(src, w, h, 17), info, F.GpuOp.Unpack)
    with pytest.raises(F.InvalidArgument):
        ctx.run(F.GpuFrame.from_f32(f32[:10], w, h), info, F.GpuOp.Normalize)
    packed = F.GpuFrame.from_packed(src, w, h, 12)
    with pytest.raises(F.NotAvailable):
        ctx.prepare(packed, info, F.DevelopSpec(F.Rect.full(w, h)))
    xt = F.FrameInfo(w, h, F.CfaPattern.parse(XTRANS), F.BlackLevel.uniform(512.0), F.WhiteLevel.uniform(16383))
    d_in = torch.zeros((h, w), dtype=torch.int16, device="cuda")
    dst = torch.zeros(w * h * 4, dtype=torch.uint8, device="cuda")
    with pytest.raises(F.NotAvailable):
        ctx.batch([(F.GpuFrame.from_u16(d_in, w, h), xt, F.DevelopSpec(F.Rect.full(w, h)), dst)])
    with pytest.raises(F.NotAvailable):  # the AHD of a non-Bayer pattern
        ctx.run(F.GpuFrame.from_u16(d_in.cpu().numpy().view(np.uint16), w, h), xt,
                F.GpuOp.Develop(F.DevelopSpec(F.Rect.full(w, h), edge=dataclasses.replace(
                    F.ahd_spec(info), canon=(0, 1, 1, 2)))))


# ---------------------------------------------------------------- display tone and output curves


@pytest.mark.parametrize("path", ["bilinear", "ahd", "edge", "xtrans", "half"])
def test_tone_and_curve(ctx, path):
    """GpuTone and GpuCurve on the display outputs of every develop path, bit for bit; linear
    outputs (also as the second output of the same pass) are untouched."""
    w, h = 140, 96
    tone = N.synthetic_tone(ev=0.5)
    lut8, lut16 = N.synthetic_curve()
    gtone = F.GpuTone(tone["table"], float(tone["xmax"]), float(tone["scale"]), tone["to_prophoto"], tone["white"])
    gcurve = F.GpuCurve(lut8, lut16)
    cfa = XTRANS if path == "xtrans" else "GRBG"
    info = _calibrated(w, h, _pat(cfa), 1)
    arr = _synthetic(w, h, 1, 13)
    origin = (1, 2)
    tables = synth.edge_tables(3)
    edge, oedge, demosaic = None, None, "bilinear"
    if path == "ahd":
        edge, demosaic = F.ahd_spec(info, origin), "ahd"
    elif path == "edge":
        cx, cy = F.edge_plan(info, origin)
        edge = F.EdgeSpec(period=2, cx=cx, cy=cy, beta=0.1, halo=8, canon=F.EDGE_CANON, g0=tables[0], dir=tables[1])
        oedge, demosaic = dict(g0=tables[0], dir=tables[1], cx=cx, cy=cy, beta=0.1), "edge"
    frame = F.GpuFrame.from_u16(arr, w, h).with_origin(origin)
    for curve in (None, gcurve):
        for out, orientation in ((F.DevelopOutput.Rgba8, 1), (F.DevelopOutput.Rgba8, 6), (F.DevelopOutput.Rgb16, 3)):
            spec = F.DevelopSpec(F.Rect.full(w, h), wb=WB, cam_to_rgb=MATRIX, output=out,
                                 orientation=F.Orientation(orientation), tone=gtone, curve=curve,
                                 half_size=path == "half", edge=edge, output2=F.DevelopOutput.LinearRgb16)
            run = ctx.run(frame, info, F.GpuOp.Develop(spec))
            kw = dict(origin=origin, wb=WB, matrix=MATRIX, demosaic=demosaic, edge=oedge, half_size=path == "half",
                      **_oracle_kw(info))
            want = O.develop(arr, _opat(cfa), orientation=orientation, encoding=ENC[out], tone=tone,
                             curve=None if curve is None else (lut8, lut16), **kw)
            got = run.image.download()
            assert np.array_equal(got, want), (path, curve is not None, out.name, orientation)
            want2 = O.develop(arr, _opat(cfa), encoding="linear16", **kw)
            assert np.array_equal(run.image2.download(), want2), (path, "linear16 untouched")
