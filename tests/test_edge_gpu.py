"""GPU edge-aware demosaic (Bayer) parity: the HIP kernel through the C ABI against the CPU oracle.

Mirrors the cases of the reference's crates/flashraw-gpu/tests/edge_aware.rs (the four Bayer
phases, sensor origins, clean and smoothed frames, crops, region windows, frames too small for
an interior, several bands, the 8-bit output) with the bar raised to bit-exactness: the kernel
performs the oracle's f32 operations in the oracle's order (no contraction, correctly rounded
division and square root), so every encoding must match byte for byte. Tables: the synthetic
ones of the package (the sparsity of trained tables: full 7x7 filters) and sparse random ones
(zero taps skipped, edgeaware/mod.rs:130-160).
"""
import numpy as np
import pytest

import flashraw_amd as F
import np_restatement as N
from flashraw_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ENC = {
    F.DevelopOutput.Rgba8: "rgba8",
    F.DevelopOutput.LinearRgbF32: "linear_f32",
    F.DevelopOutput.Rgb16: "srgb16",
    F.DevelopOutput.LinearRgb16: "linear16",
    F.DevelopOutput.LinearRgbF16: "linear_f16",
}
MATRIX = ((1.5, -0.3, -0.2), (-0.2, 1.4, -0.2), (0.05, -0.4, 1.35))
PHASES = ["RGGB", "BGGR", "GRBG", "GBRG"]


def _tables(kind, seed=0):
    if kind == "dense":
        return synth.edge_tables(seed)
    return N.edge_tables(seed, sparse=True)


def _spec(info, origin, tables, beta):
    cx, cy = F.edge_plan(info, origin)
    g0, d = tables
    return F.EdgeSpec(period=2, cx=cx, cy=cy, beta=beta, halo=8, canon=(0, 1, 1, 2), g0=g0, dir=d, ahd=False)


def _compare(ctx, mos, cfa, *, tables, beta=0.0, output=F.DevelopOutput.Rgba8, origin=(0, 0), window=None,
             crop=None, black=(1, 1, 1, [256.0]), white=(4095,), wb=(2.0, 1.0, 1.5, 1.0), matrix=MATRIX,
             orientation=1, levels=None, output2=None, what=""):
    H, W = mos.shape
    win = window or F.Rect(0, 0, W, H)
    sub = mos[win.y:win.y + win.height, win.x:win.x + win.width]
    sub_origin = (origin[0] + win.x, origin[1] + win.y)
    rows, cols, samples, values = black
    info = F.FrameInfo(win.width, win.height, F.CfaPattern.bayer(cfa),
                       F.BlackLevel(rows, cols, samples, tuple(values), (), ()), F.WhiteLevel(tuple(white)))
    edge = _spec(info, sub_origin, tables, beta)
    crop_r = F.Rect(*crop) if crop else F.Rect.full(win.width, win.height)
    spec = F.DevelopSpec(crop_r, wb=wb, cam_to_rgb=matrix, orientation=F.Orientation(orientation), output=output,
                         levels=levels, edge=edge, output2=output2)
    frame = F.GpuFrame.from_u16(mos, W, H).with_origin(origin).with_window(win)
    run = ctx.run(frame, info, F.GpuOp.Develop(spec))
    oedge = dict(g0=tables[0], dir=tables[1], cx=edge.cx, cy=edge.cy, beta=beta)
    outs = [(output, run.image)] + ([(output2, run.image2)] if output2 is not None else [])
    for out, img in outs:
        got = img.download()
        want = O.develop(sub, cfa, origin=sub_origin, black=black, white=white, wb=wb, matrix=matrix,
                         demosaic="edge", edge=oedge, encoding=ENC[out], orientation=orientation, crop=crop,
                         levels=levels)
        assert got.shape == want.shape, (what, out, got.shape, want.shape)
        if out == F.DevelopOutput.LinearRgbF32:
            bad = np.count_nonzero(got.view(np.uint32) != want.view(np.uint32))
        else:
            bad = np.count_nonzero(got.astype(np.int64) != want.astype(np.int64))
        if bad:
            d = np.abs(got.astype(np.float64) - want.astype(np.float64))
            idx = np.unravel_index(np.argmax(d), d.shape)
            raise AssertionError(f"{what} {out.name}: {bad} samples differ, max {d.max()} at {idx}")
    return run


def _mosaic(w, h, cfa, seed, noisy=False):
    return synth.mosaic(w, h, cfa, black=256.0, white=4095.0, gain=0.5, read_sigma=8.0 if noisy else 1.0,
                        edges=30, discs=10, stripes=6, seed=seed)


@pytest.mark.parametrize("cfa", PHASES)
@pytest.mark.parametrize("origin", [(0, 0), (1, 3)])
@pytest.mark.parametrize("beta", [0.0, 0.35])
def test_phases_origins_and_smoothing(ctx, cfa, origin, beta):
    """edge_aware.rs:199-220: 203 x 141 frames, crop (5, 4, w - 9, h - 7), linear f32 and RGBA8."""
    w, h = 203, 141
    mos = _mosaic(w, h, cfa, seed=7 + origin[0], noisy=beta > 0)
    tables = _tables("dense", seed=origin[1])
    for out in (F.DevelopOutput.LinearRgbF32, F.DevelopOutput.Rgba8):
        _compare(ctx, mos, cfa, tables=tables, beta=beta, output=out, origin=origin, crop=(5, 4, w - 9, h - 7),
                 what=f"{cfa} {origin} beta {beta}")


@pytest.mark.parametrize("cfa", PHASES)
def test_sparse_tables_every_encoding(ctx, cfa):
    mos = _mosaic(97, 75, cfa, seed=3)
    tables = _tables("sparse", seed=len(cfa) + ord(cfa[0]))
    for out in ENC:
        _compare(ctx, mos, cfa, tables=tables, beta=0.2, output=out, origin=(1, 0), what=f"{cfa} sparse")


@pytest.mark.parametrize("cfa", PHASES)
def test_window_and_small_frames(ctx, cfa):
    """edge_aware.rs:223-240: a region of the raster (the halo is measured from the window border)
    and frames too small for any interior pixel or just large enough for a few."""
    tables = _tables("dense", seed=5)
    mos = _mosaic(203, 141, cfa, seed=21, noisy=True)
    _compare(ctx, mos, cfa, tables=tables, beta=0.3, output=F.DevelopOutput.LinearRgbF32, origin=(2, 1),
             window=F.Rect(13, 9, 150, 111), what=f"{cfa} window")
    for w, h in [(1, 1), (2, 3), (15, 16), (16, 16), (17, 17), (18, 19), (19, 40), (40, 18), (33, 9)]:
        small = _mosaic(w, h, cfa, seed=w * 100 + h)
        for out in (F.DevelopOutput.LinearRgbF32, F.DevelopOutput.Rgba8):
            _compare(ctx, small, cfa, tables=tables, beta=0.25, output=out, what=f"{cfa} {w}x{h}")


@pytest.mark.parametrize("orientation", [1, 3, 6, 8])
def test_orientation_levels_dual_output(ctx, orientation):
    mos = _mosaic(120, 90, "GRBG", seed=9)
    tables = _tables("dense", seed=2)
    _compare(ctx, mos, "GRBG", tables=tables, beta=0.1, output=F.DevelopOutput.Rgb16, orientation=orientation,
             what=f"orientation {orientation}")
    _compare(ctx, mos, "GRBG", tables=tables, beta=0.0, output=F.DevelopOutput.LinearRgb16, levels=(1.37, 0.02),
             output2=F.DevelopOutput.Rgba8, orientation=orientation, what="levels + dual")


def test_calibration_crops(ctx):
    """A 2x2 black pattern and per-colour white levels, with crops that cut the halo."""
    mos = _mosaic(180, 130, "BGGR", seed=5)
    tables = _tables("dense", seed=4)
    for crop in [None, (0, 0, 150, 111), (13, 9, 101, 77), (3, 2, 140, 100), (100, 60, 80, 70)]:
        _compare(ctx, mos, "BGGR", tables=tables, beta=0.15, output=F.DevelopOutput.LinearRgbF32, origin=(1, 0),
                 black=(2, 2, 1, [250.0, 256.0, 251.0, 259.0]), white=(4000, 4095, 3900), crop=crop,
                 what=f"crop {crop}")


def test_banded_equals_whole(ctx):
    """Forced bands (override_limits, context.rs:704) give the bytes of one launch."""
    mos = _mosaic(300, 257, "RGGB", seed=4, noisy=True)
    info = F.FrameInfo(300, 257, F.CfaPattern.RGGB, F.BlackLevel.uniform(256.0), F.WhiteLevel.uniform(4095))
    spec = F.DevelopSpec(F.Rect.full(300, 257), wb=synth.WB, cam_to_rgb=MATRIX,
                         edge=_spec(info, (0, 0), _tables("dense"), 0.3))
    frame = F.GpuFrame.from_u16(mos, 300, 257)
    whole = ctx.run(frame, info, F.GpuOp.Develop(spec))
    assert whole.timings.bands == 1
    ctx.override_limits(40 * 600)
    try:
        banded = ctx.run(frame, info, F.GpuOp.Develop(spec))
    finally:
        ctx.override_limits(0)
    assert banded.timings.bands > 1
    assert np.array_equal(whole.image.download(), banded.image.download())


def test_differs_from_bilinear_and_ahd(ctx):
    """edge_aware.rs:136-141: the comparison is only meaningful if edge-aware is not bilinear."""
    mos = _mosaic(160, 120, "RGGB", seed=12)
    info = F.FrameInfo(160, 120, F.CfaPattern.RGGB, F.BlackLevel.uniform(256.0), F.WhiteLevel.uniform(4095))
    out = {}
    for name, edge in (("bilinear", None), ("ahd", F.ahd_spec(info)),
                       ("edge", _spec(info, (0, 0), _tables("dense"), 0.0))):
        spec = F.DevelopSpec(F.Rect.full(160, 120), output=F.DevelopOutput.LinearRgbF32, edge=edge)
        out[name] = ctx.run(F.GpuFrame.from_u16(mos, 160, 120), info, F.GpuOp.Develop(spec)).image.download()
    assert np.abs(out["edge"] - out["bilinear"]).max() > 1e-2
    assert np.abs(out["edge"] - out["ahd"]).max() > 1e-2


def test_tables_are_cached_per_contents(ctx):
    """Two specs with the same tables reuse one device copy; other tables give other bytes."""
    mos = _mosaic(90, 70, "RGGB", seed=1)
    info = F.FrameInfo(90, 70, F.CfaPattern.RGGB, F.BlackLevel.uniform(256.0), F.WhiteLevel.uniform(4095))
    frame = F.GpuFrame.from_u16(mos, 90, 70)
    res = []
    for seed in (0, 0, 1):
        spec = F.DevelopSpec(F.Rect.full(90, 70), output=F.DevelopOutput.LinearRgbF32,
                             edge=_spec(info, (0, 0), _tables("dense", seed), 0.0))
        res.append(ctx.run(frame, info, F.GpuOp.Develop(spec)).image.download())
    assert np.array_equal(res[0], res[1])
    assert not np.array_equal(res[0], res[2])


def test_errors(ctx):
    mos = _mosaic(64, 64, "RGGB", seed=1)
    info = F.FrameInfo(64, 64, F.CfaPattern.RGGB, F.BlackLevel.uniform(256.0), F.WhiteLevel.uniform(4095))
    frame = F.GpuFrame.from_u16(mos, 64, 64)
    g0, d = _tables("dense")
    good = _spec(info, (0, 0), (g0, d), 0.0)

    def run(edge):
        ctx.run(frame, info, F.GpuOp.Develop(F.DevelopSpec(F.Rect.full(64, 64), edge=edge)))

    run(good)
    import dataclasses
    for bad in (dataclasses.replace(good, g0=g0[:3]), dataclasses.replace(good, dir=d[:, :2]),
                dataclasses.replace(good, beta=1.5), dataclasses.replace(good, beta=float("nan")),
                dataclasses.replace(good, halo=7), dataclasses.replace(good, cx=1 - good.cx),
                dataclasses.replace(good, g0=np.where(g0 > 0.01, np.inf, g0).astype(np.float32))):
        with pytest.raises(F.InvalidArgument):
            run(bad)
    with pytest.raises(F.NotAvailable):
        run(dataclasses.replace(good, period=6))
    # the batched and served paths develop bilinear and AHD frames
    import torch
    d_in = torch.from_numpy(mos.view(np.int16)).to("cuda")
    dst = torch.zeros(64 * 64 * 4, dtype=torch.uint8, device="cuda")
    with pytest.raises(F.NotAvailable):
        ctx.batch([(F.GpuFrame.from_u16(d_in, 64, 64), info, F.DevelopSpec(F.Rect.full(64, 64), edge=good), dst)])


def test_24mp_full_frame(ctx):
    """The 24 MP workload (6048x4024 RGGB 14-bit) with full 7x7 tables and smoothing, every byte."""
    mos = synth.frame_24mp()
    info = F.FrameInfo(6048, 4024, F.CfaPattern.RGGB, F.BlackLevel.uniform(512.0), F.WhiteLevel.uniform(16383))
    tables = _tables("dense")
    edge = _spec(info, (0, 0), tables, 0.25)
    spec = F.DevelopSpec(F.Rect.full(6048, 4024), wb=synth.WB, cam_to_rgb=synth.MATRIX, edge=edge)
    run = ctx.run(F.GpuFrame.from_u16(mos, 6048, 4024), info, F.GpuOp.Develop(spec))
    got = run.image.download()
    want = O.develop(mos, "RGGB", black=(1, 1, 1, [512.0]), white=(16383,), wb=synth.WB, matrix=synth.MATRIX,
                     demosaic="edge", edge=dict(g0=tables[0], dir=tables[1], cx=0, cy=0, beta=0.25),
                     encoding="rgba8")
    assert np.array_equal(got, want), np.count_nonzero(got != want)
    print(f"24 MP edge-aware: kernel {run.timings.kernel_ms:.3f} ms")


def test_untamed_inputs(ctx):
    """Inputs that do not bound the intermediate values (a white balance beyond 1e6, a tap of
    1e13) take the kernel that follows the tap lists and the masked sums literally; it matches the
    oracle bit for bit too."""
    mos = _mosaic(150, 97, "GBRG", seed=8)
    g0, d = _tables("dense", seed=6)
    _compare(ctx, mos, "GBRG", tables=(g0, d), beta=0.3, output=F.DevelopOutput.LinearRgbF32,
             wb=(2.5e6, 1.0, 1.5, 1.0), what="huge wb")
    g0b, db = g0.copy(), d.copy()
    g0b[0, 3] = 1e13
    db[1, 0, 2, 10] = -1e13
    for out in (F.DevelopOutput.LinearRgbF32, F.DevelopOutput.Rgba8):
        _compare(ctx, mos, "GBRG", tables=(g0b, db), beta=0.0, output=out, what="huge taps")
    sparse = _tables("sparse", seed=9)
    _compare(ctx, mos, "GBRG", tables=sparse, beta=0.1, output=F.DevelopOutput.LinearRgbF32,
             black=(1, 1, 1, [2e6]), what="huge black")


@pytest.mark.parametrize("cfa", PHASES)
def test_edge_spec_end_to_end(ctx, cfa):
    """edge_aware.rs:146-160: the spec of flashraw::gpu::edge_spec (smoothing weight estimated
    from the normalised rows of the frame), then the GPU develop against the oracle with it."""
    w, h = 203, 141
    origin = (1, 3)
    tables = _tables("dense", seed=3)
    betas = []
    for read_sigma in (0.5, 40.0):
        mos = synth.mosaic(w, h, cfa, black=256.0, white=4095.0, gain=8.0, read_sigma=read_sigma, edges=30,
                           discs=10, seed=11)
        info = F.FrameInfo(w, h, F.CfaPattern.bayer(cfa), F.BlackLevel.uniform(256.0), F.WhiteLevel.uniform(4095))
        norm = O.normalize(mos, cfa, origin=origin, black=(1, 1, 1, [256.0]), white=(4095,))
        spec = F.edge_spec(info, origin, (w, h), (2.0, 1.0, 1.5, 1.0), lambda y: norm[y], tables)
        assert (spec.cx, spec.cy) == F.edge_plan(info, origin)
        betas.append(spec.beta)
        for out in (F.DevelopOutput.LinearRgbF32, F.DevelopOutput.Rgba8):
            _compare(ctx, mos, cfa, tables=tables, beta=spec.beta, output=out, origin=origin,
                     what=f"{cfa} sigma {read_sigma} beta {spec.beta}")
    assert betas[1] > betas[0]
