"""GPU edge-aware demosaic of 6x6 (X-Trans) mosaics: the HIP kernel through the C ABI against
the CPU oracle, byte for byte.

The X-Trans form of crates/flashraw/src/edgeaware/mod.rs: 36 canonical positions (the canonical
arrangement GGRGGB / GGBGGR / BRGRBG / GGBGGR / GGRGGB / RBGBRG, mod.rs:168-185), pass 1 the 3x3
median of the differences of the filtered planes (mod.rs:455-474, median_row 825-855), pass 2's
green re-estimate at red and blue sites with weight clamp(1 - 2 beta, 0, 1) (mod.rs:547-592); the
pixels within the halo of the frame border take the generic CFA mean (process.rs:1098-1126).
The cases mirror the Bayer ones of the reference's crates/flashraw-gpu/tests/edge_aware.rs
(shifted patterns through the sensor origin, clean and smoothed frames, crops, region windows,
small frames) on synthetic tables with the trained tables' layout (xtrans.bin is not shipped).
"""
import dataclasses

import numpy as np
import pytest

import flashraw_amd as F
import np_restatement as N
from flashraw_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

XT = N.XTRANS
XT_CFA = N.parse_cfa(XT)  # the pattern as rows of colours, for the oracle
ENC = {
    F.DevelopOutput.Rgba8: "rgba8",
    F.DevelopOutput.LinearRgbF32: "linear_f32",
    F.DevelopOutput.Rgb16: "srgb16",
    F.DevelopOutput.LinearRgb16: "linear16",
    F.DevelopOutput.LinearRgbF16: "linear_f16",
}
MATRIX = ((1.5, -0.3, -0.2), (-0.2, 1.4, -0.2), (0.05, -0.4, 1.35))


def _mosaic(w, h, seed, noisy=False):
    """The package's seeded scene sampled through the X-Trans pattern (one plane per colour)."""
    planes = [synth.mosaic(w, h, [[c, c], [c, c]], black=256.0, white=4095.0, gain=0.5,
                           read_sigma=8.0 if noisy else 1.0, edges=30, discs=10, stripes=6, seed=seed)
              for c in range(3)]
    return np.choose(N.colours(XT, w, h), planes).astype(np.uint16)


def _info(w, h, black=(1, 1, 1, [256.0]), white=(4095,)):
    rows, cols, samples, values = black
    return F.FrameInfo(w, h, F.CfaPattern.parse(XT), F.BlackLevel(rows, cols, samples, tuple(values), (), ()),
                       F.WhiteLevel(tuple(white)))


def _spec(info, origin, tables, beta):
    cx, cy, canon, period = F.edge_plan(info, origin, with_pattern=True)
    assert period == 6 and tuple(canon) == F.EDGE_CANON_XTRANS
    g0, d = tables
    return F.EdgeSpec(period=6, cx=cx, cy=cy, beta=beta, halo=8, canon=canon, g0=g0, dir=d, ahd=False)


def _compare(ctx, mos, *, tables, beta=0.0, output=F.DevelopOutput.Rgba8, origin=(0, 0), window=None, crop=None,
             black=(1, 1, 1, [256.0]), white=(4095,), wb=(2.0, 1.0, 1.5, 1.0), matrix=MATRIX, orientation=1,
             levels=None, output2=None, what=""):
    H, W = mos.shape
    win = window or F.Rect(0, 0, W, H)
    sub = mos[win.y:win.y + win.height, win.x:win.x + win.width]
    sub_origin = (origin[0] + win.x, origin[1] + win.y)
    info = _info(win.width, win.height, black, white)
    edge = _spec(info, sub_origin, tables, beta)
    assert (edge.cx, edge.cy) == N.edge_plan_p(XT, sub_origin)[:2]
    crop_r = F.Rect(*crop) if crop else F.Rect.full(win.width, win.height)
    spec = F.DevelopSpec(crop_r, wb=wb, cam_to_rgb=matrix, orientation=F.Orientation(orientation), output=output,
                         levels=levels, edge=edge, output2=output2)
    frame = F.GpuFrame.from_u16(mos, W, H).with_origin(origin).with_window(win)
    run = ctx.run(frame, info, F.GpuOp.Develop(spec))
    oedge = dict(g0=tables[0], dir=tables[1], cx=edge.cx, cy=edge.cy, beta=beta, canon=list(edge.canon), period=6)
    outs = [(output, run.image)] + ([(output2, run.image2)] if output2 is not None else [])
    for out, img in outs:
        got = img.download()
        want = O.develop(sub, XT_CFA, origin=sub_origin, black=black, white=white, wb=wb, matrix=matrix,
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


@pytest.mark.parametrize("origin", [(0, 0), (2, 3), (5, 1)])
@pytest.mark.parametrize("beta", [0.0, 0.2, 0.6])
def test_origins_and_smoothing(ctx, origin, beta):
    """Every shift of the pattern class through the sensor origin; beta 0 (no smoothing, full
    green re-estimate), 0.2 and 0.6 (no re-estimate: xg = 0); crop (5, 4, w - 9, h - 7)."""
    w, h = 203, 141
    mos = _mosaic(w, h, seed=7 + origin[0], noisy=beta > 0)
    tables = N.edge_tables_xtrans(origin[1])
    for out in (F.DevelopOutput.LinearRgbF32, F.DevelopOutput.Rgba8):
        _compare(ctx, mos, tables=tables, beta=beta, output=out, origin=origin, crop=(5, 4, w - 9, h - 7),
                 what=f"{origin} beta {beta}")


def test_every_encoding_and_dual_output(ctx):
    mos = _mosaic(97, 75, seed=3)
    tables = N.edge_tables_xtrans(11)
    for out in ENC:
        _compare(ctx, mos, tables=tables, beta=0.1, output=out, origin=(4, 2), what="encodings")
    _compare(ctx, mos, tables=tables, beta=0.0, output=F.DevelopOutput.LinearRgb16, levels=(1.37, 0.02),
             output2=F.DevelopOutput.Rgba8, orientation=6, what="levels + dual + rotated")


def test_window_calibration_and_small_frames(ctx):
    """A region of the raster (the halo counts from the window border), a 6x6 black pattern and
    per-colour white levels, and frames too small for any interior pixel or just large enough."""
    tables = N.edge_tables_xtrans(5)
    mos = _mosaic(203, 141, seed=21, noisy=True)
    _compare(ctx, mos, tables=tables, beta=0.3, output=F.DevelopOutput.LinearRgbF32, origin=(2, 1),
             window=F.Rect(13, 9, 150, 111), what="window")
    black6 = (6, 6, 1, [250.0 + (i * 7) % 11 for i in range(36)])
    _compare(ctx, mos, tables=tables, beta=0.05, output=F.DevelopOutput.LinearRgbF32, origin=(3, 0), black=black6,
             white=(4000, 4095, 3900), crop=(13, 9, 101, 77), what="6x6 black")
    for w, h in [(1, 1), (5, 7), (16, 16), (17, 17), (18, 19), (19, 40), (40, 18), (33, 9)]:
        small = _mosaic(w, h, seed=w * 100 + h)
        for out in (F.DevelopOutput.LinearRgbF32, F.DevelopOutput.Rgba8):
            _compare(ctx, small, tables=tables, beta=0.25, output=out, what=f"{w}x{h}")


def test_untamed_inputs(ctx):
    """A white balance beyond 1e6 and a tap of 1e13: the tap lists are followed literally."""
    mos = _mosaic(150, 97, seed=8)
    g0, d = N.edge_tables_xtrans(6)
    _compare(ctx, mos, tables=(g0, d), beta=0.3, output=F.DevelopOutput.LinearRgbF32, wb=(2.5e6, 1.0, 1.5, 1.0),
             what="huge wb")
    g0b, db = g0.copy(), d.copy()
    g0b[7, 9] = 1e13
    db[20, 0, 2, 10] = -1e13
    _compare(ctx, mos, tables=(g0b, db), beta=0.0, output=F.DevelopOutput.LinearRgbF32, what="huge taps")


def test_banded_equals_whole(ctx):
    mos = _mosaic(300, 257, seed=4, noisy=True)
    info = _info(300, 257)
    spec = F.DevelopSpec(F.Rect.full(300, 257), wb=synth.WB, cam_to_rgb=MATRIX,
                         edge=_spec(info, (0, 0), N.edge_tables_xtrans(0), 0.3))
    frame = F.GpuFrame.from_u16(mos, 300, 257)
    whole = ctx.run(frame, info, F.GpuOp.Develop(spec))
    ctx.override_limits(40 * 600)
    try:
        banded = ctx.run(frame, info, F.GpuOp.Develop(spec))
    finally:
        ctx.override_limits(0)
    assert banded.timings.bands > 1
    assert np.array_equal(whole.image.download(), banded.image.download())


def test_edge_spec_end_to_end(ctx):
    """flashraw::gpu::edge_spec on an X-Trans frame (green pairs one and two samples apart give
    the noise estimate), then the GPU develop against the oracle with that spec."""
    w, h = 203, 141
    origin = (1, 3)
    tables = N.edge_tables_xtrans(3)
    betas = []
    for read_sigma in (0.5, 40.0):
        planes = [synth.mosaic(w, h, [[c, c], [c, c]], black=256.0, white=4095.0, gain=8.0, read_sigma=read_sigma,
                               edges=30, discs=10, seed=11) for c in range(3)]
        mos = np.choose(N.colours(XT, w, h, origin), planes).astype(np.uint16)
        info = _info(w, h)
        norm = O.normalize(mos, XT_CFA, origin=origin, black=(1, 1, 1, [256.0]), white=(4095,))
        spec = F.edge_spec(info, origin, (w, h), (2.0, 1.0, 1.5, 1.0), lambda y: norm[y], tables)
        cx, cy = N.edge_plan_p(XT, origin)[:2]
        assert (spec.period, spec.cx, spec.cy) == (6, cx, cy)
        want_beta = O.edge_beta(norm, cx, cy, 1.0, canon=F.EDGE_CANON_XTRANS)
        assert np.float32(spec.beta) == np.float32(want_beta)
        betas.append(spec.beta)
        _compare(ctx, mos, tables=tables, beta=spec.beta, origin=origin, what=f"edge_spec sigma {read_sigma}")
    assert betas[0] < betas[1]


def test_errors(ctx):
    mos = _mosaic(64, 64, seed=1)
    info = _info(64, 64)
    frame = F.GpuFrame.from_u16(mos, 64, 64)
    g0, d = N.edge_tables_xtrans(0)
    good = _spec(info, (0, 0), (g0, d), 0.0)

    def run(edge, inf=info):
        ctx.run(frame, inf, F.GpuOp.Develop(F.DevelopSpec(F.Rect.full(64, 64), edge=edge)))

    run(good)
    for bad in (dataclasses.replace(good, g0=g0[:35]), dataclasses.replace(good, dir=d[:, :2]),
                dataclasses.replace(good, canon=good.canon[:4]), dataclasses.replace(good, cx=(good.cx + 1) % 6),
                dataclasses.replace(good, beta=-0.5), dataclasses.replace(good, halo=4)):
        with pytest.raises(F.InvalidArgument):
            run(bad)
    # X-Trans tables on a Bayer frame, Bayer tables on an X-Trans frame, AHD on X-Trans
    bayer = F.FrameInfo(64, 64, F.CfaPattern.RGGB, F.BlackLevel.uniform(256.0), F.WhiteLevel.uniform(4095))
    with pytest.raises(F.NotAvailable):
        run(good, bayer)
    bg0, bd = synth.edge_tables(0)
    with pytest.raises(F.NotAvailable):
        run(F.EdgeSpec(period=2, cx=0, cy=0, beta=0.0, halo=8, canon=(0, 1, 1, 2), g0=bg0, dir=bd))
    with pytest.raises(F.NotAvailable):
        run(F.EdgeSpec(period=2, cx=0, cy=0, beta=0.0, halo=6, canon=(0, 1, 1, 2), ahd=True))
