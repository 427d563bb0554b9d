"""The edge-aware demosaic on the device with the reference's trained filter tables (bayer.bin,
xtrans.bin of crates/flashraw/src/edgeaware/, kept under tests/golden/ as f32 data), against the
oracle byte for byte: the four Bayer phases, the X-Trans shifts and the 24 MP Bayer frame. Each
case records which kernel took it (the register-window kernel for inputs that bound every value,
or the general tap-list kernel) and the general kernel is forced once to show both give the
same bytes on these tables."""
import os

import numpy as np
import pytest

import flashraw_amd as F
from flashraw_amd import synth
from oracle import oracle as O
import test_edge_gpu as EB
import test_edge_xtrans_gpu as EX

pytestmark = pytest.mark.gpu

BAYER = synth.trained_edge_tables("bayer")
XTRANS = synth.trained_edge_tables("xtrans")


def _kernel(ctx, mos, info, spec):
    import torch
    h, w = mos.shape
    d = torch.from_numpy(np.ascontiguousarray(mos).view(np.int16)).to("cuda")
    return ctx.prepare(F.GpuFrame.from_u16(d, w, h), info, spec).kernel


def _bayer_kernel(ctx, mos, cfa, beta):
    h, w = mos.shape
    info = F.FrameInfo(w, h, F.CfaPattern.bayer(cfa), F.BlackLevel.uniform(256.0), F.WhiteLevel.uniform(4095))
    spec = F.DevelopSpec(F.Rect.full(w, h), wb=(2.0, 1.0, 1.5, 1.0), cam_to_rgb=EB.MATRIX,
                         edge=EB._spec(info, (0, 0), BAYER, beta))
    return _kernel(ctx, mos, info, spec)


@pytest.mark.parametrize("cfa", EB.PHASES)
@pytest.mark.parametrize("beta", [0.0, 0.35])
def test_bayer_phases(ctx, cfa, beta):
    w, h = 203, 141
    mos = EB._mosaic(w, h, cfa, seed=5, noisy=beta > 0)
    for out in (F.DevelopOutput.LinearRgbF32, F.DevelopOutput.Rgba8):
        EB._compare(ctx, mos, cfa, tables=BAYER, beta=beta, output=out, origin=(1, 0), crop=(5, 4, w - 9, h - 7),
                    what=f"trained {cfa} beta {beta}")
    kernel = _bayer_kernel(ctx, mos, cfa, beta)
    print(f"{cfa} beta {beta}: {kernel}")
    assert kernel == "edge_tame"


def test_general_kernel_gives_the_same_bytes(ctx):
    w, h = 203, 141
    mos = EB._mosaic(w, h, "GRBG", seed=9, noisy=True)
    os.environ["FLASHRAW_EDGE_IMPL"] = "general"
    try:
        assert _bayer_kernel(ctx, mos, "GRBG", 0.3) == "edge_general"
        for out in (F.DevelopOutput.LinearRgbF32, F.DevelopOutput.Rgba8):
            EB._compare(ctx, mos, "GRBG", tables=BAYER, beta=0.3, output=out, what="trained general kernel")
    finally:
        del os.environ["FLASHRAW_EDGE_IMPL"]
    assert _bayer_kernel(ctx, mos, "GRBG", 0.3) == "edge_tame"


@pytest.mark.parametrize("origin", [(0, 0), (2, 3), (5, 1)])
@pytest.mark.parametrize("beta", [0.0, 0.2])
def test_xtrans_shifts(ctx, origin, beta):
    w, h = 203, 141
    mos = EX._mosaic(w, h, seed=11 + origin[0], noisy=beta > 0)
    for out in (F.DevelopOutput.LinearRgbF32, F.DevelopOutput.Rgba8):
        run = EX._compare(ctx, mos, tables=XTRANS, beta=beta, output=out, origin=origin, crop=(5, 4, w - 9, h - 7),
                          what=f"trained X-Trans {origin} beta {beta}")
    info = EX._info(w, h)
    spec = F.DevelopSpec(F.Rect.full(w, h), wb=(2.0, 1.0, 1.5, 1.0), cam_to_rgb=EX.MATRIX,
                         edge=EX._spec(info, (0, 0), XTRANS, beta))
    assert _kernel(ctx, mos, info, spec) == "edge_general"  # period 6 takes the tap-list kernel
    assert run is not None


def test_24mp_bayer_trained(ctx):
    """The 24 MP workload (6048x4024 RGGB 14-bit) through the trained Bayer table, every byte."""
    mos = synth.frame_24mp()
    W, H = 6048, 4024
    info = F.FrameInfo(W, H, F.CfaPattern.RGGB, F.BlackLevel.uniform(512.0), F.WhiteLevel.uniform(16383))
    cx, cy = F.edge_plan(info, (0, 0))
    edge = F.EdgeSpec(period=2, cx=cx, cy=cy, beta=0.0, halo=8, canon=(0, 1, 1, 2), g0=BAYER[0], dir=BAYER[1],
                      ahd=False)
    spec = F.DevelopSpec(F.Rect.full(W, H), wb=synth.WB, cam_to_rgb=synth.MATRIX, edge=edge)
    assert _kernel(ctx, mos, info, spec) == "edge_tame"
    run = ctx.run(F.GpuFrame.from_u16(mos, W, H), info, F.GpuOp.Develop(spec))
    got = run.image.download()
    want = O.develop(mos, "RGGB", black=(1, 1, 1, [512.0]), white=(16383,), wb=synth.WB, matrix=synth.MATRIX,
                     demosaic="edge", encoding="rgba8", edge=dict(g0=BAYER[0], dir=BAYER[1], cx=cx, cy=cy, beta=0.0))
    print(f"24 MP trained edge-aware: kernel {run.timings.kernel_ms:.3f} ms")
    assert np.array_equal(got, want)
