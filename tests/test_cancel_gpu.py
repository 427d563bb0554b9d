"""Cancellation between stages and the bounded per-frame table cache, on the device.

Cancellation follows the reference's GPU facade (crates/flashraw/src/gpu.rs:292-299, 341-344,
462-492): the token is checked between stages and launches, never inside a kernel. A token whose
poll answers "cancel" on its k-th check stops the call after exactly k - 1 checks passed, so every
stage boundary is reached deterministically from the host; a token that never trips leaves the
result bit-identical to a run without one.
"""
import ctypes

import numpy as np
import pytest
import torch

import flashraw_amd as F
from flashraw_amd import gpu as G
from flashraw_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MATRIX = ((1.6, -0.35, -0.25), (-0.2, 1.45, -0.25), (0.05, -0.45, 1.4))
WB = (2.1, 1.0, 1.6, 1.0)


class _Countdown:
    """A poll that answers True on its n-th call."""

    def __init__(self, n):
        self.n, self.calls = n, 0

    def __call__(self):
        self.calls += 1
        return self.calls >= self.n


def _frame(w=256, h=192, seed=11):
    mos = synth.mosaic(w, h, "RGGB", seed=seed)
    info = F.FrameInfo(w, h, F.CfaPattern.RGGB, F.BlackLevel.uniform(512.0), F.WhiteLevel.uniform(16383))
    return mos, info


def _want(mos, demosaic="ahd"):
    return O.develop(mos, "RGGB", black=(1, 1, 1, [512.0]), white=(16383,), wb=WB, matrix=MATRIX,
                     demosaic=demosaic, encoding="rgba8")


def test_precancelled_develop_returns_cancelled(ctx):
    mos, info = _frame()
    tok = F.CancelToken()
    tok.cancel()
    spec = F.DevelopSpec(F.Rect.full(*mos.shape[::-1]), wb=WB, cam_to_rgb=MATRIX, edge=F.ahd_spec(info), cancel=tok)
    with pytest.raises(F.Cancelled):
        ctx.run(F.GpuFrame.from_u16(mos, mos.shape[1], mos.shape[0]), info, F.GpuOp.Develop(spec))


def test_run_develop_stops_at_every_stage_boundary(ctx):
    """Host frame, banded (4 bands): checks at entry, after the upload, before each band and after
    the develop. Cancelling at check k for every k short of the total raises Cancelled; a token that
    never trips gives the oracle's bytes."""
    mos, info = _frame(256, 192)
    h, w = mos.shape
    ctx.override_limits(w * 4 * 48)  # 48 RGBA8 rows (and 60 input rows) per band: 4 bands
    try:
        outcomes = []
        for k in range(1, 12):
            poll = _Countdown(k)
            spec = F.DevelopSpec(F.Rect.full(w, h), wb=WB, cam_to_rgb=MATRIX, edge=F.ahd_spec(info),
                                 cancel=F.CancelToken(poll))
            try:
                run = ctx.run(F.GpuFrame.from_u16(mos, w, h), info, F.GpuOp.Develop(spec))
                outcomes.append(("ok", k, poll.calls, run.timings.bands))
                got = run.image.download()
                assert np.array_equal(got, _want(mos)), k
            except F.Cancelled:
                outcomes.append(("cancelled", k, poll.calls, None))
        print(outcomes)
        first_ok = next(i for i, o in enumerate(outcomes) if o[0] == "ok")
        # entry + after upload + 4 bands + after the develop = 7 checks: k = 1..7 cancel
        assert all(o[0] == "cancelled" for o in outcomes[:first_ok])
        assert all(o[0] == "ok" for o in outcomes[first_ok:])
        assert first_ok == 7 and outcomes[first_ok][3] == 4, outcomes
    finally:
        ctx.override_limits(0)


def test_develop_into_cancelled_and_not(ctx):
    mos, info = _frame(200, 150, seed=4)
    h, w = mos.shape
    d_in = torch.from_numpy(mos.view(np.int16)).to("cuda")
    dst = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
    frame = F.GpuFrame.from_u16(d_in, w, h)
    tok = F.CancelToken()
    spec = F.DevelopSpec(F.Rect.full(w, h), wb=WB, cam_to_rgb=MATRIX, edge=F.ahd_spec(info), cancel=tok)
    ctx.develop_into(frame, info, spec, dst)  # not cancelled: runs
    torch.cuda.synchronize()
    assert np.array_equal(dst.cpu().numpy(), _want(mos))
    tok.cancel()
    dst.zero_()
    torch.cuda.synchronize()
    with pytest.raises(F.Cancelled):
        ctx.develop_into(frame, info, spec, dst)
    torch.cuda.synchronize()
    assert int(dst.cpu().numpy().max()) == 0  # nothing was launched


def test_spec_without_the_cancel_field_is_accepted(ctx):
    """A develop spec whose size ends before `cancel` (the layout without the field) develops as
    before and reads no token."""
    mos, info = _frame(128, 96, seed=8)
    h, w = mos.shape
    d_in = torch.from_numpy(mos.view(np.int16)).to("cuda")
    dst = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
    tok = F.CancelToken()
    tok.cancel()  # would stop the develop if it were read
    spec = F.DevelopSpec(F.Rect.full(w, h), wb=WB, cam_to_rgb=MATRIX, edge=F.ahd_spec(info), cancel=tok)
    f, k1 = F.GpuFrame.from_u16(d_in, w, h)._c()
    i, k2 = info._c()
    s, k3 = spec._c()
    s.size = G._DevelopSpec.cancel.offset
    G.check(F.lib().frh_develop_into(ctx.handle, ctypes.byref(f), ctypes.byref(i), ctypes.byref(s),
                                     dst.data_ptr(), dst.numel(), None, 0, ctx.stream_handle(None)))
    torch.cuda.synchronize()
    s.size = G._DevelopSpec.cancel.offset - 1
    with pytest.raises(F.InvalidArgument):
        G.check(F.lib().frh_develop_into(ctx.handle, ctypes.byref(f), ctypes.byref(i), ctypes.byref(s),
                                         dst.data_ptr(), dst.numel(), None, 0, ctx.stream_handle(None)))
    del k1, k2, k3
    assert np.array_equal(dst.cpu().numpy(), _want(mos))


def test_noise_reduction_stops_between_stages(ctx):
    """GpuContext::denoise with a token: checks at entry, after the pyramids, after the noise
    estimate, after the passes, after the guided filter and after the encode (6). k = 1..6 cancel;
    k = 7 completes with the same bytes as a run without a token."""
    w, h = 160, 120
    mos = synth.mosaic(w, h, "RGGB", black=256.0, white=4095.0, gain=0.6, read_sigma=14.0, seed=3)
    info = F.FrameInfo(w, h, F.CfaPattern.RGGB, F.BlackLevel.uniform(256.0), F.WhiteLevel.uniform(4095))
    dspec = F.DevelopSpec(F.Rect.full(w, h), wb=(2.0, 1.0, 1.6, 1.0), cam_to_rgb=MATRIX,
                          output=F.DevelopOutput.LinearRgbF32)
    img = ctx.run(F.GpuFrame.from_u16(mos, w, h), info, F.GpuOp.Develop(dspec)).image
    plain = ctx.denoise(img, F.denoise_spec(w, h, 0.5, 0.5)).image.download()
    for k in range(1, 8):
        poll = _Countdown(k)
        spec = F.denoise_spec(w, h, 0.5, 0.5)
        spec.cancel = F.CancelToken(poll)
        if k <= 6:
            with pytest.raises(F.Cancelled):
                ctx.denoise(img, spec)
            assert poll.calls == k
        else:
            got = ctx.denoise(img, spec).image.download()
            assert poll.calls == 6
            assert np.array_equal(got, plain)


def test_table_cache_stays_bounded_and_pinned_plans_survive(ctx):
    """Frames with distinct black-level row deltas (1 MiB each) do not grow the context's table
    cache past its 64 MiB bound; a prepared plan made before the flood keeps its tables and still
    develops the oracle's bytes after it."""
    w, h = 16, 262144
    rng = np.random.default_rng(5)
    mos = rng.integers(600, 15000, size=(h, w), dtype=np.uint16)
    d_in = torch.from_numpy(mos.view(np.int16)).to("cuda")
    dst = torch.empty((h, w, 4), dtype=torch.uint8, device="cuda")
    frame = F.GpuFrame.from_u16(d_in, w, h)

    def info_for(seed):
        rd = tuple(float(v) for v in np.random.default_rng(seed).uniform(-8, 8, h).astype(np.float32))
        return F.FrameInfo(w, h, F.CfaPattern.RGGB,
                           F.BlackLevel(1, 1, 1, (512.0,), rd, ()), F.WhiteLevel.uniform(16383)), rd

    spec = F.DevelopSpec(F.Rect.full(w, h), wb=WB, cam_to_rgb=MATRIX)
    info0, rd0 = info_for(0)
    plan = ctx.prepare(frame, info0, spec)
    want0 = O.develop(mos, "RGGB", black=(1, 1, 1, [512.0]), white=(16383,), wb=WB, matrix=MATRIX,
                      demosaic="bilinear", encoding="rgba8", row_deltas=rd0)
    peak = 0
    for s in range(1, 90):
        info, _ = info_for(s)
        ctx.develop_into(frame, info, spec, dst)
        peak = max(peak, ctx.table_cache()[1])
    torch.cuda.synchronize()
    info, rd = info_for(89)
    want = O.develop(mos, "RGGB", black=(1, 1, 1, [512.0]), white=(16383,), wb=WB, matrix=MATRIX,
                     demosaic="bilinear", encoding="rgba8", row_deltas=rd)
    assert np.array_equal(dst.cpu().numpy(), want)
    entries, nbytes = ctx.table_cache()
    print(f"table cache: {entries} entries, {nbytes} bytes (peak {peak})")
    assert peak <= (64 << 20) + (1 << 20)
    dst.zero_()
    plan.launch(d_in, dst)
    torch.cuda.synchronize()
    assert np.array_equal(dst.cpu().numpy(), want0)
    del plan
