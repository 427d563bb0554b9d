"""frh_batch_develop_noalloc: the batched develop that allocates no device memory (its descriptors,
work lists and counters live in the library's fixed device array). Two calls in a row, the second
with fewer and smaller frames, each byte-identical to the batch object's launch
(frh_batch_create + frh_batch_launch) of the same frames and to the oracle; the limits are
refused before anything reaches the device."""
import numpy as np
import pytest
import torch

import flashraw_amd as F
from flashraw_amd import synth
from oracle import oracle as O
import test_batch_gpu as TB

pytestmark = pytest.mark.gpu


def _noalloc_equals_batch(ctx, jobs, stream):
    # the existing entry point
    b = ctx.batch([(j["frame"], j["info"], j["spec"], j["dst"]) for j in jobs])
    b.launch(stream=stream)
    stream.synchronize()
    want = [j["dst"].cpu().numpy().copy() for j in jobs]
    del b
    for j in jobs:
        j["dst"].fill_(0xA5)
    torch.cuda.synchronize()
    ctx.batch_develop_noalloc([(j["frame"], j["info"], j["spec"], j["dst"]) for j in jobs], stream=stream)
    stream.synchronize()
    for k, (j, w) in enumerate(zip(jobs, want)):
        got = j["dst"].cpu().numpy()
        assert np.array_equal(got, w), (k, j["mos"].shape, j["cfa"], j["ahd"])
    return want


def test_two_calls_in_a_row_second_smaller(ctx):
    s = torch.cuda.Stream()
    first = TB._jobs(11, 12, max_side=900)
    _noalloc_equals_batch(ctx, first, s)
    for j in first[:3]:  # and against the oracle
        want = O.develop(j["mos"], j["cfa"], origin=j["origin"], black=(1, 1, 1, [j["black"]]), white=(j["white"],),
                         wb=j["wb"], matrix=j["matrix"], demosaic="ahd" if j["ahd"] else "bilinear", encoding="rgba8")
        assert np.array_equal(j["dst"].cpu().numpy(), want.reshape(-1))
    kept = [j["dst"].cpu().numpy().copy() for j in first]
    second = TB._jobs(12, 4, max_side=300)
    _noalloc_equals_batch(ctx, second, s)
    # the second call writes only its own frames' outputs
    for j, k in zip(first, kept):
        assert np.array_equal(j["dst"].cpu().numpy(), k)


def test_strip_heights_keep_bytes(ctx):
    s = torch.cuda.Stream()
    jobs = TB._jobs(13, 6, max_side=500)
    ref = _noalloc_equals_batch(ctx, jobs, s)
    for rows in ((2, 2), (40, 12)):
        for j in jobs:
            j["dst"].zero_()
        torch.cuda.synchronize()
        ctx.batch_develop_noalloc([(j["frame"], j["info"], j["spec"], j["dst"]) for j in jobs], *rows, stream=s)
        s.synchronize()
        for k, j in enumerate(jobs):
            assert np.array_equal(j["dst"].cpu().numpy(), ref[k]), rows


def test_limits_refused_before_the_device(ctx):
    mos = synth.mosaic(64, 48, "RGGB", seed=1)
    d_in = torch.from_numpy(mos.view(np.int16)).to("cuda")
    dst = torch.zeros(64 * 48 * 4, dtype=torch.uint8, device="cuda")
    frame = F.GpuFrame.from_u16(d_in, 64, 48)
    spec = F.DevelopSpec(F.Rect.full(64, 48))
    plain = F.FrameInfo(64, 48, F.CfaPattern.RGGB, F.BlackLevel.uniform(512.0), F.WhiteLevel.uniform(16383))
    with pytest.raises(F.LimitExceeded):
        ctx.batch_develop_noalloc([(frame, plain, spec, dst)] * 257)
    entries = ctx.table_cache()[0]
    deltas = F.FrameInfo(64, 48, F.CfaPattern.RGGB, F.BlackLevel(1, 1, 1, (512.0,), (1.5,) * 48, ()),
                         F.WhiteLevel.uniform(16383))
    with pytest.raises(F.NotAvailable):
        ctx.batch_develop_noalloc([(frame, deltas, spec, dst)])
    tone = F.DevelopSpec(F.Rect.full(64, 48), curve=F.GpuCurve(np.arange(256, dtype=np.uint8), ()))
    with pytest.raises(F.NotAvailable):
        ctx.batch_develop_noalloc([(frame, plain, tone, dst)])
    assert ctx.table_cache()[0] == entries  # nothing was copied to the device
    with pytest.raises(F.NotAvailable):
        ctx.batch_develop_noalloc([(F.GpuFrame.from_u16(mos, 64, 48), plain, spec, dst)])
    with pytest.raises(F.InvalidArgument):
        ctx.batch_develop_noalloc([])
