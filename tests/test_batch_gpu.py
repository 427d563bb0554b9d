"""Batched develop: many frames of mixed size, Bayer phase, bit depth, white balance, colour matrix
and demosaic in one persistent launch per demosaic kind. Each frame's bytes must equal its own
single-frame develop (and the oracle), alone or in any mix, eagerly and from a replayed HIP graph."""
import numpy as np
import pytest

import flashraw_amd as F
from flashraw_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PHASES = ["RGGB", "BGGR", "GRBG", "GBRG"]


def _jobs(seed, n, max_side=700, outputs=(F.DevelopOutput.Rgba8,)):
    import torch
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(n):
        w = int(rng.integers(40, max_side))
        h = int(rng.integers(30, max_side))
        cfa = PHASES[int(rng.integers(0, 4))]
        bits = 12 if rng.random() < 0.5 else 14
        white = (1 << bits) - 1
        black = 64.0 * (bits - 10)
        mos = synth.mosaic(w, h, cfa, black=black, white=float(white), seed=seed * 100 + k, stripes=int(rng.integers(0, 20)))
        wb = (float(rng.uniform(1.2, 2.8)), 1.0, float(rng.uniform(1.2, 2.8)), 1.0)
        m = np.asarray(synth.MATRIX) * rng.uniform(0.9, 1.1, (3, 3))
        matrix = tuple(tuple(float(v) for v in row) for row in m)
        ahd = rng.random() < 0.5
        out = outputs[int(rng.integers(0, len(outputs)))]
        origin = (int(rng.integers(0, 2)), int(rng.integers(0, 2)))
        info = F.FrameInfo(w, h, F.CfaPattern.bayer(cfa), F.BlackLevel.uniform(black), F.WhiteLevel.uniform(white))
        spec = F.DevelopSpec(F.Rect.full(w, h), wb=wb, cam_to_rgb=matrix, output=out,
                             edge=F.ahd_spec(info, origin) if ahd else None)
        d_in = torch.from_numpy(mos.view(np.int16)).to("cuda")
        frame = F.GpuFrame.from_u16(d_in, w, h).with_origin(origin)
        _, _, _, nbytes = F.context().output_size(frame, spec, out)
        dst = torch.zeros(nbytes, dtype=torch.uint8, device="cuda")
        jobs.append(dict(frame=frame, info=info, spec=spec, dst=dst, mos=mos, cfa=cfa, black=black, white=white,
                         wb=wb, matrix=matrix, ahd=ahd, out=out, origin=origin, d_in=d_in))
    return jobs


def _single(ctx, j):
    img = ctx.run(F.GpuFrame.from_u16(j["mos"], j["mos"].shape[1], j["mos"].shape[0]).with_origin(j["origin"]),
                  j["info"], F.GpuOp.Develop(j["spec"])).image
    return img.download().reshape(-1).view(np.uint8)


def _batch(ctx, jobs):
    return ctx.batch([(j["frame"], j["info"], j["spec"], j["dst"]) for j in jobs])


def _check(ctx, jobs, what=""):
    for k, j in enumerate(jobs):
        got = j["dst"].cpu().numpy()
        want = _single(ctx, j)
        assert np.array_equal(got, want), (what, k, j["mos"].shape, j["cfa"], j["ahd"], j["out"])


def test_mixed_batch_equals_single_frames(ctx):
    import torch
    jobs = _jobs(1, 14, outputs=(F.DevelopOutput.Rgba8,))
    b = _batch(ctx, jobs)
    assert b.items > 0
    b.launch()
    torch.cuda.synchronize()
    _check(ctx, jobs, "eager")
    # and two frames against the oracle directly
    for j in jobs[:2]:
        want = O.develop(j["mos"], j["cfa"], origin=j["origin"], black=(1, 1, 1, [j["black"]]), white=(j["white"],),
                         wb=j["wb"], matrix=j["matrix"], demosaic="ahd" if j["ahd"] else "bilinear", encoding="rgba8")
        assert np.array_equal(j["dst"].cpu().numpy(), want.reshape(-1))


def test_mixed_outputs_use_the_general_path(ctx):
    import torch
    jobs = _jobs(2, 8, max_side=400,
                 outputs=(F.DevelopOutput.Rgba8, F.DevelopOutput.LinearRgb16, F.DevelopOutput.LinearRgbF32))
    b = _batch(ctx, jobs)
    b.launch()
    torch.cuda.synchronize()
    _check(ctx, jobs, "general")


def test_order_and_company_do_not_matter(ctx):
    import torch
    jobs = _jobs(3, 10, max_side=500)
    _batch(ctx, jobs).launch()
    torch.cuda.synchronize()
    first = [j["dst"].cpu().numpy().copy() for j in jobs]
    for j in jobs:
        j["dst"].zero_()
    rev = list(reversed(jobs))
    _batch(ctx, rev).launch()
    torch.cuda.synchronize()
    for j, f in zip(jobs, first):
        assert np.array_equal(j["dst"].cpu().numpy(), f)
    # alone
    for j, f in zip(jobs[:3], first[:3]):
        j["dst"].zero_()
        _batch(ctx, [j]).launch()
        torch.cuda.synchronize()
        assert np.array_equal(j["dst"].cpu().numpy(), f)


def test_graph_replays(ctx):
    """Capture once, replay three times with eager work and synchronisation in between; every
    replay reproduces the eager bytes."""
    import torch
    jobs = _jobs(4, 9, max_side=600)
    b = _batch(ctx, jobs)
    s = torch.cuda.Stream()
    b.launch(stream=s)
    s.synchronize()
    ref = [j["dst"].cpu().numpy().copy() for j in jobs]
    b.capture(s)
    for rep in range(3):
        for j in jobs:
            j["dst"].zero_()
        torch.cuda.synchronize()
        b.launch(stream=s)
        s.synchronize()
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        float(junk)
        for k, j in enumerate(jobs):
            assert np.array_equal(j["dst"].cpu().numpy(), ref[k]), (rep, k)


def test_autotune_keeps_bytes(ctx):
    import torch
    jobs = _jobs(5, 6, max_side=800)
    b = _batch(ctx, jobs)
    b.launch()
    torch.cuda.synchronize()
    ref = [j["dst"].cpu().numpy().copy() for j in jobs]
    ra, rb = b.autotune(iters=1)
    assert 16 <= ra <= 128 and 8 <= rb <= 64
    for j in jobs:
        j["dst"].zero_()
    b.launch()
    torch.cuda.synchronize()
    for k, j in enumerate(jobs):
        assert np.array_equal(j["dst"].cpu().numpy(), ref[k])
    for rows in ((2, 2), (18, 6)):
        b.set_strip_rows(*rows)
        for j in jobs:
            j["dst"].zero_()
        b.launch()
        torch.cuda.synchronize()
        for k, j in enumerate(jobs):
            assert np.array_equal(j["dst"].cpu().numpy(), ref[k]), rows


def test_rejections(ctx):
    import torch
    mos = synth.mosaic(64, 48, "RGGB", seed=1)
    d_in = torch.from_numpy(mos.view(np.int16)).to("cuda")
    dst = torch.zeros(64 * 48 * 4, dtype=torch.uint8, device="cuda")
    frame = F.GpuFrame.from_u16(d_in, 64, 48)
    spec = F.DevelopSpec(F.Rect.full(64, 48))
    deltas = F.FrameInfo(64, 48, F.CfaPattern.RGGB, F.BlackLevel(1, 1, 1, (512.0,), (1.0,) * 48, ()),
                         F.WhiteLevel.uniform(16383))
    with pytest.raises(F.NotAvailable):
        ctx.batch([(frame, deltas, spec, dst)])
    plain = F.FrameInfo(64, 48, F.CfaPattern.RGGB, F.BlackLevel.uniform(512.0), F.WhiteLevel.uniform(16383))
    with pytest.raises(F.InvalidArgument):
        ctx.batch([(F.GpuFrame.from_u16(mos, 64, 48), plain, spec, dst)])
    with pytest.raises(F.NotAvailable):
        ctx.batch([(frame, plain, F.DevelopSpec(F.Rect.full(64, 48), output2=F.DevelopOutput.Rgba8), dst)])


def test_many_small_frames(ctx):
    """More frames than the list words of a small batch: the lists hold 2n + 8 words."""
    import torch
    jobs = _jobs(6, 40, max_side=120)
    b = _batch(ctx, jobs)
    b.launch()
    torch.cuda.synchronize()
    _check(ctx, jobs, "many")
