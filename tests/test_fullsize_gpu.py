"""Full-size parity against the oracle, at the sizes the workloads run at.

- The batched workload: 64 frames of 12-61 MP (every Bayer phase, 12/14 bits, own WB and
  matrix), AHD to RGBA8, autotuned, captured in a HIP graph and replayed; every byte of every
  frame equals the oracle's develop of that frame (4.1 GB of mosaics kept on the host, one
  frame's output and oracle output at a time). The model is the reference's banded-equals-whole check
  (crates/flashraw-gpu/tests/equivalence.rs:477-559): size-dependent code (work lists, strip
  plans, descriptors) must not change a byte.
- The served workload: 12 requests drawn like the serving benchmark's (12-102 MP log-uniform,
  30 % AHD, 20 % denoised by the CFA-plane denoiser first). Each output equals the oracle's
  develop of its input (of the device denoiser's output for denoised requests) and its develop
  alone.
- The luma / chroma noise reduction at 24 MP and in rows wider than one gather chunk, within the
  reference's GPU-vs-CPU contract (equivalence.rs:960-1003) against the oracle whose guided-filter
  box means are summed as the reference's GPU kernels sum them, and within a stated bound against
  the CPU path's running sums.
"""
import numpy as np
import pytest
import torch

import flashraw_amd as F
from flashraw_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(600)
def test_batch_64_frames_12_61mp_every_byte(ctx):
    frames = synth.batch_frames(64)
    jobs, keep, mosaics, total = [], [], [], 0
    for f in frames:
        mos = synth.batch_mosaic(f)
        mosaics.append(mos)
        white = (1 << f["bits"]) - 1
        black = 64.0 * (f["bits"] - 10)
        info = F.FrameInfo(f["w"], f["h"], F.CfaPattern.bayer(f["cfa"]), F.BlackLevel.uniform(black),
                           F.WhiteLevel.uniform(white))
        spec = F.DevelopSpec(F.Rect.full(f["w"], f["h"]), wb=f["wb"], cam_to_rgb=synth.jittered_matrix(f["jitter"]),
                             output=F.DevelopOutput.Rgba8, edge=F.ahd_spec(info))
        d_in = torch.from_numpy(mos.view(np.int16)).to("cuda")
        dst = torch.empty((f["h"], f["w"], 4), dtype=torch.uint8, device="cuda")
        jobs.append((F.GpuFrame.from_u16(d_in, f["w"], f["h"]), info, spec, dst))
        keep.append(d_in)
        total += f["w"] * f["h"]
    print(f"{len(jobs)} frames, {total / 1e6:.0f} MP", flush=True)
    assert len(jobs) == 64 and total > 1.5e9
    stream = torch.cuda.Stream()
    batch = ctx.batch(jobs)
    rows_ahd, _ = batch.autotune(stream=stream, iters=1)
    batch.capture(stream)
    torch.cuda.synchronize()
    for _, _, _, dst in jobs:  # the autotuner's launches wrote the outputs: clear them first
        dst.fill_(0x5A)
    torch.cuda.synchronize()
    batch.launch(stream=stream)  # graph replay
    torch.cuda.synchronize()
    bad = []
    for k, (f, mos, (_, _, spec, dst)) in enumerate(zip(frames, mosaics, jobs)):
        want = O.develop(mos, f["cfa"], black=(1, 1, 1, [64.0 * (f["bits"] - 10)]), white=((1 << f["bits"]) - 1,),
                         wb=f["wb"], matrix=spec.cam_to_rgb, demosaic="ahd", encoding="rgba8")
        got = dst.cpu().numpy()
        n = int(np.count_nonzero(got != want))
        if n:
            bad.append((k, f["w"], f["h"], f["cfa"], f["bits"], n))
        del want, got
    print(f"autotuned AHD strip rows {rows_ahd}; frames differing: {bad}")
    assert not bad


def _serve_jobs(n, seed):
    rng = np.random.default_rng(seed)
    shapes = synth.serve_mosaic_shapes(seed=seed, n=n)
    out = []
    for k, sh in enumerate(shapes):
        white = (1 << sh["bits"]) - 1
        black = 64.0 * (sh["bits"] - 10)
        mos = F.pinned_empty((sh["h"], sh["w"]), np.uint16)
        mos[:] = synth.mosaic(sh["w"], sh["h"], sh["cfa"], black=black, white=float(white), seed=sh["seed"])
        ahd = bool(rng.random() < 0.3)
        denoise = bool(rng.random() < 0.2)
        if k == 0:
            denoise = True  # at least one denoised and one AHD request
        if k == 1:
            ahd = True
        wb = (float(rng.uniform(1.2, 2.8)), 1.0, float(rng.uniform(1.2, 2.8)), 1.0)
        matrix = synth.jittered_matrix(rng.uniform(0.9, 1.1, (3, 3)))
        info = F.FrameInfo(sh["w"], sh["h"], F.CfaPattern.bayer(sh["cfa"]), F.BlackLevel.uniform(black),
                           F.WhiteLevel.uniform(white))
        spec = F.DevelopSpec(F.Rect.full(sh["w"], sh["h"]), wb=wb, cam_to_rgb=matrix, output=F.DevelopOutput.Rgba8,
                             edge=F.ahd_spec(info) if ahd else None)
        dn = F.CfaDenoiseSpec(sh["w"], sh["h"], [black] * 4, 2.0, 144.0, 2.7, 4, F.DnPrecision.F16X3) \
            if denoise else None
        out.append(dict(sh, mos=mos, black=black, white=white, ahd=ahd, dn=dn, wb=wb, matrix=matrix, info=info,
                        spec=spec))
    return out


@pytest.mark.timeout(600)
def test_served_requests_12_102mp_equal_the_oracle_and_alone(ctx):
    jobs = _serve_jobs(12, seed=21)
    mp = sum(j["w"] * j["h"] for j in jobs) / 1e6
    print(f"{len(jobs)} requests, {mp:.0f} MP, sizes {[round(j['w'] * j['h'] / 1e6, 1) for j in jobs]}, "
          f"AHD {sum(j['ahd'] for j in jobs)}, denoised {sum(j['dn'] is not None for j in jobs)}", flush=True)
    assert min(j["w"] * j["h"] for j in jobs) >= 11.5e6 and max(j["w"] * j["h"] for j in jobs) <= 102.5e6
    srv = ctx.server(device_budget_bytes=64 << 30)
    outs = [F.pinned_empty((j["h"], j["w"], 4), np.uint8) for j in jobs]
    tickets = [srv.submit(F.GpuFrame.from_u16(j["mos"], j["w"], j["h"]), j["info"], j["spec"], o, j["dn"])
               for j, o in zip(jobs, outs)]
    results = [srv.wait(t, 120000) for t in tickets]
    st = srv.stats()
    srv.close()
    print(f"server: {st['batches']} batches, {st['batched_frames']} frames, peak {st['peak_device_bytes'] / 1e9:.1f} GB")
    for k, (j, o, r) in enumerate(zip(jobs, outs, results)):
        assert r is not None and r.status == 0, (k, r)
        src = j["mos"]
        if j["dn"] is not None:
            d_src = torch.from_numpy(np.ascontiguousarray(src).view(np.int16)).to("cuda")
            d_dst = torch.zeros_like(d_src)
            ctx.denoise_cfa(j["dn"], d_src, d_dst)
            torch.cuda.synchronize()
            src = d_dst.cpu().numpy().view(np.uint16)
            assert not np.array_equal(src, j["mos"])  # the denoiser did change the mosaic
        want = O.develop(src, j["cfa"], black=(1, 1, 1, [j["black"]]), white=(j["white"],), wb=j["wb"],
                         matrix=j["matrix"], demosaic="ahd" if j["ahd"] else "bilinear", encoding="rgba8")
        n = int(np.count_nonzero(o != want))
        assert n == 0, (k, j["w"], j["h"], j["cfa"], j["ahd"], j["dn"] is not None, n)
        alone = ctx.run(F.GpuFrame.from_u16(src, j["w"], j["h"]), j["info"], F.GpuOp.Develop(j["spec"]))
        assert np.array_equal(alone.image.download(), o), k


def _nr_stats(a, b):
    d = np.abs(a.astype(np.float64) - b.astype(np.float64)).reshape(-1)
    return d.mean(), np.percentile(d, 99.9), d.max()


def _nr_case(ctx, w, h, luma, chroma, mos=None):
    if mos is None:
        mos = synth.mosaic(w, h, "RGGB", black=256.0, white=4095.0, gain=0.6, read_sigma=14.0, edges=40, discs=12,
                           seed=w + h)
        info = F.FrameInfo(w, h, F.CfaPattern.RGGB, F.BlackLevel.uniform(256.0), F.WhiteLevel.uniform(4095))
    else:
        info = F.FrameInfo(w, h, F.CfaPattern.RGGB, F.BlackLevel.uniform(512.0), F.WhiteLevel.uniform(16383))
    spec = F.DevelopSpec(F.Rect.full(w, h), wb=synth.WB, cam_to_rgb=synth.MATRIX, output=F.DevelopOutput.LinearRgbF32)
    img = ctx.run(F.GpuFrame.from_u16(mos, w, h), info, F.GpuOp.Develop(spec)).image
    lin = img.download()
    got = ctx.denoise(img, F.denoise_spec(w, h, luma, chroma)).image.download()
    assert got.shape == (h, w, 3)
    # the oracle with the guided filter's box means summed as the reference's GPU kernels sum them
    # (direct window sums, denoise.wgsl:191-218): the reference's own contract
    direct = O.nr_apply(lin, luma, chroma, box_direct=True)
    mean, p999, mx = _nr_stats(got, direct)
    moved = _nr_stats(direct, lin)[0]
    print(f"{w}x{h} luma {luma} chroma {chroma} vs direct-sum oracle: mean {mean:.2e} p99.9 {p999:.2e} "
          f"max {mx:.2e} (moves {moved:.2e})")
    assert mean <= 2e-6 and p999 <= 1e-4, (w, h, luma, chroma, mean, p999)
    assert moved > 1e-3
    # the oracle with the CPU path's running sums (mod.rs:280-335), whose rounding drifts along row
    # segments of w / 8 samples at half resolution: the reference's 99.9th-percentile bound, and a
    # mean bound of 1.5x the 4.3e-6 measured at 24 MP (the reference states 2e-6 for its 347x229
    # test frame, whose segments are 43 samples long)
    want = O.nr_apply(lin, luma, chroma)
    mean, p999, mx = _nr_stats(got, want)
    print(f"{w}x{h} luma {luma} chroma {chroma} vs running-sum oracle: mean {mean:.2e} p99.9 {p999:.2e} "
          f"max {mx:.2e}")
    assert mean <= 6.5e-6 and p999 <= 1e-4, (w, h, luma, chroma, mean, p999)


@pytest.mark.parametrize("w,h", [(2500, 200), (4100, 96)])
def test_noise_reduction_rows_wider_than_a_gather_chunk(ctx, w, h):
    """Rows of 3 and 5 chunks of 1024 samples in the gather, and tile grids wider than 6 tiles."""
    _nr_case(ctx, w, h, 0.5, 0.5)


@pytest.mark.timeout(300)
def test_noise_reduction_24mp(ctx):
    """The 6048x4024 read-sigma-12 frame developed to linear RGB, strengths 0.5 / 0.5."""
    _nr_case(ctx, 6048, 4024, 0.5, 0.5, mos=synth.frame_24mp_noisy())
