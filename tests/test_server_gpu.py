"""Serving loop: jobs admitted by device memory, batched continuously, copied back. Each job's
output must equal its own develop (and, for denoised jobs, the denoiser followed by the develop)
alone, whatever else is in flight; admission must respect the device budget."""
import numpy as np
import pytest

import flashraw_amd as F
from flashraw_amd import synth

pytestmark = pytest.mark.gpu

PHASES = ["RGGB", "BGGR", "GRBG", "GBRG"]


def _jobs(seed, n, max_side=640):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        w = int(rng.integers(5, max_side // 8)) * 8
        h = int(rng.integers(4, max_side // 8)) * 8
        cfa = PHASES[int(rng.integers(0, 4))]
        bits = 12 if rng.random() < 0.5 else 14
        white = (1 << bits) - 1
        black = 64.0 * (bits - 10)
        raster = synth.mosaic(w + 8, h + 6, cfa, black=black, white=float(white), seed=seed * 1000 + k)
        win = F.Rect(int(rng.integers(0, 8)), int(rng.integers(0, 6)), w, h)
        info = F.FrameInfo(w, h, F.CfaPattern.bayer(cfa), F.BlackLevel.uniform(black), F.WhiteLevel.uniform(white))
        origin = (win.x % 2, win.y % 2)  # sensor position of the window: the phase seen by the window
        ahd = rng.random() < 0.4
        spec = F.DevelopSpec(F.Rect.full(w, h), wb=(float(rng.uniform(1.2, 2.8)), 1.0, float(rng.uniform(1.2, 2.8)), 1.0),
                             cam_to_rgb=tuple(tuple(float(v) for v in r) for r in np.asarray(synth.MATRIX) *
                                              rng.uniform(0.9, 1.1, (3, 3))),
                             output=F.DevelopOutput.Rgba8, edge=F.ahd_spec(info, origin) if ahd else None)
        dn = None
        if rng.random() < 0.3:
            dn = F.CfaDenoiseSpec(w, h, [black] * 4, 2.0, 144.0, 2.7, 4, F.DnPrecision.F16X3)
        frame = F.GpuFrame.from_u16(raster, w + 8, h + 6).with_window(win).with_origin(origin)
        out.append(dict(frame=frame, info=info, spec=spec, dn=dn, raster=raster, win=win, origin=origin, w=w, h=h))
    return out


def _alone(ctx, j):
    """The job developed by itself: the single-frame path (after the denoiser when asked)."""
    import torch
    win = j["win"]
    sub = np.ascontiguousarray(j["raster"][win.y:win.y + j["h"], win.x:win.x + j["w"]])
    if j["dn"] is not None:
        src = torch.from_numpy(sub.view(np.int16)).to("cuda")
        dst = torch.zeros_like(src)
        ctx.denoise_cfa(j["dn"], src, dst)
        torch.cuda.synchronize()
        sub = dst.cpu().numpy().view(np.uint16)
    run = ctx.run(F.GpuFrame.from_u16(sub, j["w"], j["h"]).with_origin(j["origin"]), j["info"],
                  F.GpuOp.Develop(j["spec"]))
    return run.image.download()


def _serve(srv, jobs):
    outs = [np.zeros((j["h"], j["w"], 4), dtype=np.uint8) for j in jobs]
    tickets = [srv.submit(j["frame"], j["info"], j["spec"], o, j["dn"]) for j, o in zip(jobs, outs)]
    results = [srv.wait(t, 60000) for t in tickets]
    return outs, results


def test_mixed_jobs_equal_their_own_develop(ctx):
    jobs = _jobs(1, 16)
    srv = ctx.server()
    outs, results = _serve(srv, jobs)
    st = srv.stats()
    srv.close()
    for k, (j, o, r) in enumerate(zip(jobs, outs, results)):
        assert r is not None and r.status == 0, (k, r)
        assert r.latency_ms > 0 and r.batch_frames >= 1
        assert np.array_equal(o, _alone(ctx, j)), (k, j["w"], j["h"], j["dn"] is not None)
    assert st["jobs_done"] == len(jobs) and st["jobs_failed"] == 0 and st["batches"] >= 1


def test_alone_through_the_server_equals_the_mix(ctx):
    jobs = _jobs(2, 10, max_side=400)
    srv = ctx.server()
    mix, _ = _serve(srv, jobs)
    for j, m in zip(jobs[:4], mix[:4]):
        alone, res = _serve(srv, [j])
        assert res[0].status == 0 and res[0].batch_frames == 1
        assert np.array_equal(alone[0], m)
    srv.close()


def test_admission_respects_the_budget(ctx):
    jobs = _jobs(3, 12, max_side=512)
    need = [j["w"] * j["h"] * (2 * (2 if j["dn"] else 1) + 4) for j in jobs]
    budget = 2 * max(need)
    srv = ctx.server(device_budget_bytes=budget)
    outs, results = _serve(srv, jobs)
    st = srv.stats()
    srv.close()
    assert all(r.status == 0 for r in results)
    assert 0 < st["peak_device_bytes"] <= budget
    for j, o in zip(jobs, outs):
        assert np.array_equal(o, _alone(ctx, j))


def test_job_larger_than_budget_fails_alone(ctx):
    jobs = _jobs(4, 3, max_side=256)
    small = min(j["w"] * j["h"] for j in jobs) * 6 + 1
    srv = ctx.server(device_budget_bytes=small)
    big = max(range(3), key=lambda k: jobs[k]["w"] * jobs[k]["h"] * (8 if jobs[k]["dn"] else 6))
    outs, results = _serve(srv, jobs)
    srv.close()
    assert results[big].status == 8  # FRH_ERROR_LIMIT_EXCEEDED


def test_submit_rejections(ctx):
    import torch
    srv = ctx.server()
    mos = synth.mosaic(60, 40, "RGGB", seed=1)
    info = F.FrameInfo(60, 40, F.CfaPattern.RGGB, F.BlackLevel.uniform(512.0), F.WhiteLevel.uniform(16383))
    spec = F.DevelopSpec(F.Rect.full(60, 40))
    out = np.zeros((40, 60, 4), dtype=np.uint8)
    with pytest.raises(F.NotAvailable):  # the MFMA denoiser needs sides that are multiples of 8
        srv.submit(F.GpuFrame.from_u16(mos, 60, 40), info, spec, out, F.CfaDenoiseSpec(60, 40, [512.0] * 4))
    d = torch.from_numpy(mos.view(np.int16)).to("cuda")
    with pytest.raises(F.InvalidArgument):
        srv.submit(F.GpuFrame.from_u16(d, 60, 40), info, spec, out)
    with pytest.raises(F.InvalidArgument):
        srv.submit(F.GpuFrame.from_u16(mos, 60, 40), info, spec, np.zeros(10, dtype=np.uint8))
    srv.close()
