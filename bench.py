"""Throughput of the fused RAW->sRGB develop on one MI355X.

Default workload: one 24 MP frame (6048x4024 RGGB, 14-bit, black 512, white 16383, seeded
synthetic scene with Poisson-Gaussian noise), AHD demosaic, WB (2.1, 1, 1.6), a fixed
camera->sRGB matrix, RGBA8 output. A step is one develop of that frame with its input already
resident in HBM and the output written to HBM. Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--demosaic ahd|bilinear] [--strip-rows R]
  python bench.py --mode copy   # a plain device copy of 1 GiB (PMC byte-counter calibration)
"""
import argparse
import json
import os
import re
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--demosaic", default="ahd", choices=["ahd", "bilinear"])
    p.add_argument("--mode", default="develop", choices=["develop", "copy"])
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--strip-rows", type=int, default=0,
                   help="use this strip height instead of autotuning (profiled runs, where the autotuner's "
                        "own launches would mix into the per-kernel statistics)")
    return p.parse_args()


def _traffic(kernel_tag):
    """HBM bytes per launch from the committed PMC summary, when there is one."""
    path = os.path.join(ROOT, "profiles", f"pmc_{kernel_tag}.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def copy_mode(args):
    import torch
    n = 1 << 30
    a = torch.empty(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 10
    print(json.dumps({"mode": "copy", "bytes_each_way": n, "ms": dt * 1e3, "GBps": 2 * n / dt / 1e9}))


def main():
    args = _parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (each frame develops on one "
              "GPU; see DESIGN.md for how the path would shard)", file=sys.stderr)
        sys.exit(2)
    if args.mode == "copy":
        return copy_mode(args)

    import numpy as np
    import torch

    import flashraw_amd as F
    from flashraw_amd import synth

    if args.strip_rows:
        os.environ["FLASHRAW_AHD_ROWS" if args.demosaic == "ahd" else "FLASHRAW_BILINEAR_ROWS"] = str(args.strip_rows)
    ctx = F.context()
    W, H = 6048, 4024
    mos = synth.frame_24mp()
    info = F.FrameInfo(W, H, F.CfaPattern.RGGB, F.BlackLevel.uniform(512.0), F.WhiteLevel.uniform(16383))
    edge = F.ahd_spec(info) if args.demosaic == "ahd" else None
    spec = F.DevelopSpec(F.Rect.full(W, H), wb=synth.WB, cam_to_rgb=synth.MATRIX, output=F.DevelopOutput.Rgba8,
                         edge=edge)
    d_in = torch.from_numpy(mos.view(np.int16)).to("cuda")
    dst = torch.empty((H, W, 4), dtype=torch.uint8, device="cuda")
    frame = F.GpuFrame.from_u16(d_in, W, H)
    plan = ctx.prepare(frame, info, spec)
    stream = torch.cuda.Stream()  # the launch stream; the HIP events below are recorded on it
    # one-time setup outside the timed region: the autotuner picks the strip height
    rows = args.strip_rows or plan.autotune(d_in, dst, stream=stream)

    for _ in range(args.warmup):
        plan.launch(d_in, dst, stream=stream)
    torch.cuda.synchronize()
    # Timed region: K back-to-back steps; per-launch HIP events on the launch stream.
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s, e in evs:
        s.record(stream)
        plan.launch(d_in, dst, stream=stream)
        e.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms = [s.elapsed_time(e) for s, e in evs]
    ms_per_step = wall * 1e3 / args.steps
    mp = W * H / 1e6
    value = mp / (ms_per_step / 1e3)
    k_avg = statistics.mean(kms)
    algo_bytes = W * H * (2 + 4)  # u16 in, RGBA8 out
    achieved = algo_bytes / (k_avg / 1e3) / 1e9
    got = dst.cpu().numpy()

    cpu = None
    max_diff = None
    if not args.no_cpu:
        from oracle import oracle as O
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
        want = None
        times = []
        for i in range(6):
            t = time.perf_counter()
            want = O.develop(mos, "RGGB", black=(1, 1, 1, [512.0]), white=(16383,), wb=synth.WB,
                             matrix=synth.MATRIX, demosaic=args.demosaic, encoding="rgba8", threads=threads)
            if i:
                times.append(time.perf_counter() - t)
        cpu_s = statistics.median(times)
        max_diff = int(np.abs(got.astype(np.int32) - want.astype(np.int32)).max())
        cpu = {"value": round(mp / cpu_s, 2), "unit": "MP/s", "cores": threads, "kind": "port",
               "sample": f"the same 6048x4024 frame, median of 5 runs after 1 warm-up, "
                         f"{args.demosaic} RGBA8, C restatement of the reference CPU path"}
    tag = f"develop_{args.demosaic}_6048x4024_rgba8"
    line = {
        "metric": "MP/s RAW->sRGB develop (24 MP AHD, % HBM roofline)",
        "value": round(value, 1),
        "unit": "MP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"6048x4024 RGGB 14-bit u16 -> RGBA8, {args.demosaic} demosaic, resident in HBM",
                   "frame_mp": round(mp, 3), "demosaic": args.demosaic, "output": "RGBA8",
                   "max_abs_diff_vs_oracle_lsb": max_diff, "strip_rows": rows, "strip_rows_autotuned": not args.strip_rows},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": _traffic(tag),
                     "kernel_ms_avg": round(k_avg, 5), "kernel_ms_p50": round(statistics.median(kms), 5),
                     "algorithmic_bytes_per_launch": algo_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
