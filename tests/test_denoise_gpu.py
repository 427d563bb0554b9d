"""GPU CFA-plane DCT denoiser against the CPU oracle.

The f32 scalar variant performs the oracle's float statement per block, so it matches the
oracle's f32 path to float rounding of the overlap-add. The MFMA variants (f16 and bf16 operands
in the Kronecker-form kernel, split f16 operands in the separable one; f32 accumulation) are
compared with the oracle's float64 statement; their error bounds are stated below in DN on a
14-bit scale (black 512), about 1.5 times the errors measured on MI355X: f16 max 44 DN, mean
0.091 DN on the 24 MP frame (up to 72 / 2.0 DN on small frames dense with edges, where the 16-bit
operands round large local contrast); bf16 max 67 / mean 0.98 (94 / 11.6); split f16 max 0.003
DN except where a coefficient within float rounding of the threshold flips (one pixel of the 24 MP
frame, 12.7 DN), mean 0.0002 DN; scalar f32 max 0.004 DN.
"""
import numpy as np
import pytest

import flashraw_amd as F
from flashraw_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

BLACK = [512.0, 510.0, 514.0, 511.0]
# Error bounds against the float64 oracle on the test frames (DN): max |err|, mean |err|.
BOUNDS = {F.DnPrecision.F16: (100.0, 3.0), F.DnPrecision.BF16: (142.0, 14.0), F.DnPrecision.F16X3: (0.01, 0.002),
          F.DnPrecision.F32: (0.01, 0.002)}
# The 24 MP frame: split f16 may flip a threshold decision on a handful of coefficients.
BOUNDS_24MP = {F.DnPrecision.F16: (66.0, 0.14), F.DnPrecision.BF16: (100.0, 1.5), F.DnPrecision.F16X3: (40.0, 0.001)}
# Threshold 0 (identity) and infinite threshold (block means): max |err| in DN.
BOUNDS_ID = {F.DnPrecision.F16: (10.0, 7.0), F.DnPrecision.BF16: (55.0, 21.0), F.DnPrecision.F16X3: (0.01, 0.01)}
MFMA = [F.DnPrecision.F16, F.DnPrecision.BF16, F.DnPrecision.F16X3]


def _frame(w, h, seed, sigma=12.0):
    return synth.mosaic(w, h, "RGGB", read_sigma=sigma, seed=seed)


def _run(ctx, mos, precision, sig_a=2.0, sig_b=144.0, thr=2.7, step=4, seg_rows=0):
    import torch
    h, w = mos.shape
    src = torch.from_numpy(mos.view(np.int16)).to("cuda")
    dst = torch.zeros((h, w), dtype=torch.int16, device="cuda")
    dstf = torch.zeros((h, w), dtype=torch.float32, device="cuda")
    spec = F.CfaDenoiseSpec(w, h, BLACK, sig_a, sig_b, thr, step, precision, seg_rows=seg_rows)
    ctx.denoise_cfa(spec, src, dst, dstf)
    torch.cuda.synchronize()
    return dst.cpu().numpy().view(np.uint16), dstf.cpu().numpy()


def _stats(got, want):
    d = np.abs(got.astype(np.float64) - want.astype(np.float64))
    rmse = float(np.sqrt(np.mean(d * d)))
    psnr = 20 * np.log10((16383 - 512) / rmse) if rmse > 0 else float("inf")
    return float(d.max()), float(d.mean()), psnr


@pytest.mark.parametrize("precision", MFMA + [F.DnPrecision.F32])
@pytest.mark.parametrize("size", [(64, 48), (80, 56), (72, 40), (256, 200), (1000, 712)])
def test_against_float64_oracle(ctx, precision, size):
    w, h = size
    mos = _frame(w, h, seed=w + h)
    u, f = _run(ctx, mos, precision)
    of, ou = O.dct_denoise(mos, BLACK, 2.0, 144.0, 2.7, 4, precision=64)
    mx, mean, psnr = _stats(f, of)
    print(f"{precision.name} {w}x{h}: max {mx:.4f} DN, mean {mean:.5f} DN, PSNR {psnr:.1f} dB vs float64")
    bmax, bmean = BOUNDS[precision]
    assert mx <= bmax and mean <= bmean, (mx, mean, psnr)
    # the u16 output is the f32 output with black added, rounded and clamped
    want_u = np.clip(np.floor(f.astype(np.float64) + np.tile(np.array(BLACK).reshape(2, 2), (h // 2, w // 2)) + 0.5),
                     0, 65535)
    assert np.abs(u.astype(np.int64) - want_u.astype(np.int64)).max() <= 1


def test_f32_matches_oracle_f32_statement(ctx):
    """Per-block values are the oracle's f32 ones; only the overlap-add (f32 here, f64 in the
    oracle) differs, by float rounding."""
    mos = _frame(96, 80, seed=5)
    for step in (2, 3, 4, 5, 8):
        _, f = _run(ctx, mos, F.DnPrecision.F32, step=step)
        of, _ = O.dct_denoise(mos, BLACK, 2.0, 144.0, 2.7, step, precision=32)
        assert np.all(np.abs(f - of) <= 4 * np.spacing(np.abs(of)) + 1e-6), step


def test_golden_fixture(ctx):
    import pathlib
    g = np.load(pathlib.Path(__file__).parent / "golden" / "dct_small.npz")
    a, b, thr, step = (float(v) for v in g["params"])
    mos = g["mosaic"]
    global BLACK
    saved, BLACK = BLACK, [float(v) for v in g["black"]]
    try:
        for precision in MFMA + [F.DnPrecision.F32]:
            _, f = _run(ctx, mos, precision, a, b, thr, int(step))
            mx, mean, _ = _stats(f, g["out_f32"])
            assert mx <= BOUNDS[precision][0] and mean <= BOUNDS[precision][1], (precision, mx, mean)
    finally:
        BLACK = saved


@pytest.mark.parametrize("precision", MFMA)
def test_identity_and_dc_only(ctx, precision):
    """Threshold 0 keeps every coefficient (the planes come back, to operand rounding);
    an infinite threshold keeps DC only (overlap-added block means)."""
    mos = _frame(128, 96, seed=3)
    _, f = _run(ctx, mos, precision, thr=0.0)
    planes = mos.astype(np.float64) - np.tile(np.array(BLACK).reshape(2, 2), (48, 64))
    e0 = float(np.abs(f - planes).max())
    _, f = _run(ctx, mos, precision, thr=1e30)
    of, _ = O.dct_denoise(mos, BLACK, 2.0, 144.0, 1e30, 4, precision=64)
    e1 = float(np.abs(f - of).max())
    print(f"{precision.name}: identity max {e0:.3f} DN, DC-only max {e1:.3f} DN")
    assert e0 <= BOUNDS_ID[precision][0] and e1 <= BOUNDS_ID[precision][1], (e0, e1)


def test_segments_do_not_change_the_result(ctx):
    """The wave decomposition (tile rows per wave) does not change a bit."""
    mos = _frame(512, 400, seed=9)
    for precision in MFMA:
        ref = _run(ctx, mos, precision, seg_rows=16)[1]
        for seg in (1, 3, 7, 64):
            assert np.array_equal(_run(ctx, mos, precision, seg_rows=seg)[1], ref), (precision, seg)


def test_noisy_24mp_frame(ctx):
    """The 6048x4024 RGGB frame with read sigma 12 DN: the MFMA variants against float64."""
    mos = synth.frame_24mp_noisy()
    of, _ = O.dct_denoise(mos, BLACK, 2.0, 144.0, 2.7, 4, precision=64)
    for precision in MFMA:
        _, f = _run(ctx, mos, precision)
        mx, mean, psnr = _stats(f, of)
        print(f"24 MP {precision.name}: max {mx:.3f} DN, mean {mean:.5f} DN, PSNR {psnr:.1f} dB vs float64")
        assert mx <= BOUNDS_24MP[precision][0] and mean <= BOUNDS_24MP[precision][1], (precision, mx, mean, psnr)


def test_rejects_unsupported(ctx):
    import torch
    src = torch.zeros((60, 60), dtype=torch.int16, device="cuda")
    dst = torch.zeros_like(src)
    with pytest.raises(F.NotAvailable):
        ctx.denoise_cfa(F.CfaDenoiseSpec(60, 60, step=4), src, dst)  # sides not multiples of 8
    with pytest.raises(F.NotAvailable):
        ctx.denoise_cfa(F.CfaDenoiseSpec(64, 48, step=2), src, dst)
    with pytest.raises(F.InvalidArgument):
        ctx.denoise_cfa(F.CfaDenoiseSpec(64, 48), src)  # no output
    with pytest.raises(F.InvalidArgument):
        ctx.denoise_cfa(F.CfaDenoiseSpec(14, 48, precision=F.DnPrecision.F32), src, dst)
