"""GPU luma / chroma noise reduction (GpuContext::denoise) against the CPU oracle of
crates/flashraw/src/denoise/mod.rs, through the C ABI.

The contract is the reference's own (crates/flashraw-gpu/tests/equivalence.rs:960-1003,
gpu_noise_reduction_matches_cpu): not bit-identical, because a coefficient on the hard threshold
can flip with the rounding of the transforms and the overlap-add sums run in another order. On
linear RGB: mean |diff| <= 2e-6 and 99.9th percentile <= 1e-4 for the strengths (0.5, 0), (0, 0.5),
(0.2, 0), (1, 1), (0.5, 0.3); on a toned, rotated RGBA8 output every sample within 3 codes and at
most 0.1 % more than 1 code off. The noise models themselves are computed from blocks measured the
same way, so they must equal the oracle's bit for bit (checked through the library's
Model::from_stats).
"""
import numpy as np
import pytest

import flashraw_amd as F
import np_restatement as N
from flashraw_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MATRIX = ((1.6, -0.35, -0.25), (-0.2, 1.45, -0.25), (0.05, -0.45, 1.4))
STRENGTHS = [(0.5, 0.0), (0.0, 0.5), (0.2, 0.0), (1.0, 1.0), (0.5, 0.3)]


def _noisy_develop(ctx, w, h, seed=3):
    """A developed linear RGB image of a noisy mosaic (noisy_develop of the reference's tests)."""
    mos = synth.mosaic(w, h, "RGGB", black=256.0, white=4095.0, gain=0.6, read_sigma=14.0, edges=25, discs=8,
                       seed=seed)
    info = F.FrameInfo(w, h, F.CfaPattern.RGGB, F.BlackLevel.uniform(256.0), F.WhiteLevel.uniform(4095))
    spec = F.DevelopSpec(F.Rect.full(w, h), wb=(2.0, 1.0, 1.6, 1.0), cam_to_rgb=MATRIX,
                         output=F.DevelopOutput.LinearRgbF32)
    run = ctx.run(F.GpuFrame.from_u16(mos, w, h), info, F.GpuOp.Develop(spec))
    return run.image


def _stats(a, b):
    d = np.abs(a.astype(np.float64) - b.astype(np.float64)).reshape(-1)
    return d.mean(), np.percentile(d, 99.9), d.max()


@pytest.mark.parametrize("luma,chroma", STRENGTHS)
def test_matches_the_oracle_on_linear_rgb(ctx, luma, chroma):
    w, h = 347, 229
    img = _noisy_develop(ctx, w, h)
    lin = img.download()
    spec = F.denoise_spec(w, h, luma, chroma)
    run = ctx.denoise(img, spec)
    got = run.image.download()
    want = O.nr_apply(lin, luma, chroma)
    mean, p999, mx = _stats(got, want)
    change = _stats(want, lin)[0]
    print(f"luma {luma} chroma {chroma}: mean {mean:.2e} p99.9 {p999:.2e} max {mx:.2e} (moves {change:.2e})")
    assert got.shape == (h, w, 3)
    assert mean <= 2e-6 and p999 <= 1e-4, (luma, chroma, mean, p999)
    assert change > 1e-3  # the noise reduction does change the image


def test_display_output_toned_and_rotated(ctx):
    """The reference test's display case: strengths (0.5, 0.5), the standard tone, Rotate90."""
    w, h = 347, 229
    img = _noisy_develop(ctx, w, h, seed=5)
    lin = img.download()
    tone = N.synthetic_tone(ev=0.0)
    gtone = F.GpuTone(tone["table"], float(tone["xmax"]), float(tone["scale"]), tone["to_prophoto"], tone["white"])
    spec = F.denoise_spec(w, h, 0.5, 0.5, F.DevelopOutput.Rgba8)
    spec.orientation = F.Orientation.Rotate90
    spec.tone = gtone
    got = ctx.denoise(img, spec).image.download()
    den = O.nr_apply(lin, 0.5, 0.5)
    toned = N.tone_apply(den, tone["table"], tone["xmax"], tone["scale"], tone["to_prophoto"], tone["white"])
    want = N.orient(N.rgba8(toned, O.srgb8_table()), int(F.Orientation.Rotate90))
    assert got.shape == want.shape == (w, h, 4)
    d = np.abs(got.astype(np.int64) - want.astype(np.int64))
    over = int(np.count_nonzero(d > 1))
    print(f"rgba8: {over} of {d.size} samples more than 1 code off, at most {d.max()}")
    assert d.max() <= 3 and over * 1000 <= d.size


def test_every_encoding_follows_the_linear_result(ctx):
    """The other encodings are the develop's encoders of the denoised linear RGB (denoise_encode)."""
    w, h = 160, 120
    img = _noisy_develop(ctx, w, h, seed=8)
    lin_out = ctx.denoise(img, F.denoise_spec(w, h, 0.7, 0.6)).image.download()
    for out, enc in ((F.DevelopOutput.LinearRgb16, N.linear16), (F.DevelopOutput.Rgba8, None)):
        got = ctx.denoise(img, F.denoise_spec(w, h, 0.7, 0.6, out)).image.download()
        want = enc(lin_out) if enc else N.rgba8(lin_out, O.srgb8_table())
        assert np.array_equal(got, want), out


def test_models_equal_the_oracle(ctx):
    """The blocks measured on the device give the oracle's model bit for bit."""
    w, h = 300, 200
    rng = np.random.default_rng(4)
    y = (0.3 + 0.02 * rng.standard_normal((h, w))).astype(np.float32)
    want = O.nr_model_from_stats(O.nr_stats(y, y))
    got = F.nr_model_from_stats(O.nr_stats(y, y))
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


@pytest.mark.parametrize("n", [20, 40, 500, 1153, 5000, 42336, 55300])
def test_device_fit_equals_the_oracle(ctx, n):
    """Model::from_stats fitted on the device (the rank selections GpuContext.denoise runs) equals
    the oracle's bit for bit: below 32 blocks there is no model; 1153 blocks make 12 bins of 97 (the
    last shorter); 55300 blocks exceed what the fit kernels hold per bin and take the host form.
    Repeated means (ties broken by block order), zero coefficients and a few non-finite blocks are
    mixed in."""
    rng = np.random.default_rng(n)
    st = (rng.standard_normal((n, 64)) ** 2).astype(np.float32) * 1e-4
    st[:, 0] = rng.choice(np.linspace(0.05, 0.9, 37), n).astype(np.float32)  # many equal means
    st[rng.random((n, 64)) < 0.01] = 0.0
    if n > 100:
        st[3, 0] = np.nan
        st[7, 9] = np.inf
    want = O.nr_model_from_stats(st)
    got = F.nr_model_fit(ctx, st)
    if want is None:
        assert got is None
        return
    assert got is not None
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


def test_resident_form_equals_the_image_form(ctx):
    import torch
    w, h = 203, 150
    img = _noisy_develop(ctx, w, h, seed=9)
    lin = torch.from_numpy(img.download()).to("cuda")
    spec = F.denoise_spec(w, h, 0.6, 0.4, F.DevelopOutput.Rgba8)
    dst = torch.zeros(h * w * 4, dtype=torch.uint8, device="cuda")
    ctx.denoise_into(lin, w, h, spec, dst)
    torch.cuda.synchronize()
    want = ctx.denoise(img, spec).image.download()
    assert np.array_equal(dst.cpu().numpy().reshape(h, w, 4), want)


def test_deterministic_and_small_frames(ctx):
    w, h = 129, 97
    img = _noisy_develop(ctx, w, h, seed=2)
    spec = F.denoise_spec(w, h, 1.0, 1.0)
    a = ctx.denoise(img, spec).image.download()
    b = ctx.denoise(img, spec).image.download()
    assert np.array_equal(a, b)
    # below 8 samples a side nothing is denoised: the split and merge only
    tiny = _noisy_develop(ctx, 7, 30, seed=1)
    lin = tiny.download()
    got = ctx.denoise(tiny, F.denoise_spec(7, 30, 1.0, 1.0)).image.download()
    assert np.abs(got - lin).max() < 1e-6
    # zero strengths: no plane is denoised
    z = ctx.denoise(img, F.denoise_spec(w, h, 0.0, 0.0)).image.download()
    assert np.abs(z - img.download()).max() < 1e-6


def test_errors(ctx):
    w, h = 64, 48
    img = _noisy_develop(ctx, w, h)
    spec = F.denoise_spec(w, h, 0.5, 0.5)
    import dataclasses
    bad_steps = dataclasses.replace(spec, luma=dataclasses.replace(spec.luma, steps=(3, 2)))
    with pytest.raises(F.InvalidArgument):
        ctx.denoise(img, bad_steps)
    with pytest.raises(F.InvalidArgument):
        ctx.denoise(img, dataclasses.replace(spec, output=99))
    # the input must be linear f32 RGB
    mos = synth.mosaic(w, h, "RGGB", black=256.0, white=4095.0, seed=1)
    info = F.FrameInfo(w, h, F.CfaPattern.RGGB, F.BlackLevel.uniform(256.0), F.WhiteLevel.uniform(4095))
    rgba = ctx.run(F.GpuFrame.from_u16(mos, w, h), info,
                   F.GpuOp.Develop(F.DevelopSpec(F.Rect.full(w, h)))).image
    with pytest.raises(F.InvalidArgument):
        ctx.denoise(rgba, spec)


def test_megapixel_frame(ctx):
    """A 3 MP frame: the full pyramid (six chroma levels) against the oracle."""
    w, h = 2048, 1536
    img = _noisy_develop(ctx, w, h, seed=12)
    lin = img.download()
    spec = F.denoise_spec(w, h, 0.6, 0.6)
    assert spec.chroma.levels == 6 and spec.luma.levels == 4
    run = ctx.denoise(img, spec)
    mean, p999, mx = _stats(run.image.download(), O.nr_apply(lin, 0.6, 0.6))
    print(f"3 MP: mean {mean:.2e} p99.9 {p999:.2e} max {mx:.2e}, {run.timings.kernel_ms:.2f} ms")
    assert mean <= 2e-6 and p999 <= 1e-4
