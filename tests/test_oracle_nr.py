"""Oracle pinning for the luma / chroma noise reduction of developed linear RGB
(crates/flashraw/src/denoise/mod.rs, restated in oracle/nr.c).

Each piece is checked against what the mathematics or a library fixes: the 2-D DCT against
scipy's orthonormal DCT-II (and the reference's dct_is_orthonormal_and_inverts, mod.rs:1149-1163),
box_mean against scipy's uniform filter with repeated edges, down / the bilinear upsampling against
their closed forms, the noise estimate against white noise of known deviation (the reference's
the_estimate_finds_the_deviation_of_white_noise, mod.rs:1165-1178), Model::from_stats against
constant-block closed forms and a numpy statement of its bins and medians, one DCT pass against its
limits (threshold 0 keeps every block: the identity; a huge threshold keeps the block means; a
Wiener pass with zero noise: the identity), and apply against the reference's
flat_noise_is_removed_and_the_mean_kept (mod.rs:1180-1199) and
zero_strength_changes_nothing_and_edges_survive (mod.rs:1201-1219). A planted mistake (a
transposed coefficient layout) fails the DCT check.
"""
import numpy as np
import pytest
import scipy.fft
import scipy.ndimage

from oracle import oracle as O

F32 = np.float32


def test_dct_block_is_scipy_orthonormal_dct():
    rng = np.random.default_rng(7)
    for _ in range(20):
        b = rng.standard_normal((8, 8)).astype(F32)
        got = O.nr_dct_block(b)
        want = scipy.fft.dctn(b.astype(np.float64), norm="ortho")  # [vertical k, horizontal l]
        assert np.abs(got - want.T).max() < 2e-6
        # a planted transposition of the layout is caught
        assert np.abs(got - want).max() > 1e-3
        # orthonormal: energy kept, and the inverse gives the block back
        assert abs(float((got.astype(np.float64) ** 2).sum() - (b.astype(np.float64) ** 2).sum())) < 1e-3 * float(
            (b.astype(np.float64) ** 2).sum())
        assert np.abs(O.nr_inverse_block(got) - b).max() < 1e-5


def test_dct_horizontal_ramp_has_horizontal_frequencies_only():
    """Coefficient [l][k] is horizontal frequency l, vertical k: a horizontal ramp has only l > 0."""
    b = np.tile(np.arange(8, dtype=F32), (8, 1))
    c = O.nr_dct_block(b)
    assert np.abs(c[1:, 1:]).max() < 1e-5 and np.abs(c[0, 1:]).max() < 1e-5 and np.abs(c[1:, 0]).max() > 1.0


@pytest.mark.parametrize("r", [1, 2, 5])
@pytest.mark.parametrize("shape", [(64, 77), (33, 200), (130, 9)])
def test_box_mean_is_the_uniform_filter(r, shape):
    rng = np.random.default_rng(r)
    p = rng.random(shape).astype(F32)
    got = O.nr_box_mean(p, r)
    want = scipy.ndimage.uniform_filter(p.astype(np.float64), size=2 * r + 1, mode="nearest")
    assert np.abs(got - want).max() < 2e-6


@pytest.mark.parametrize("r", [1, 3])
def test_box_mean_direct_order_is_the_gpu_kernels_sum(r):
    """The direct-sum order (the reference GPU's dn_box_v / dn_box_h, denoise.wgsl:191-218):
    vertical sums over d = -r..r from 0, then horizontal sums of those, divided by (2r + 1)^2 --
    restated sample by sample in float32, bit for bit; within float error of the uniform filter."""
    rng = np.random.default_rng(7 + r)
    p = rng.random((37, 53)).astype(F32)
    h, w = p.shape
    got = O.nr_box_mean(p, r, direct=True)
    vs = np.zeros_like(p)
    for d in range(-r, r + 1):
        vs = (vs + p[np.clip(np.arange(h) + d, 0, h - 1), :]).astype(F32)
    hs = np.zeros_like(p)
    for d in range(-r, r + 1):
        hs = (hs + vs[:, np.clip(np.arange(w) + d, 0, w - 1)]).astype(F32)
    want = (hs / F32((2 * r + 1) ** 2)).astype(F32)
    assert np.array_equal(got, want)
    ref = scipy.ndimage.uniform_filter(p.astype(np.float64), size=2 * r + 1, mode="nearest")
    assert np.abs(got - ref).max() < 2e-6


def test_box_mean_running_sums_drift_along_long_rows():
    """The CPU path's running sums (mod.rs:280-335) drift along a row segment while direct sums do
    not: on a long row of values around 0.5 the running-sum error grows with the segment length.
    The two orders therefore differ by more on wide images, which is why the full-size GPU
    comparison states its mean bound against both."""
    rng = np.random.default_rng(3)
    w = 4096
    p = (0.5 + 0.1 * rng.standard_normal((4, w))).astype(F32)
    ref = scipy.ndimage.uniform_filter(p.astype(np.float64), size=5, mode="nearest")
    run = np.abs(O.nr_box_mean(p, 2) - ref)
    direct = np.abs(O.nr_box_mean(p, 2, direct=True) - ref)
    seg = w // 4
    early, late = run[:, :64].mean(), run[:, seg - 64:seg].mean()
    assert late > 2 * early  # the running sum's error grows along its segment
    assert direct.max() < 1e-6 and direct[:, seg - 64:seg].mean() < late / 4


@pytest.mark.parametrize("shape", [(10, 12), (11, 13), (1, 7), (48, 49)])
def test_down_is_the_2x2_mean(shape):
    rng = np.random.default_rng(1)
    p = rng.random(shape).astype(F32)
    h, w = shape
    q = np.pad(p, ((0, h % 2), (0, w % 2)), mode="edge")
    want = F32(0.25) * (((q[0::2, 0::2] + q[0::2, 1::2]) + q[1::2, 0::2]) + q[1::2, 1::2])
    assert np.array_equal(O.nr_down(p), want)


def test_add_up_closed_forms():
    """Upsampling a constant adds the constant; a linear ramp's 2x2 means upsample back to the
    ramp away from the borders (pixel centres aligned)."""
    d = np.zeros((20, 26), F32)
    assert np.array_equal(O.nr_add_up(d, np.full((10, 13), 0.375, F32)), np.full((20, 26), 0.375, F32))
    yy, xx = np.mgrid[0:20, 0:26].astype(F32)
    ramp = (0.01 * xx + 0.02 * yy).astype(F32)
    up = O.nr_add_up(np.zeros_like(ramp), O.nr_down(ramp))
    assert np.abs(up[1:-1, 1:-1] - ramp[1:-1, 1:-1]).max() < 1e-5


@pytest.mark.parametrize("sigma", [0.01, 0.05])
def test_estimate_finds_white_noise(sigma):
    rng = np.random.default_rng(11)
    w = h = 512
    p = (0.3 + sigma * rng.standard_normal((h, w))).astype(F32)
    g = np.full((h, w), 0.3, F32)
    means, sig = O.nr_model_from_stats(O.nr_stats(p, g))
    mean = float(np.mean(sig[:, 1:].mean(axis=1)))
    assert abs(mean / sigma - 1.0) < 0.04, mean


def test_estimate_grid_and_levels():
    assert O.nr_estimate_step(40, 40) == 0  # 25 blocks < 32
    assert O.nr_estimate_step(64, 32) == 1
    # 6048 x 4024: 756 * 503 blocks, ceil(sqrt(380268 / 50000)) = 3
    assert O.nr_estimate_step(6048, 4024) == 3
    assert O.nr_stats(np.zeros((4024, 6048), F32)[:64, :64], np.zeros((64, 64), F32)).shape == (64, 64)
    assert O.nr_level_count(6048, 4024, 6) == 6 and O.nr_level_count(6048, 4024, 4) == 4
    assert O.nr_level_count(100, 100, 6) == 2  # 50 >= 48, 25 < 48
    assert O.nr_level_count(94, 200, 6) == 1


def _from_stats_np(stats):
    """Model::from_stats (mod.rs:803-848) in numpy."""
    st = stats.astype(F32)
    act = np.zeros(len(st), F32)
    for k in range(1, 64):
        act = (act + st[:, k]).astype(F32)
    ok = np.isfinite(st[:, 0]) & np.isfinite(act)
    idx = np.nonzero(ok)[0]
    if len(idx) < 32:
        return None
    order = idx[np.argsort(st[idx, 0], kind="stable")]
    n = len(order)
    bins = min(max(n // 96, 1), 12)
    per = -(-n // bins)
    means, sig = [], []
    for b0 in range(0, n, per):
        part = order[b0:b0 + per]
        a = act[part]
        cut = np.sort(a)[len(a) // 2]
        flat = part[a <= cut]
        s = np.zeros(64, F32)
        for k in range(1, 64):
            med = np.sort(st[flat, k])[len(flat) // 2]
            s[k] = np.sqrt(F32(F32(med / F32(0.4549364)) * F32(1.10)))
        means.append(st[part[len(part) // 2], 0])
        sig.append(s)
    sig = np.asarray(sig, F32)
    if len(sig) > 2:
        o = sig.copy()
        for i in range(len(sig)):
            a, b = max(i - 1, 0), min(i + 1, len(sig) - 1)
            sig[i] = (F32(0.25) * o[a] + F32(0.5) * o[i]) + F32(0.25) * o[b]
    return np.asarray(means, F32), sig


@pytest.mark.parametrize("n", [31, 32, 95, 97, 500, 1300])
def test_model_from_stats(n):
    rng = np.random.default_rng(n)
    st = rng.random((n, 64)).astype(F32) ** 3
    st[:, 0] = np.round(rng.random(n) * 20) / 20  # ties in the means: the stable order matters
    got = O.nr_model_from_stats(st)
    want = _from_stats_np(st)
    if want is None:
        assert got is None
        return
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    assert len(got[0]) == min(max(n // 96, 1), 12)


def test_model_from_stats_constant_blocks():
    """Every block alike: each deviation is sqrt(c / chi2 median * 1.1) and no bin differs."""
    st = np.tile(np.linspace(1e-4, 2e-3, 64, dtype=F32), (400, 1))
    st[:, 0] = np.linspace(0.1, 0.9, 400, dtype=F32)
    means, sig = O.nr_model_from_stats(st)
    assert len(means) == 4
    want = np.sqrt((st[0, 1:].astype(np.float64) / 0.4549364) * 1.10)
    assert np.abs(sig[:, 1:] - want).max() < 1e-7 and np.all(sig[:, 0] == 0)
    # bin means: the middle block of each quarter of the sorted means
    assert np.array_equal(means, st[[50, 150, 250, 350], 0])
    bad = st.copy()
    bad[3, 0] = np.nan  # non-finite blocks are left out
    assert len(O.nr_model_from_stats(bad)[0]) == 4


def _plane(w, h, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    return (0.3 + 0.2 * np.sin(xx / 9.0) * np.cos(yy / 7.0) + 0.02 * rng.standard_normal((h, w))).astype(F32)


@pytest.mark.parametrize("step", [2, 4])
def test_dct_pass_limits(step):
    w, h = 61, 45
    p = _plane(w, h, 2)
    means = np.asarray([0.1, 0.5], F32)
    sig = np.full((2, 64), 0.01, F32)
    sig[:, 0] = 0.0
    # threshold 0: every coefficient kept, every block reconstructs itself
    same = O.nr_dct_pass(p, p, None, means, sig, gain=1.0, hard=0.0, step=step)
    assert np.abs(same - p).max() < 2e-6
    # a huge threshold keeps the DC only: the mean over the blocks covering each sample of the
    # blocks' means on the symmetrically padded plane
    dc = O.nr_dct_pass(p, p, None, means, sig, gain=1.0, hard=1e30, step=step)
    pw = w + 16
    fw = -(-pw // 8) * 8
    pad = np.pad(p.astype(np.float64), ((8, 8), (8, 8 + fw - pw)), mode="symmetric")
    # the extra columns past the symmetric padding repeat the pad map's clamp
    xs = list(range(0, pw - 8 + 1, step))
    xs += [pw - 8] if xs[-1] != pw - 8 else []
    ys = list(range(0, h + 16 - 8 + 1, step))
    ys += [h + 16 - 8] if ys[-1] != h + 16 - 8 else []
    acc = np.zeros(pad.shape)
    cnt = np.zeros(pad.shape)
    for y0 in ys:
        for x0 in xs:
            acc[y0:y0 + 8, x0:x0 + 8] += pad[y0:y0 + 8, x0:x0 + 8].mean()
            cnt[y0:y0 + 8, x0:x0 + 8] += 1
    want = (acc / np.maximum(cnt, 1))[8:8 + h, 8:8 + w]
    assert np.abs(dc - want).max() < 2e-6
    # the Wiener pass with zero noise deviations and the plane as its own pilot: the identity
    z = np.zeros((2, 64), F32)
    wi = O.nr_dct_pass(p, p, p, means, z, gain=1.0, step=step)
    assert np.abs(wi - p).max() < 2e-6


def _rgb_noise(w, h, means, sd, seed):
    rng = np.random.default_rng(seed)
    return np.stack([m + sd * rng.standard_normal((h, w)) for m in means], -1).astype(F32)


def test_flat_noise_is_removed_and_the_mean_kept():
    w, h = 256, 192
    rgb = _rgb_noise(w, h, (0.4, 0.3, 0.2), 0.03, 3)
    out = O.nr_apply(rgb, 1.0, 1.0)
    for c, m in enumerate((0.4, 0.3, 0.2)):
        before = np.sqrt(((rgb[..., c] - m) ** 2).mean())
        after = np.sqrt(((out[..., c] - m) ** 2).mean())
        assert after < before * 0.25, (c, before, after)
        assert abs(out[..., c].mean() - m) < 2e-3


def test_zero_strength_changes_nothing_and_edges_survive():
    w = h = 128
    rng = np.random.default_rng(5)
    v = np.where(np.arange(w) < w // 2, 0.1, 0.8).astype(F32)[None, :, None]
    rgb = (v + 0.01 * rng.standard_normal((h, w, 3))).astype(F32)
    assert np.array_equal(O.nr_apply(rgb, 0.0, 0.0), rgb)
    out = O.nr_apply(rgb, 1.0, 1.0)
    assert abs(out[:, w // 2 - 1, 1].mean() - 0.1) < 0.02 and abs(out[:, w // 2, 1].mean() - 0.8) < 0.02


def test_luma_only_keeps_the_colour_differences_and_chroma_only_keeps_luma():
    w, h = 150, 110
    rgb = _rgb_noise(w, h, (0.35, 0.3, 0.25), 0.02, 8)
    kr, kg, kb = 0.2126, 0.7152, 0.0722

    def ybr(x):
        y = kr * x[..., 0] + kg * x[..., 1] + kb * x[..., 2]
        return y, x[..., 2] - y, x[..., 0] - y

    y0, u0, v0 = ybr(rgb.astype(np.float64))
    y1, u1, v1 = ybr(O.nr_apply(rgb, 1.0, 0.0).astype(np.float64))
    assert np.abs(u1 - u0).max() < 1e-5 and np.abs(v1 - v0).max() < 1e-5 and np.abs(y1 - y0).std() > 5e-3
    y2, u2, v2 = ybr(O.nr_apply(rgb, 0.0, 1.0).astype(np.float64))
    assert np.abs(y2 - y0).max() < 1e-5 and np.abs(u2 - u0).std() > 5e-3
    # strengths clamp to [0, 1]; tiny frames are left alone
    assert np.array_equal(O.nr_apply(rgb, 3.0, 1.0), O.nr_apply(rgb, 1.0, 1.0))
    tiny = _rgb_noise(7, 30, (0.3, 0.3, 0.3), 0.02, 1)
    assert np.array_equal(O.nr_apply(tiny, 1.0, 1.0), tiny)
