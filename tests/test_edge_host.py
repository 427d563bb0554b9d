"""Host side of the edge-aware spec: the canonical position of a frame (Plan::new,
edgeaware/mod.rs:204-230), the smoothing weight estimated from the noise of the frame
(Plan::set_noise_from, mod.rs:369-409) and the user level's floor (Plan::apply_level,
mod.rs:413-419; NoiseReduction::floor, options.rs:70-78).

The package's estimate (numpy, host code) must equal the oracle's C restatement bit for bit; the
oracle is pinned by closed forms: a frame whose green pairs all give the same value fixes the
constants, one with two populations fixes the rank (len / 5), and the reference test's property
(edge_aware.rs:214: beta > 0 exactly on the noisy frames) holds on its mosaic model.
"""
import numpy as np
import pytest

import flashraw_amd as F
import np_restatement as N
from flashraw_amd import synth
from oracle import oracle as O

PHASES = ["RGGB", "BGGR", "GRBG", "GBRG"]
F32 = np.float32


def _info(w, h, cfa):
    return F.FrameInfo(w, h, F.CfaPattern.bayer(cfa), F.BlackLevel.uniform(256.0), F.WhiteLevel.uniform(4095))


@pytest.mark.parametrize("cfa", PHASES)
@pytest.mark.parametrize("origin", [(0, 0), (1, 0), (0, 1), (3, 5)])
def test_edge_plan(cfa, origin):
    assert F.edge_plan(_info(8, 8, cfa), origin) == N.edge_plan(cfa, origin)


def test_edge_plan_rejects_other_patterns():
    assert F.edge_plan(F.FrameInfo(8, 8, F.CfaPattern(2, 2, (0, 1, 1, 1))), (0, 0)) is None


@pytest.mark.parametrize("cfa", PHASES)
@pytest.mark.parametrize("read_sigma", [0.5, 4.0, 20.0])
@pytest.mark.parametrize("g_gain", [1.0, 1.7])
def test_beta_matches_oracle(cfa, read_sigma, g_gain):
    w, h = 517, 333
    mos = synth.mosaic(w, h, cfa, black=256.0, white=4095.0, gain=2.0, read_sigma=read_sigma, seed=int(read_sigma * 7))
    for origin in ((0, 0), (1, 1)):
        norm = O.normalize(mos, cfa, origin=origin, black=(1, 1, 1, [256.0]), white=(4095,))
        cx, cy = N.edge_plan(cfa, origin)
        want = O.edge_beta(norm, cx, cy, g_gain)
        got = F.estimate_beta(w, h, g_gain, lambda y: norm[y], cx, cy)
        assert F32(got) == F32(want), (cfa, origin, got, want)


def _beta_closed_form(q):
    sigma = q / (np.sqrt(2.0) * 0.2533)
    nu = sigma * sigma
    return 0.5 * min(max((nu - 2e-4) / 6e-4, 0.0), 1.0)


@pytest.mark.parametrize("nu", [1e-4, 3.5e-4, 5e-4, 7.5e-4, 1e-3])
def test_beta_known_answer(nu):
    """Every pair gives |a - b| / sqrt(mid) = q, so beta = 0.5 clamp((nu - 2e-4) / 6e-4, 0, 1)."""
    w, h = 200, 120
    q = np.sqrt(nu) * np.sqrt(2.0) * 0.2533
    # greens g, g + d, g, g + d, ...: every pair has |a - b| = d and mid = g + d / 2
    mid = 0.25
    d = q * np.sqrt(mid)
    f = np.zeros((h, w), dtype=F32)
    for y in range(h):
        xs = np.arange((y + 1) % 2, w, 2)
        f[y, xs] = np.where(np.arange(len(xs)) % 2 == 0, mid - d / 2, mid + d / 2).astype(F32)
    want = _beta_closed_form(q)
    got_o = O.edge_beta(f, 0, 0, 1.0)
    got_p = F.estimate_beta(w, h, 1.0, lambda y: f[y], 0, 0)
    assert F32(got_o) == F32(got_p)
    assert abs(got_o - want) < 2e-3, (got_o, want)


@pytest.mark.parametrize("frac_lo,expect", [(0.3, "lo"), (0.1, "hi")])
def test_beta_rank_is_the_lower_quintile(frac_lo, expect):
    """With a fraction of small differences above 1/5 the estimate comes from them, below 1/5 from
    the large ones."""
    w, h = 240, 100
    mid = 0.25
    lo, hi = 0.004, 0.012  # |a - b|
    f = np.zeros((h, w), dtype=F32)
    for y in range(h):
        xs = np.arange((y + 1) % 2, w, 2)
        n = len(xs)
        # values g0, g0 + d1, g0, g0 + d2, ...; pair i (samples i, i + 1) differs by d of its run
        d = np.where(np.arange(n) < frac_lo * n, lo, hi)
        vals = np.where(np.arange(n) % 2 == 0, mid - d / 2, mid + d / 2)
        f[y, xs] = vals.astype(F32)
    got = O.edge_beta(f, 0, 0, 1.0)
    assert F32(got) == F32(F.estimate_beta(w, h, 1.0, lambda y: f[y], 0, 0))
    want = _beta_closed_form((lo if expect == "lo" else hi) / np.sqrt(mid))
    assert abs(got - want) < 5e-3, (got, want)


def test_beta_zero_cases():
    f = np.full((15, 400), 0.3, F32)
    assert O.edge_beta(f, 0, 0, 1.0) == 0.0 and F.estimate_beta(400, 15, 1.0, lambda y: f[y], 0, 0) == 0.0
    # flat frame: every difference is 0, nu = 0 < 2e-4
    f = np.full((64, 400), 0.3, F32)
    assert O.edge_beta(f, 0, 0, 1.0) == 0.0 and F.estimate_beta(400, 64, 1.0, lambda y: f[y], 0, 0) == 0.0
    # every mean outside (0.03, 0.6): fewer than 1000 values
    rng = np.random.default_rng(1)
    f = (0.7 + 0.4 * rng.random((64, 400))).astype(F32)
    assert O.edge_beta(f, 0, 0, 1.0) == 0.0 and F.estimate_beta(400, 64, 1.0, lambda y: f[y], 0, 0) == 0.0
    # ... unless the green gain brings them inside
    assert O.edge_beta(f, 0, 0, 0.5) > 0.0


def _reference_test_mosaic(w, h, cfa, noise, seed):
    """The noise model of edge_aware.rs:39-80 (photon-like noise of variance noise^2 * value on a
    scene of flat shapes in [0, 1], mosaicked and scaled to black 256 / white 4095), with numpy's
    generator in place of that test's LCG."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    disc = np.hypot(xx - 70, yy - 60) < 48
    diag = ((xx * 0.8 + yy * 0.6).astype(np.int64) % 18) < 6
    rgb = np.stack([0.2 + 0.5 * disc + 0.2 * diag, 0.3 + 0.3 * diag, 0.25 + 0.4 * disc], -1)
    col = np.asarray(synth.BAYER[cfa]).reshape(2, 2)[yy.astype(int) % 2, xx.astype(int) % 2]
    v = np.take_along_axis(rgb, col[..., None], -1)[..., 0]
    if noise > 0:
        v = v + noise * np.sqrt(v) * rng.standard_normal(v.shape)
    return np.clip(np.round(256 + v * (4095 - 256)), 0, 4095).astype(np.uint16)


@pytest.mark.parametrize("cfa", PHASES)
def test_beta_positive_exactly_on_noisy_frames(cfa):
    """edge_aware.rs:214: beta > 0 on the frames with noise 0.04, 0 on the clean ones."""
    w, h = 203, 141
    for noise in (0.0, 0.04):
        mos = _reference_test_mosaic(w, h, cfa, noise, seed=7)
        norm = O.normalize(mos, cfa, black=(1, 1, 1, [256.0]), white=(4095,))
        cx, cy = N.edge_plan(cfa, (0, 0))
        beta = F.estimate_beta(w, h, 1.0, lambda y: norm[y], cx, cy)
        assert (beta > 0.0) == (noise > 0.0), (cfa, noise, beta)


def test_edge_spec_levels():
    """apply_level: Off clears the estimate, Low / Medium / High are floors, Auto keeps it."""
    w, h = 300, 200
    mos = synth.mosaic(w, h, "RGGB", black=256.0, white=4095.0, gain=2.0, read_sigma=6.0, seed=2)
    norm = O.normalize(mos, "RGGB", black=(1, 1, 1, [256.0]), white=(4095,))
    info = _info(w, h, "RGGB")
    tables = synth.edge_tables()
    est = F.estimate_beta(w, h, 1.0, lambda y: norm[y], 0, 0)
    assert 0.0 < est < 0.5
    got = {lv: F.edge_spec(info, (0, 0), (w, h), (2.0, 1.0, 1.5, 1.0), lambda y: norm[y], tables, lv).beta
           for lv in F.NoiseReduction}
    assert got[F.NoiseReduction.Auto] == pytest.approx(est, abs=0)
    assert got[F.NoiseReduction.Off] == 0.0
    assert got[F.NoiseReduction.Low] == max(est, 0.25)
    assert got[F.NoiseReduction.Medium] == pytest.approx(max(est, 0.5))
    assert got[F.NoiseReduction.High] == pytest.approx(0.8)
    spec = F.edge_spec(info, (1, 0), (w, h), (2.0, 1.0, 1.5, 1.0), lambda y: norm[y], tables)
    assert (spec.period, spec.halo, tuple(spec.canon), spec.ahd) == (2, 8, (0, 1, 1, 2), False)
    assert (spec.cx, spec.cy) == N.edge_plan("RGGB", (1, 0))
    assert F.edge_spec(F.FrameInfo(w, h, F.CfaPattern(2, 2, (0, 1, 1, 1))), (0, 0), (w, h), (1, 1, 1, 1),
                       lambda y: norm[y], tables) is None
