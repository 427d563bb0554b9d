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


# ---------------------------------------------------------------- X-Trans (period 6)

XT = N.XTRANS


def _xt_info(w, h):
    return F.FrameInfo(w, h, F.CfaPattern.parse(XT), F.BlackLevel.uniform(256.0), F.WhiteLevel.uniform(4095))


def test_xtrans_canon_is_the_reference_arrangement():
    """mod.rs:168-185: GGRGGB / GGBGGR / BRGRBG / GGBGGR / GGRGGB / RBGBRG with R, G, B = 0, 1, 2."""
    rows = ["GGRGGB", "GGBGGR", "BRGRBG", "GGBGGR", "GGRGGB", "RBGBRG"]
    assert F.EDGE_CANON_XTRANS == tuple("RGB".index(c) for r in rows for c in r)
    assert list(F.EDGE_CANON_XTRANS) == N.edge_plan_p(XT, (0, 0))[2]


@pytest.mark.parametrize("origin", [(0, 0), (1, 0), (0, 1), (2, 3), (5, 5), (7, 13)])
def test_xtrans_edge_plan(origin):
    cx, cy, canon, period = F.edge_plan(_xt_info(12, 12), origin, with_pattern=True)
    assert (cx, cy, list(canon), period) == tuple(N.edge_plan_p(XT, origin))
    # the plan puts every frame sample at a canonical position of its own colour
    col = N.colours(XT, 12, 12, origin)
    for y in range(12):
        for x in range(12):
            assert canon[((y + cy) % 6) * 6 + (x + cx) % 6] == col[y, x]


def test_xtrans_edge_plan_shifted_patterns():
    """Plan::new accepts every shifted copy of the canonical arrangement (mod.rs:211-216) and
    nothing else of period 6."""
    p = N.parse_cfa(XT)
    for sy in range(6):
        for sx in range(6):
            shifted = np.roll(np.roll(p, -sy, 0), -sx, 1)
            info = F.FrameInfo(12, 12, F.CfaPattern(6, 6, tuple(int(v) for v in shifted.reshape(-1))))
            cx, cy = F.edge_plan(info, (0, 0))
            assert all(F.EDGE_CANON_XTRANS[((y + cy) % 6) * 6 + (x + cx) % 6] == shifted[y, x]
                       for y in range(6) for x in range(6))
    other = p.copy()
    other[2, 1], other[2, 2] = other[2, 2], other[2, 1]  # R and G swapped at one site
    assert F.edge_plan(F.FrameInfo(12, 12, F.CfaPattern(6, 6, tuple(int(v) for v in other.reshape(-1)))), (0, 0)) is None


def _pairs_brute(norm, cx, cy, canon, P, g_gain):
    """Plan::set_noise_from's values written out sample by sample (tiny frames only)."""
    h, w = norm.shape
    f32 = np.float32
    vals = []
    for n in range(48):
        y = 4 + n * (h - 9) // 47
        for x in range(w - 2):
            if canon[((y + cy) % P) * P + (x + cx) % P] != 1:
                continue
            for d in (1, 2):
                if canon[((y + cy) % P) * P + (x + d + cx) % P] != 1:
                    continue
                a, b = f32(norm[y, x]) * f32(g_gain), f32(norm[y, x + d]) * f32(g_gain)
                mid = f32(0.5) * (a + b)
                if f32(0.03) < mid < f32(0.6):
                    vals.append(f32(abs(a - b)) / f32(np.sqrt(mid)))
    return vals


@pytest.mark.parametrize("origin", [(0, 0), (3, 1)])
@pytest.mark.parametrize("read_sigma", [2.0, 20.0])
def test_xtrans_beta_brute_force(origin, read_sigma):
    """The oracle and the package against the sample-by-sample statement: the same values and the
    value of rank len / 5."""
    w, h = 160, 90
    rng = np.random.default_rng(int(read_sigma))
    norm = (0.3 + read_sigma * 1e-3 * rng.standard_normal((h, w))).astype(F32)
    cx, cy, canon, P = N.edge_plan_p(XT, origin)
    vals = _pairs_brute(norm, cx, cy, canon, P, 1.3)
    assert len(vals) >= 1000
    # both kinds of pair are present: one and two samples apart
    q = np.sort(np.asarray(vals, F32))[len(vals) // 5]
    sigma = F32(q) / (F32(np.sqrt(2.0)) * F32(0.2533))
    want = F32(0.5) * F32(min(max((sigma * sigma - F32(2e-4)) / F32(6e-4), F32(0)), F32(1)))
    got_o = O.edge_beta(norm, cx, cy, 1.3, canon=canon)
    got_p = F.estimate_beta(w, h, 1.3, lambda y: norm[y], cx, cy, canon)
    assert F32(got_o) == want and F32(got_p) == want, (got_o, got_p, want)


def test_xtrans_beta_closed_form_and_pair_kinds():
    """Greens alternating lo / hi along each row's chain of green samples: every pair (one or two
    apart) differs by d with mean g, so q = d / sqrt(g) and beta has its closed form; keeping only
    pairs two apart (the Bayer rule) halves the count, which drops below the 1000 values the
    estimate needs on this frame size."""
    canon = np.asarray(F.EDGE_CANON_XTRANS).reshape(6, 6)
    mid, d = 0.25, 0.01
    w, h = 56, 64
    f = np.zeros((h, w), F32)
    for y in range(h):
        greens = [x for x in range(w) if canon[y % 6, x % 6] == 1]
        for i, x in enumerate(greens):
            f[y, x] = mid - d / 2 if i % 2 == 0 else mid + d / 2
    vals = _pairs_brute(f, 0, 0, list(F.EDGE_CANON_XTRANS), 6, 1.0)
    n1 = sum(1 for y in range(h) for x in range(w - 2) if canon[y % 6, x % 6] == 1 and canon[y % 6, (x + 1) % 6] == 1)
    assert 1000 <= len(vals) < 2000 and n1 > 0
    got = O.edge_beta(f, 0, 0, 1.0, canon=F.EDGE_CANON_XTRANS)
    assert F32(got) == F32(F.estimate_beta(w, h, 1.0, lambda y: f[y], 0, 0, F.EDGE_CANON_XTRANS))
    assert abs(got - _beta_closed_form(d / np.sqrt(mid))) < 2e-3, got
    # every value is equal, so the rank cannot tell the pair kinds apart; the count can
    two_apart = sum(1 for n in range(48) for x in range(w - 2)
                    if canon[(4 + n * (h - 9) // 47) % 6, x % 6] == 1
                    and canon[(4 + n * (h - 9) // 47) % 6, (x + 2) % 6] == 1)
    assert two_apart < 1000 <= len(vals)


def test_xtrans_edge_spec():
    w, h = 300, 200
    rng = np.random.default_rng(3)
    norm = (0.3 + 0.02 * rng.standard_normal((h, w))).astype(F32)
    g0, d = N.edge_tables_xtrans(0)
    spec = F.edge_spec(_xt_info(w, h), (2, 1), (w, h), (2.0, 1.2, 1.5, 1.0), lambda y: norm[y], (g0, d))
    cx, cy, canon, P = N.edge_plan_p(XT, (2, 1))
    assert (spec.period, spec.cx, spec.cy, list(spec.canon), spec.halo, spec.ahd) == (6, cx, cy, canon, 8, False)
    assert F32(spec.beta) == F32(O.edge_beta(norm, cx, cy, 1.2, canon=canon)) and spec.beta > 0
    assert spec.g0.shape == (36, 49) and spec.dir.shape == (36, 3, 4, 49)
