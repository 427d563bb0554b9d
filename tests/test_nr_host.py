"""Host side of the noise reduction: flashraw::denoise::gpu_spec (crates/flashraw/src/denoise/mod.rs:
155-193) as the library states it (frh_nr_spec_for), against the reference's formulas: the pyramid
levels (level_count, four for luma and six for chroma while both sides stay at least 48), the luma
keep (1 - s)^1.2, the Wiener pass skipped at full resolution up to luma 0.25, the chroma gain
1 + 2 max(s - 0.5, 0), keep max(1 - 2s, 0), steps 4 on the two finest chroma levels, and the
guided filter's radius ceil(round(6 (0.5 + s)) / 2) at half resolution; and the library's
Model::from_stats against the oracle's bit for bit.
"""
import numpy as np
import pytest

import flashraw_amd as F
from oracle import oracle as O

F32 = np.float32


@pytest.mark.parametrize("w,h", [(6048, 4024), (347, 229), (100, 100), (94, 200), (8, 8), (7, 30)])
@pytest.mark.parametrize("luma,chroma", [(0.5, 0.0), (0.0, 0.5), (0.2, 0.0), (1.0, 1.0), (0.5, 0.3), (0.25, 0.75),
                                         (-1.0, 2.0)])
def test_spec_for(w, h, luma, chroma):
    s = F.denoise_spec(w, h, luma, chroma)
    lu, ch = min(max(luma, 0.0), 1.0), min(max(chroma, 0.0), 1.0)
    assert (s.luma is not None) == (lu > 0) and (s.chroma is not None) == (ch > 0)
    if s.luma is not None:
        assert s.luma.levels == O.nr_level_count(w, h, 4)
        assert (s.luma.gain, tuple(s.luma.steps), s.luma.hard_only0) == (1.0, (2, 2), lu <= 0.25)
        # powf of the f32 operands, correctly rounded (the f64 power of the same operands, rounded)
        assert F32(s.luma.keep) == F32(np.float64(F32(1.0) - F32(lu)) ** np.float64(F32(1.2)))
    if s.chroma is not None:
        assert s.chroma.levels == O.nr_level_count(w, h, 6)
        assert F32(s.chroma.gain) == F32(1.0) + F32(2.0) * max(F32(ch) - F32(0.5), F32(0.0))
        assert (tuple(s.chroma.steps), s.chroma.hard_only0) == ((4, 4), True)
        assert F32(s.chroma.keep) == max(F32(1.0) - F32(2.0) * F32(ch), F32(0.0))
    r = max(int(np.floor(F32(6.0) * (F32(0.5) + F32(ch)) + F32(0.5))), 1)
    assert s.chroma_radius == max(-(-r // 2), 1)
    assert (F32(s.hard), s.max_blocks, s.output) == (F32(2.7), 50000, F.DevelopOutput.LinearRgbF32)


@pytest.mark.parametrize("n", [20, 32, 200, 1300, 4000])
def test_model_from_stats_equals_the_oracle(n):
    rng = np.random.default_rng(n)
    st = (rng.random((n, 64)) ** 3).astype(F32)
    st[:, 0] = np.round(rng.random(n) * 30) / 30
    st[n // 3, 5] = np.inf  # a non-finite block is left out
    got, want = F.nr_model_from_stats(st), O.nr_model_from_stats(st)
    if want is None:
        assert got is None
        return
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
