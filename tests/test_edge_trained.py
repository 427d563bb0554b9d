"""The oracle's edge-aware demosaic on the reference's trained filter tables (bayer.bin and
xtrans.bin of crates/flashraw/src/edgeaware/, kept under tests/golden/ as f32 data and read in
the layout of load(), edgeaware/mod.rs:131-160): bit-exact against the independent per-plane
numpy statement, flat fields stay flat (every trained filter sums to 1, mod.rs:13-14), and the
tables' properties the fast GPU kernel relies on hold."""
import numpy as np
import pytest

import np_restatement as N
from flashraw_amd import synth
from oracle import oracle as O

BAYERS = ["RGGB", "BGGR", "GRBG", "GBRG"]
WB = (2.1, 1.0, 1.6, 1.0)
MATRIX = ((1.62, -0.38, -0.24), (-0.21, 1.58, -0.37), (-0.12, -0.45, 1.57))


def test_table_files_have_the_reference_layout():
    g0, d = synth.trained_edge_tables("bayer")
    assert g0.shape == (4, 49) and d.shape == (4, 3, 4, 49)
    # canonical RGGB (bayer(), mod.rs:163-166): green positions estimate G0 by the sample itself
    assert np.count_nonzero(g0[1]) == 1 and np.count_nonzero(g0[2]) == 1
    assert g0[1, 24] == 1.0 and g0[2, 24] == 1.0
    # the own colour's directional filters are zero, the others sum to 1 in every direction
    canon = (0, 1, 1, 2)
    for ph in range(4):
        for c in range(3):
            s = d[ph, c].sum(axis=-1)
            if c == canon[ph]:
                assert np.all(d[ph, c] == 0)
            else:
                assert np.allclose(s, 1.0, atol=1e-5), (ph, c, s)
    assert np.allclose(g0.sum(axis=1), 1.0, atol=1e-5)
    xg0, xd = synth.trained_edge_tables("xtrans")
    assert xg0.shape == (36, 49) and xd.shape == (36, 3, 4, 49)
    assert np.allclose(xg0.sum(axis=1), 1.0, atol=1e-5)
    assert np.isfinite(xd).all() and np.isfinite(d).all()


@pytest.mark.parametrize("cfa", BAYERS)
@pytest.mark.parametrize("origin", [(0, 0), (1, 1)])
@pytest.mark.parametrize("beta", [0.0, 0.35])
def test_bayer_trained_equals_per_plane_statement(cfa, origin, beta):
    w, h = 61, 45
    mos = N.lcg_stream(w, h, seed=41 + origin[0])
    g0, d = synth.trained_edge_tables("bayer")
    cx, cy = N.edge_plan(cfa, origin)
    edge = dict(g0=g0, dir=d, cx=cx, cy=cy, beta=beta)
    kw = dict(origin=origin, black=(1, 1, 1, [64.0]), white=(4095,), wb=WB, matrix=MATRIX)
    got = O.develop(mos, cfa, demosaic="edge", encoding="linear_f32", edge=edge, **kw)
    want = N.develop_edge(mos, cfa, g0=g0, dirt=d, beta=beta, **kw)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("origin", [(0, 0), (2, 3), (5, 1)])
@pytest.mark.parametrize("beta", [0.0, 0.2])
def test_xtrans_trained_equals_per_plane_statement(origin, beta):
    w, h = 58, 47
    mos = N.lcg_stream(w, h, seed=23 + origin[0])
    g0, d = synth.trained_edge_tables("xtrans")
    cx, cy, canon, period = N.edge_plan_p(N.XTRANS, origin)
    edge = dict(g0=g0, dir=d, cx=cx, cy=cy, beta=beta, canon=canon, period=period)
    kw = dict(origin=origin, black=(1, 1, 1, [64.0]), white=(4095,), wb=WB, matrix=MATRIX)
    got = O.develop(mos, N.parse_cfa(N.XTRANS).tolist(), demosaic="edge", encoding="linear_f32", edge=edge, **kw)
    want = N.develop_edge(mos, N.XTRANS, g0=g0, dirt=d, beta=beta, **kw)
    assert np.array_equal(got, want)


def test_trained_tables_keep_a_flat_field_flat():
    g0, d = synth.trained_edge_tables("bayer")
    mos = np.full((40, 48), 1200, dtype=np.uint16)
    edge = dict(g0=g0, dir=d, cx=0, cy=0, beta=0.3)
    got = O.develop(mos, "RGGB", black=(1, 1, 1, [0.0]), white=(4000,), demosaic="edge", encoding="linear_f32",
                    edge=edge)
    assert np.abs(got - 0.3).max() < 1e-5


def test_trained_and_synthetic_tables_differ_in_output():
    """The trained tables are what the comparisons use: a develop with them differs from one with
    the package's synthetic tables of the same sparsity."""
    mos = N.lcg_stream(61, 45, seed=3)
    kw = dict(black=(1, 1, 1, [64.0]), white=(4095,), wb=WB, matrix=MATRIX, demosaic="edge",
              encoding="linear_f32")
    g0, d = synth.trained_edge_tables("bayer")
    a = O.develop(mos, "RGGB", edge=dict(g0=g0, dir=d, cx=0, cy=0, beta=0.0), **kw)
    sg0, sd = synth.edge_tables(0)
    b = O.develop(mos, "RGGB", edge=dict(g0=sg0, dir=sd, cx=0, cy=0, beta=0.0), **kw)
    assert not np.array_equal(a[8:-8, 8:-8], b[8:-8, 8:-8])
