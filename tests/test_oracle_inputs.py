"""Oracle pinning for the input side and the develop layouts beyond Bayer mosaics:

- packed payloads (unpack_row_scalar, crates/flashraw-codecs/src/packed.rs:137-180) against the
  GPU kernels' own statement (kernels.rs:333-357) over the reference's case matrix
  (crates/flashraw-gpu/tests/equivalence.rs:66-90: every depth 1..16, every order, odd widths,
  padded strides, truncated payloads), round trips through a packer written from the format
  definitions, and known answers;
- the linearisation curve (run.rs:38-39);
- normalisation of multi-component, CFA-less and f32 frames (Calib, process.rs:256-320) against a
  numpy statement, with the reference's calibrated_info (equivalence.rs:166-177);
- half-size develop (process.rs:975-1036) and the no-demosaic layouts (process.rs:1038-1046)
  against numpy statements, bit for bit.
"""
import numpy as np
import pytest

import np_restatement as N
from oracle import oracle as O

F32 = np.float32
ORDERS = [0, 1, 2, 3]
XTRANS = "GGRGGB/GGBGGR/BRGRBG/GGBGGR/GGRGGB/RBGBRG"


def pseudo(n, seed):
    """equivalence.rs:39-50: xorshift bytes."""
    x = (seed | 1) & 0xFFFFFFFFFFFFFFFF
    out = bytearray(n)
    for i in range(n):
        x ^= (x << 13) & 0xFFFFFFFFFFFFFFFF
        x ^= x >> 7
        x ^= (x << 17) & 0xFFFFFFFFFFFFFFFF
        out[i] = x & 0xFF
    return np.frombuffer(bytes(out), dtype=np.uint8)


@pytest.mark.parametrize("bits", range(1, 17))
def test_unpack_matches_the_gpu_statement(bits):
    for order in ORDERS:
        for w in (1, 3, 8, 13, 64, 101):
            h = 5
            row_bytes = (w * bits + 7) // 8
            for stride in (row_bytes, row_bytes + 3, row_bytes + 8):
                for cut in (0, 5):
                    n = max(stride * h - cut, 0)
                    src = pseudo(n, bits * 7919 + w + stride)
                    got = O.unpack(src, w, h, bits, order, stride)
                    want = N.packed_model(src, w, h, bits, order, stride)
                    assert np.array_equal(got, want), (bits, order, w, stride, cut)


@pytest.mark.parametrize("order", ORDERS)
def test_unpack_round_trip(order):
    rng = np.random.default_rng(order)
    for bits in range(1, 17):
        wsz = 2 if order == 2 else (4 if order == 3 else 1)
        for w in (8, 16, 32, 40, 96):
            if ((w * bits + 7) // 8) % wsz:
                continue
            samples = rng.integers(0, 1 << bits, (4, w)).astype(np.uint16)
            row_bytes = (w * bits + 7) // 8
            for stride in (row_bytes, row_bytes + 6):
                src = N.pack(samples, bits, order, stride)
                assert np.array_equal(O.unpack(src, w, 4, bits, order, stride), samples), (bits, w, stride)


def test_unpack_known_answers():
    b = np.array([0xAB, 0xCD, 0xEF], np.uint8)
    assert O.unpack(b, 2, 1, 12, "msb_bytes", 3).tolist() == [[0xABC, 0xDEF]]
    assert O.unpack(b, 2, 1, 12, "lsb_bytes", 3).tolist() == [[0xDAB, 0xEFC]]
    # 16-bit words stored little-endian, MSB first inside: bytes 34 12 78 56 -> 0x1234 0x5678
    w = np.array([0x34, 0x12, 0x78, 0x56], np.uint8)
    assert O.unpack(w, 2, 1, 16, "msb16", 4).tolist() == [[0x1234, 0x5678]]
    assert O.unpack(w, 2, 1, 16, "msb32", 4).tolist() == [[0x5678, 0x1234]]
    # one 16-bit sample in a 2-byte row: its logical bytes sit at physical 3 and 2, past the row
    assert O.unpack(w, 1, 1, 16, "msb32", 4).tolist() == [[0]]
    assert O.unpack(w, 4, 1, 8, "msb32", 4).tolist() == [[0x56, 0x78, 0x12, 0x34]]
    # 10-bit MSB: 0b1111111111 0b0000000001 -> bytes FF C0 10
    assert O.unpack(np.array([0xFF, 0xC0, 0x10], np.uint8), 2, 1, 10, "msb_bytes", 3).tolist() == [[1023, 1]]
    # a truncated payload reads zeros
    assert O.unpack(np.array([0xAB], np.uint8), 2, 1, 12, "msb_bytes", 3).tolist() == [[0xAB0, 0]]
    # bytes past row_bytes read zero even when the stride holds data (word orders)
    p = np.array([0x01, 0x80, 0xFF, 0xFF], np.uint8)
    assert O.unpack(p, 1, 1, 12, "msb16", 4).tolist() == [[0x800]]


def test_linearize():
    v = np.array([0, 1, 5, 11999, 12000, 65535], np.uint16)
    lut = np.array([(i * 3 // 2 + 7) & 0xFFFF for i in range(12000)], np.uint16)
    got = O.linearize(v, lut)
    want = [int(lut[min(int(x), 11999)]) for x in v]
    assert got.tolist() == want
    assert O.linearize(v, np.zeros(0, np.uint16)).tolist() == v.tolist()


def _calibrated(w, h):
    """equivalence.rs:166-177: 2x2 black pattern, row and column deltas, per-colour whites."""
    black = (2, 2, 1, [510.0, 512.0, 509.0, 515.0])
    rowd = [(y % 5) * 0.5 for y in range(h + 8)]
    cold = [(x % 3) * -0.25 for x in range(w + 8)]
    return black, (16000, 16383, 15800), rowd, cold


def _synthetic(w, h, cpp, seed):
    """equivalence.rs:179-189: gradient plus hashed noise, below black to above white."""
    i = np.arange(w * h * cpp, dtype=np.uint64)
    x = (i // cpp) % w
    y = (i // cpp) // w
    n = ((i * np.uint64(2654435761) + np.uint64(seed)) & np.uint64(0xFFFFFFFF)) >> np.uint64(22)
    return np.minimum(300 + x * 40 + y * 25 + n, 16383).astype(np.uint16).reshape(h, w * cpp)


@pytest.mark.parametrize("cfa,cpp", [("RGGB", 1), (None, 3), (None, 1), (XTRANS, 1), (None, 4)])
def test_normalise_layouts(cfa, cpp):
    w, h = 123, 77
    black, white, rowd, cold = _calibrated(w, h)
    arr = _synthetic(w, h, cpp, 5)
    kw = dict(origin=(3, 1), black=black, white=white, row_deltas=rowd, col_deltas=cold)
    pat = N.parse_cfa(cfa) if cfa is not None else None
    got = O.normalize(arr, pat, cpp=cpp, **kw)
    want = N.normalise_c(arr, cpp, cfa, **kw)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    clamped = O.normalize(arr, pat, cpp=cpp, clamped=True, **kw)
    assert np.array_equal(clamped, np.maximum(want, F32(0.0)))


def test_normalise_f32_samples():
    w, h = 64, 20
    rng = np.random.default_rng(3)
    arr = (rng.random((h, w * 3)) * 2.0 - 0.5).astype(F32)
    black = (1, 1, 3, [0.01, 0.02, 0.03])
    got = O.normalize(arr, None, cpp=3, black=black, white=(1,))
    want = N.normalise_c(arr, 3, None, black=black, white=(1,))
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


MATRIX = ((1.5, -0.3, -0.2), (-0.2, 1.4, -0.2), (0.05, -0.4, 1.35))
WB = (2.1, 1.0, 1.6, 1.0)


@pytest.mark.parametrize("cfa,cpp", [("RGGB", 1), ("GBRG", 1), (XTRANS, 1), (None, 3), (None, 1), (None, 2)])
@pytest.mark.parametrize("crop", [(0, 0, 157, 93), (5, 4, 148, 86), (6, 3, 151, 90)])
def test_half_size(cfa, cpp, crop):
    w, h = 157, 93
    black, white, rowd, cold = _calibrated(w, h)
    arr = _synthetic(w, h, cpp, 11)
    origin = (1, 3)
    pat = N.parse_cfa(cfa) if cfa is not None else None
    got = O.develop(arr, pat, cpp=cpp, origin=origin, black=black, white=white, row_deltas=rowd, col_deltas=cold,
                    wb=WB, matrix=MATRIX, crop=crop, half_size=True, encoding="linear_f32")
    norm = N.normalise_c(arr, cpp, cfa, origin, black, white, rowd, cold)
    col = N.colours(cfa, w, h, origin) if cfa is not None else None
    want = N.finish(N.half_size(norm, cpp, col, crop), WB, MATRIX)
    assert got.shape == want.shape
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("cpp", [1, 2, 3, 4])
def test_no_demosaic_layouts(cpp):
    """Frames without a CFA: RGB from the first three components (cpp >= 3), else grey from the
    first component (process.rs:1038-1046)."""
    w, h = 61, 29
    black, white, rowd, cold = _calibrated(w, h)
    arr = _synthetic(w, h, cpp, 2)
    got = O.develop(arr, None, cpp=cpp, black=black, white=white, row_deltas=rowd, col_deltas=cold, wb=WB,
                    matrix=MATRIX, encoding="linear_f32", crop=(2, 1, 50, 25))
    norm = N.normalise_c(arr, cpp, None, (0, 0), black, white, rowd, cold)
    comps = [norm[:, c::cpp] if cpp >= 3 else norm[:, 0::cpp] for c in range(3)]
    want = N.finish(np.stack(comps, -1)[1:26, 2:52], WB, MATRIX)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_half_size_flat_field_known_answer():
    """A flat field stays flat at half size: (1000 - 0) / 4000 = 0.25 in every channel."""
    mos = np.full((40, 48), 1000, np.uint16)
    got = O.develop(mos, "RGGB", black=(1, 1, 1, [0.0]), white=(4000,), encoding="linear_f32", half_size=True)
    assert got.shape == (20, 24, 3) and np.all(got == F32(0.25))
