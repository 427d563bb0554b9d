"""Writes tests/golden/develop_small.npz and tests/golden/dct_small.npz from the CPU oracle.

Inputs restate the reference tests' deterministic mosaics (the LCG stream of
crates/flashraw/src/process.rs:2371-2375); outputs are the oracle's, for every Bayer phase x
{bilinear, AHD} x {RGBA8, linear 16-bit}, plus the float64 CFA-plane DCT denoise of a small
frame. The oracle is pinned by tests/test_oracle.py; these fixtures freeze its outputs so that
a later change to the oracle or the GPU path that alters a byte is seen.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import np_restatement as N  # noqa: E402
from oracle import oracle as O  # noqa: E402

W, H = 48, 40
BLACK = (2, 2, 1, [60.0, 64.0, 62.0, 66.0])
WHITE = (4095,)
WB = (2.1, 1.0, 1.6, 1.0)
MATRIX = ((1.62, -0.38, -0.24), (-0.21, 1.58, -0.37), (-0.12, -0.45, 1.57))


def main():
    out = {"black": np.asarray(BLACK[3], dtype=np.float32), "wb": np.asarray(WB, dtype=np.float32),
           "matrix": np.asarray(MATRIX, dtype=np.float32), "white": np.asarray(WHITE, dtype=np.uint32)}
    for k, cfa in enumerate(["RGGB", "BGGR", "GRBG", "GBRG"]):
        mos = N.lcg_stream(W, H, seed=1000 + k)
        out[f"mosaic_{cfa}"] = mos
        for dm in ("bilinear", "ahd"):
            for enc in ("rgba8", "linear16"):
                out[f"{cfa}_{dm}_{enc}"] = O.develop(mos, cfa, origin=(1, 0), black=BLACK, white=WHITE, wb=WB,
                                                     matrix=MATRIX, demosaic=dm, encoding=enc, threads=1)
    np.savez_compressed(os.path.join(HERE, "develop_small.npz"), **out)
    rng = np.random.default_rng(77)
    base = 2000 + 700 * np.sin(np.arange(64)[None, :] / 6.0) * np.cos(np.arange(48)[:, None] / 9.0)
    mos = np.clip(base + rng.normal(0, 45, (48, 64)), 0, 16383).astype(np.uint16)
    black = np.asarray([512.0, 510.0, 514.0, 511.0], dtype=np.float32)
    of, ou = O.dct_denoise(mos, black, 2.0, 144.0, 2.7, 4, precision=64, threads=1)
    np.savez_compressed(os.path.join(HERE, "dct_small.npz"), mosaic=mos, black=black,
                        params=np.asarray([2.0, 144.0, 2.7, 4], dtype=np.float32), out_f32=of, out_u16=ou)


if __name__ == "__main__":
    main()
