"""Writes tests/golden/rng_golden.npz from the reference's own RNG.

The reference's src/isaac64.{h,c} and src/rng.{h,c} compile on their own
(`make -C oracle ref` builds them into oracle/_ref/libref_rng.so); this script
records their outputs so that the oracle's RNG stays pinned on machines where
the reference tree is absent.

Contents (seeds 0, 1, 2, 12345):
  isaac_u32_<seed>  first 2048 rng_uint draws (four ISAAC64 result blocks)
  isaac_dbl_<seed>  first 1024 rng_dbl draws
  gauss_<seed>      first 4096 rng_gauss values
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

from oracle import pyoracle  # noqa: E402

SEEDS = [0, 1, 2, 12345]


def main():
    pyoracle.build()
    ref = pyoracle.ref_rng()
    if ref is None:
        raise SystemExit("oracle/_ref/libref_rng.so not built (reference tree missing?)")
    out = {}
    for s in SEEDS:
        u = np.zeros(2048, np.uint32)
        ref.ref_isaac64_u32(s, len(u), u.ctypes.data)
        d = np.zeros(1024, np.float64)
        ref.ref_isaac64_dbl(s, len(d), d.ctypes.data)
        g = np.zeros(4096, np.float64)
        ref.ref_rng_gauss(s, len(g), g.ctypes.data)
        out["isaac_u32_%d" % s] = u
        out["isaac_dbl_%d" % s] = d
        out["gauss_%d" % s] = g
    path = os.path.join(HERE, "rng_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
