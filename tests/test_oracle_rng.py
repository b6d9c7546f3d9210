"""Pins the oracle's random draws.

* ISAAC64 and rng_gauss against the reference's own compiled RNG
  (oracle/_ref, when the reference tree is present) and against the golden
  vectors recorded from it (tests/golden/rng_golden.npz).
* Philox4x32-10 against the published Random123 known-answer vectors.
* Ziggurat table identities and moments of the Gaussian.
"""
import ctypes
import os

import numpy as np
import pytest

from oracle import pyoracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "rng_golden.npz")
SEEDS = [0, 1, 2, 12345]


def _isaac(lib, seed, n):
    out = np.zeros(n, np.uint32)
    lib.orc_isaac64_u32(seed, n, out.ctypes.data)
    return out


def _gauss(lib, seed, n):
    out = np.zeros(n)
    lib.orc_gauss_isaac(seed, n, out.ctypes.data)
    return out


@pytest.mark.parametrize("seed", SEEDS)
def test_isaac64_matches_golden(oracle_lib, seed):
    g = np.load(GOLDEN, allow_pickle=False)
    assert np.array_equal(_isaac(oracle_lib, seed, 2048), g["isaac_u32_%d" % seed])


@pytest.mark.parametrize("seed", SEEDS)
def test_rng_dbl_is_dbl32_times_u32(oracle_lib, seed):
    g = np.load(GOLDEN, allow_pickle=False)
    u = g["isaac_u32_%d" % seed][:1024].astype(np.float64)
    assert np.array_equal(u * 2.3283064365386962890625e-10, g["isaac_dbl_%d" % seed])


@pytest.mark.parametrize("seed", SEEDS)
def test_gauss_matches_golden_bitwise(oracle_lib, seed):
    g = np.load(GOLDEN, allow_pickle=False)
    mine = _gauss(oracle_lib, seed, 4096)
    assert np.array_equal(mine.view(np.uint64), g["gauss_%d" % seed].view(np.uint64))


def test_against_compiled_reference_when_present(oracle_lib):
    ref = pyoracle.ref_rng()
    if ref is None:
        pytest.skip("reference RNG not built here (no /root/reference)")
    for seed in [3, 7, 99, 2 ** 31 + 5]:
        n = 20000
        a = _isaac(oracle_lib, seed, n)
        b = np.zeros(n, np.uint32)
        ref.ref_isaac64_u32(seed, n, b.ctypes.data)
        assert np.array_equal(a, b)
        ga = _gauss(oracle_lib, seed, n)
        gb = np.zeros(n)
        ref.ref_rng_gauss(seed, n, gb.ctypes.data)
        assert np.array_equal(ga.view(np.uint64), gb.view(np.uint64))


def _philox(lib, ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib.orc_philox_block(c, k, o)
    return list(o)


def test_philox_known_answers(oracle_lib):
    # Random123 kat_vectors, philox4x32_10
    assert _philox(oracle_lib, [0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert _philox(oracle_lib, [0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6,
                                                                          0x6d5451fd]
    assert _philox(oracle_lib, [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == [
        0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_philox_stream_layout(oracle_lib):
    """The per-molecule stream is blocks (block, id, iteration, replica<<8|pass)
    read x0..x3 in order."""
    n = 12
    out = np.zeros(n, np.uint32)
    seed = 0x1234_5678_9abc_def0
    oracle_lib.orc_philox_stream(seed, 77, 5, 3, 1, n, out.ctypes.data)
    key = [seed & 0xffffffff, seed >> 32]
    exp = []
    for b in range(3):
        exp += _philox(oracle_lib, [b, 77, 5, (3 << 8) | 1], key)
    assert out.tolist() == exp


def test_gauss_from_tape_consumes_like_stream(oracle_lib):
    """Feeding the ISAAC64 uint32 sequence as a tape reproduces rng_gauss."""
    g = np.load(GOLDEN, allow_pickle=False)
    tape = g["isaac_u32_1"]
    out = np.zeros(500)
    used = oracle_lib.orc_gauss_from_tape(tape.ctypes.data, len(tape), 500, out.ctypes.data)
    assert used > 500
    assert np.array_equal(out, g["gauss_1"][:500])


def test_gauss_moments_and_ks(oracle_lib):
    from scipy import stats
    x = _gauss(oracle_lib, 42, 400000)
    n = len(x)
    se = 1 / np.sqrt(n)
    assert abs(x.mean()) < 5 * se
    assert abs(x.var() - 1) < 5 * np.sqrt(2.0 / n)
    assert abs(stats.skew(x)) < 5 * np.sqrt(6.0 / n)
    assert abs(stats.kurtosis(x)) < 5 * np.sqrt(24.0 / n)
    assert stats.kstest(x, "norm").pvalue > 1e-4
    # the tail strip (|x| > R = 3.4426) is sampled with the right mass
    tail = np.mean(np.abs(x) > 3.442619855899)
    p = 2 * stats.norm.sf(3.442619855899)
    assert abs(tail - p) < 6 * np.sqrt(p / n)


def test_ziggurat_table_identities():
    """YTAB[0] = 1, strictly decreasing; every strip but the base has the same
    area v = x_i * (f(x_{i+1}) - f(x_i)) (Marsaglia & Tsang 2000)."""
    import re
    src = open(os.path.join(os.path.dirname(__file__), "..", "oracle", "ziggurat_tables.h")).read()
    y = [float(v) for v in re.search(r"ZIG_YTAB\[128\] = \{(.*?)\}", src, re.S).group(1).split(",") if v.strip()]
    w = [float(v) for v in re.search(r"ZIG_WTAB\[128\] = \{(.*?)\}", src, re.S).group(1).split(",") if v.strip()]
    y = np.array(y)
    w = np.array(w)
    assert y[0] == 1.0 and np.all(np.diff(y) < 0)
    # x at the right end of strip i (scaled by 2^-32 in WTAB)
    x = w * 2.0 ** 32
    f = np.exp(-0.5 * x ** 2)
    assert np.allclose(f[1:], y[1:], rtol=2e-9, atol=0)
    areas = x[1:] * (y[:-1] - y[1:])
    assert np.allclose(areas, areas.mean(), rtol=1e-6)
