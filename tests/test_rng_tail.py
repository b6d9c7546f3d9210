"""The Ziggurat's tail (src/rng.c:205) computes log1p(-rng_dbl()), and that
value is the Gaussian itself: the engine's log1p_neg (mcell_amd/csrc/device/
rng.h, built here for the host with the same IEEE basic operations the GPU
runs) must return the C library's log1p bit for bit on every input the tail
can give it, -k * 2^-32.  Checked on the smallest and largest 65,536 k and a
stride through all 2^32 (the full range was checked once, with no mismatch).
A wrong constant, branch or summation order shows up as mismatches here."""
import ctypes
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("rngtail") / "librngtail.so"
    subprocess.check_call([hipcc, "-O2", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
                           "-ffp-contract=off", os.path.join(HERE, "rng_shim.cpp"), "-o", str(out)])
    lib = ctypes.CDLL(str(out))
    lib.log1p_tail_mismatches.restype = ctypes.c_uint64
    lib.log1p_tail_mismatches.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                          ctypes.POINTER(ctypes.c_uint64)]
    return lib


@pytest.mark.parametrize("k0,stride,n", [
    (0, 1, 1 << 16),                    # u near 0: |x| < 2^-29 and the unreduced branch
    ((1 << 32) - (1 << 16), 1, 1 << 16),  # u near 1: the reduced branch, large k
    (1257787, 429, 10_000_000),          # a stride through the whole range
    (1257966796 - (1 << 19), 1, 1 << 20),  # around u = 0.2929, where the reduction starts
])
def test_tail_log1p_equals_libm(shim, k0, stride, n):
    first = ctypes.c_uint64()
    bad = shim.log1p_tail_mismatches(k0, stride, n, ctypes.byref(first))
    assert bad == 0, "%d mismatches, first k = %d" % (bad, first.value)
