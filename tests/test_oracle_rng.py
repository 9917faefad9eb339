"""Pins of the oracle's random streams (CPU)."""
import ctypes
import ctypes.util
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle


def test_mt19937_64_standard_known_answer(orc):
    # C++ [rand.predef]: the 10000th invocation of a default-constructed
    # std::mt19937_64 (seed 5489) produces 9981545732273789042
    s = orc.mt64_stream(5489, 10000)
    assert int(s[-1]) == 9981545732273789042


def test_lcg_matches_glibc_rand_r(orc):
    # RansacMatchFilter.cu:45-50 cites glibc's LCG. glibc's rand_r advances the
    # same recurrence x <- 1103515245 x + 12345 (mod 2^32) three times per call
    # and keeps the full 32-bit state; reduced mod 2^31 it must equal every
    # third state of the oracle's masked stream.
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.rand_r.restype = ctypes.c_int
    libc.rand_r.argtypes = [ctypes.POINTER(ctypes.c_uint)]
    for seed in (0, 1, 12345, 22695477 * 7, 2**31 - 5):
        st = ctypes.c_uint(seed)
        states = []
        for _ in range(40):
            libc.rand_r(ctypes.byref(st))
            states.append(st.value & 0x7FFFFFFF)
        got = orc.lcg_stream(seed & 0x7FFFFFFF, 120)
        assert list(got[2::3]) == states


def test_box_muller_moments(orc):
    z = orc.mt64_normals(3, 200000)
    assert abs(z.mean()) < 0.01
    assert abs(z.std() - 1.0) < 0.01
    assert abs((z ** 3).mean()) < 0.03


REF = pyoracle.REF_SAMPLER


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (reference tree absent)")
def test_sampler_matches_reference_ransac_hpp(orc):
    """TableSampler<u32,8> of the reference's ransac.hpp (compiled into oracle/_ref)
    against the oracle's restatement."""
    for n, H, seed in [(100, 50, 7), (4096, 200, 1), (9, 20, 3)]:
        out = subprocess.run([REF, "sample", str(n), str(H), str(seed)], capture_output=True, text=True, check=True)
        ref = np.array([[int(x) for x in l.split()] for l in out.stdout.strip().splitlines()], np.uint32)
        got = orc.sample_indices(seed, n, H)
        assert np.array_equal(got, ref)


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (reference tree absent)")
def test_ransac_loop_matches_reference(orc):
    """The adaptive loop (iteration cap, insane-draw cutoff, best update) of
    ransac<> against the reference's own template, with a hash score."""
    import ctypes as ct

    for n, conf, maxit, seed, mod in [(1000, 0.995, 5000, 3, 50), (300, 0.99, 2000, 9, 7), (50, 0.9, 100, 1, 2),
                                      (64, 0.999, 5000, 4, 1000)]:
        out = subprocess.run([REF, "loop", str(n), str(conf), str(maxit), str(seed), str(mod)], capture_output=True,
                             text=True, check=True).stdout.split()
        best_ref, calls_ref = int(out[0]), int(out[1])
        idx_ref = [int(x) for x in out[2:10]]
        bi = np.zeros(8, np.uint32)
        draws = ct.c_uint32(0)
        best = pyoracle._lib.or_ransac_loop_hash(n, conf, maxit, seed, mod, bi.ctypes.data_as(ct.c_void_p), ct.byref(draws))
        assert best == best_ref
        assert draws.value == calls_ref
        assert list(bi) == idx_ref
