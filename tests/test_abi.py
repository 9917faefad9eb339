"""The C ABI library loads and exports every symbol include/*.h declares (CPU)."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(rsfm_[a-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_library_exports_declared_symbols(rs):
    names = _declared()
    assert len(names) >= 10
    missing = [n for n in sorted(names) if not hasattr(rs.lib, n)]
    assert not missing, missing


def test_error_reporting_without_device(rs):
    # argument validation happens before any device call
    bad = (rs.BfTask * 1)()
    bad[0].nbDescA = 5
    rc = rs.lib.rsfm_bf_match_u8(ctypes.cast(bad, ctypes.c_void_p), 1, None)
    assert rc != 0
    assert b"descA" in rs.lib.rsfm_last_error()


def test_host_cross_check_matches_oracle(rs, orc):
    import numpy as np

    A = orc.gen_random_u8(1, 300)
    B = np.concatenate([A[::2], orc.gen_random_u8(2, 100)])
    fwd, bwd = orc.bf_match_u8(A, B)
    out = np.zeros((300, 3), np.uint32)
    n = rs.lib.rsfm_cross_check_host(fwd.ctypes.data_as(ctypes.c_void_p), 300, bwd.ctypes.data_as(ctypes.c_void_p),
                                     B.shape[0], out.ctypes.data_as(ctypes.c_void_p))
    assert np.array_equal(out[:n], orc.cross_check(fwd, bwd))
    n2 = rs.lib.rsfm_remove_conflicts_host(fwd.ctypes.data_as(ctypes.c_void_p), 300, B.shape[0],
                                          out.ctypes.data_as(ctypes.c_void_p))
    assert np.array_equal(out[:n2], orc.remove_conflicts(fwd, B.shape[0]))


def test_host_binning_matches_oracle_and_rejects_nan(rs, orc):
    """rsfm_rmf_bin_host (RansacMatchFilter.cpp:42-67) on the CPU: the oracle's
    cells and order; a non-finite keypoint is an error, not an out-of-range cell."""
    import numpy as np
    import pytest

    from rapidsfm_amd import rmf

    rng = np.random.default_rng(4)
    m = np.stack([rng.uniform(0, 4000, 500), rng.uniform(0, 3000, 500), rng.uniform(0, 4000, 500),
                  rng.uniform(0, 3000, 500)], 1).astype(np.float32)
    m[:3, :2] = [[0, 0], [4000, 3000], [-5, 3100]]  # corners and out-of-image points clamp
    reordered, bounds, order, cols, rows, cw = rmf.bin_matches(4000, 3000, m)
    ref = orc.rmf_bin(4000, 3000, m, 16)
    assert (cols, rows, cw) == tuple(ref[3:6])
    assert np.array_equal(bounds, ref[1]) and np.array_equal(reordered, ref[0]) and np.array_equal(order, ref[2])
    for bad in (np.nan, np.inf, -np.inf):
        mm = m.copy()
        mm[17, 1] = bad
        with pytest.raises(rs.RsfmError, match="non-finite"):
            rmf.bin_matches(4000, 3000, mm)
