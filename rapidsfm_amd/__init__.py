"""rapidsfm_amd - MI355X (gfx950) two-view matching hot path of RapidSFM.

The compute path is the C ABI of librsfm_amd.so (include/rsfm_amd.h): hand-written
HIP kernels. This module only binds it with ctypes and adds thin helpers that
take torch tensors for device memory and streams. There is no CPU fallback:
importing raises if the library was not built.
"""
import ctypes
import os

from . import _abi
from ._abi import (  # noqa: F401
    BestMatch,
    BfTask,
    BfTop2Task,
    RsfmError,
    ValidMatch,
    lib,
)
from .ops import (  # noqa: F401
    bf_match,
    bf_match_top2,
    desc8b_to_4b,
    filter_matches,
    ransac_f8,
    ransac_f8_finish,
    ransac_sample_indices,
    sqrnorm_u4,
    sqrnorm_u8,
)

__all__ = [
    "lib",
    "RsfmError",
    "BestMatch",
    "BfTask",
    "BfTop2Task",
    "ValidMatch",
    "bf_match",
    "bf_match_top2",
    "desc8b_to_4b",
    "filter_matches",
    "ransac_f8",
    "ransac_f8_finish",
    "ransac_sample_indices",
    "sqrnorm_u8",
    "sqrnorm_u4",
]

LIB_PATH = _abi.LIB_PATH
