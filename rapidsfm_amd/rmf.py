"""Affine RANSAC match filter over the C ABI (rsfm_rmf_*): the Python mirror of
RansacMatchFilter::getInlierMask (RansacMatchFilter.cpp:40-135)."""
import ctypes

import numpy as np
import torch

from ._abi import bind_extra, check, lib
from .ops import _ptr, _stream

c_void_p, c_u32, c_int, c_float = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_float

bind_extra("rsfm_rmf_filter", c_int,
           [c_void_p, c_void_p, c_void_p, c_u32, c_void_p, c_u32, c_u32, c_u32, c_float, c_int, c_u32, c_void_p])
bind_extra("rsfm_rmf_bin_host", c_int,
           [c_int, c_int, c_u32, c_void_p, c_u32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p])


def bin_matches(img_w, img_h, matches, cell_cols=16):
    """matches: float32 (n, 4) [left x, left y, right x, right y] on the host."""
    m = np.ascontiguousarray(matches, np.float32)
    n = m.shape[0]
    cols, rows, cw = c_u32(0), c_u32(0), c_u32(0)
    reordered = np.zeros_like(m)
    order = np.zeros(n, np.uint32)
    mx = max(img_w, img_h)
    cwv = (mx + cell_cols - 1) // cell_cols
    ncell = ((img_w + cwv - 1) // cwv) * ((img_h + cwv - 1) // cwv)
    bounds = np.zeros(ncell + 1, np.uint32)
    check(lib.rsfm_rmf_bin_host(img_w, img_h, cell_cols, m.ctypes.data_as(c_void_p), n, ctypes.byref(cols),
                                ctypes.byref(rows), ctypes.byref(cw), reordered.ctypes.data_as(c_void_p),
                                bounds.ctypes.data_as(c_void_p), order.ctypes.data_as(c_void_p)), "rmf bin")
    return reordered, bounds, order, cols.value, rows.value, cw.value


def filter_votes(reordered_dev, bounds_dev, cols, rows, n_tests, threshold, try_other, seed, stream=None):
    """Device votes (uint8 tensor, length padded to 4) and the affine scratch."""
    n = reordered_dev.shape[0]
    votes = torch.empty(((n + 3) // 4) * 4, dtype=torch.uint8, device=reordered_dev.device)
    aff = torch.empty((cols * rows, 6), dtype=torch.float32, device=reordered_dev.device)
    check(lib.rsfm_rmf_filter(_ptr(votes), _ptr(aff), _ptr(reordered_dev), n, _ptr(bounds_dev), cols, rows, n_tests,
                              float(threshold), int(try_other), seed, _stream(stream)), "rmf filter")
    return votes[:n], aff


def get_inlier_mask(img_size0, matches, min_votes=2, n_tests=256, threshold=0.15, cell_cols=16, try_other=False,
                    seed=0, device="cuda"):
    """RansacMatchFilter::getInlierMask with an explicit seed: host binning,
    device filter, mask[i] = votes >= min_votes in the original order."""
    w, h = img_size0
    reordered, bounds, order, cols, rows, cw = bin_matches(w, h, matches, cell_cols)
    rd = torch.from_numpy(reordered).to(device)
    bd = torch.from_numpy(bounds.view(np.int32)).to(device)
    votes, _ = filter_votes(rd, bd, cols, rows, n_tests, cw * threshold, try_other, seed)
    v = votes.cpu().numpy()
    mask = np.zeros(len(order), bool)
    mask[order] = v >= min_votes
    return mask


class RmfTask(ctypes.Structure):
    _fields_ = [("pts", c_void_p), ("matches", c_void_p), ("nbPtsDev", c_void_p), ("nbPts", c_u32),
                ("inlierMask", c_void_p), ("imgWidth", c_int), ("imgHeight", c_int), ("seed", c_u32)]


bind_extra("rsfm_rmf_filter_batch_workspace_size", c_int, [c_void_p, ctypes.c_size_t, c_u32,
                                                            ctypes.POINTER(ctypes.c_size_t)])
bind_extra("rsfm_rmf_filter_batch", c_int, [c_void_p, ctypes.c_size_t, c_u32, c_u32, c_float, c_u32, c_int, c_void_p,
                                            ctypes.c_size_t, c_void_p])


def filter_batch(problems, min_votes=2, n_tests=256, threshold=0.15, cell_cols=16, try_other=True, stream=None):
    """Batched affine match filter on device-resident gathered pairs.

    problems: list of (pts float32 (n, 4) CUDA, count int32 (1,) CUDA, (width, height) of the first image, seed[,
    matches int32 (n, 3) CUDA]). pts (and matches) are compacted in place and count becomes the kept count.
    Returns the masks (uint8 (n,), original order)."""
    tasks = (RmfTask * max(1, len(problems)))()
    masks = []
    for i, prob in enumerate(problems):
        pts, count, (w, h), seed = prob[:4]
        matches = prob[4] if len(prob) > 4 else None
        if pts.dtype != torch.float32 or pts.dim() != 2 or pts.shape[1] != 4 or not pts.is_contiguous():
            raise ValueError("pts must be contiguous float32 (n, 4)")
        mask = torch.empty(max(1, pts.shape[0]), dtype=torch.uint8, device=pts.device)
        tasks[i] = RmfTask(_ptr(pts), _ptr(matches) if matches is not None else None, _ptr(count), pts.shape[0],
                           _ptr(mask), int(w), int(h), int(seed) & 0xFFFFFFFF)
        masks.append(mask[: pts.shape[0]])
    tp = ctypes.cast(tasks, c_void_p)
    need = ctypes.c_size_t(0)
    check(lib.rsfm_rmf_filter_batch_workspace_size(tp, len(problems), cell_cols, ctypes.byref(need)), "rmf batch ws")
    dev = problems[0][0].device
    ws = torch.empty(need.value + 256, dtype=torch.uint8, device=dev)
    wsp = (ws.data_ptr() + 255) // 256 * 256
    if stream is not None:
        ws.record_stream(stream)
    check(lib.rsfm_rmf_filter_batch(tp, len(problems), cell_cols, n_tests, float(threshold), min_votes, int(try_other),
                                    c_void_p(wsp), need.value, _stream(stream)), "rsfm_rmf_filter_batch")
    return masks
