"""Torch-tensor helpers over the cascade-hashing C ABI (rsfm_cashash_*)."""
import ctypes

import torch

from ._abi import bind_extra, check, lib
from .ops import _ptr, _stream

c_void_p, c_u32, c_int, c_u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint64


class CasHashImage(ctypes.Structure):
    _fields_ = [("desc", c_void_p), ("bucketEnds", c_void_p), ("buckets", c_void_p), ("sqrNormDesc", c_void_p)]


class CasHashResult(ctypes.Structure):
    _fields_ = [("query", c_void_p), ("train", c_void_p)]


class CasHashRerankImage(ctypes.Structure):
    _fields_ = [("desc", c_void_p), ("codes", c_void_p), ("remap", c_void_p), ("bucketEnds", c_void_p),
                ("buckets", c_void_p), ("nbDesc", c_u32)]


bind_extra("rsfm_cashash_find_best_match_rerank", c_int, [c_void_p, c_void_p, c_void_p, c_u32, c_u32, c_u32, c_void_p])
bind_extra("rsfm_cashash_accumulate_desc", c_int, [c_void_p, c_void_p, c_int, c_void_p, c_u32, c_void_p])
bind_extra("rsfm_cashash_compute_bias", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_u64, c_u32, c_void_p])
bind_extra("rsfm_cashash_compute_hash_bucket_only", c_int,
           [c_void_p, c_void_p, c_void_p, c_u32, c_void_p, c_void_p, c_void_p])
bind_extra("rsfm_cashash_compute_hash_bucket_and_remap", c_int,
           [c_void_p, c_void_p, c_void_p, c_u32, c_void_p, c_void_p, c_void_p, c_void_p])
bind_extra("rsfm_cashash_build_histogram", c_int, [c_void_p, c_void_p, c_u32, c_void_p])
bind_extra("rsfm_cashash_compute_bucket_bound", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p])
bind_extra("rsfm_cashash_build_buckets", c_int, [c_void_p, c_u32, c_void_p, c_void_p, c_void_p])
bind_extra("rsfm_cashash_find_best_match", c_int,
           [c_void_p, c_void_p, c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_u32, c_u32, c_void_p])


def accumulate(desc, acc=None, init=True, with_sqr=False, stream=None):
    dev = desc.device
    acc = torch.zeros(128, dtype=torch.int64, device=dev) if acc is None else acc
    sqr = torch.zeros(128, dtype=torch.int64, device=dev) if with_sqr else None
    check(lib.rsfm_cashash_accumulate_desc(_ptr(acc), _ptr(sqr) if sqr is not None else None, int(init), _ptr(desc),
                                           desc.shape[0], _stream(stream)), "accumulate")
    return (acc, sqr) if with_sqr else acc


def compute_bias(weights, bias_base, acc, n_desc, stream=None):
    nch = weights.shape[0]
    out = torch.empty(nch, dtype=torch.int32, device=weights.device)
    check(lib.rsfm_cashash_compute_bias(_ptr(out), _ptr(weights), _ptr(bias_base), _ptr(acc), n_desc, nch,
                                        _stream(stream)), "compute_bias")
    return out


def compute_hash(weights, bias, desc, with_histogram=False, stream=None):
    """Returns (codes uint8 (n, 8), remap int32 (n, 4) or None, histogram int32 (8, 256) or None)."""
    n, dev = desc.shape[0], desc.device
    codes = torch.empty((n, 8), dtype=torch.uint8, device=dev)
    hist = torch.empty((8, 256), dtype=torch.int32, device=dev) if with_histogram else None
    hp = _ptr(hist) if hist is not None else None
    if weights.shape[0] == 64:
        check(lib.rsfm_cashash_compute_hash_bucket_only(_ptr(weights), _ptr(bias), _ptr(desc), n, _ptr(codes), hp,
                                                        _stream(stream)), "compute_hash")
        return codes, None, hist
    remap = torch.empty((n, 4), dtype=torch.int32, device=dev)
    check(lib.rsfm_cashash_compute_hash_bucket_and_remap(_ptr(weights), _ptr(bias), _ptr(desc), n, _ptr(codes),
                                                         _ptr(remap), hp, _stream(stream)), "compute_hash_remap")
    return codes, remap, hist


def build_histogram(codes, stream=None):
    hist = torch.empty((8, 256), dtype=torch.int32, device=codes.device)
    check(lib.rsfm_cashash_build_histogram(_ptr(hist), _ptr(codes), codes.shape[0], _stream(stream)), "histogram")
    return hist


def bucket_bounds(hist, stream=None):
    lo = torch.empty_like(hist)
    up = torch.empty_like(hist)
    mx = torch.empty(1, dtype=torch.int32, device=hist.device)
    check(lib.rsfm_cashash_compute_bucket_bound(_ptr(lo), _ptr(up), _ptr(mx), _ptr(hist), _stream(stream)), "bounds")
    return lo, up, mx


def build_buckets(codes, lower, stream=None):
    n = codes.shape[0]
    alloc = lower.clone()
    buckets = torch.empty((8, n), dtype=torch.int32, device=codes.device)
    check(lib.rsfm_cashash_build_buckets(_ptr(codes), n, _ptr(alloc), _ptr(buckets), _stream(stream)), "buckets")
    return buckets


def register_image(desc, weights, bias, stream=None):
    """hash, histogram, bounds and buckets of one image (the cascade-hash part of
    BruteForceMatcher::registerImage). Returns a dict usable by find_best_match."""
    codes, remap, _ = compute_hash(weights, bias, desc, stream=stream)
    hist = build_histogram(codes, stream)
    lo, up, mx = bucket_bounds(hist, stream)
    buckets = build_buckets(codes, lo, stream)
    return {"desc": desc, "codes": codes, "remap": remap, "ends": up, "buckets": buckets, "n": desc.shape[0],
            "max_bucket": mx}


def find_best_match(queries, train, n_schemes=8, bidirectional=True, stream=None):
    """queries: list of register_image() dicts; train: one dict. Returns a list of
    (bestQ, bestT) int32 (n, 2) tensors."""
    nq = len(queries)
    qimg = (CasHashImage * max(1, nq))()
    qsize = (ctypes.c_uint32 * max(1, nq))()
    res = (CasHashResult * max(1, nq))()
    outs = []
    for i, q in enumerate(queries):
        bq = torch.empty((q["n"], 2), dtype=torch.int32, device=q["desc"].device)
        bt = torch.empty((train["n"], 2), dtype=torch.int32, device=q["desc"].device) if bidirectional else None
        qimg[i] = CasHashImage(_ptr(q["desc"]), _ptr(q["ends"]), _ptr(q["buckets"]), None)
        qsize[i] = q["n"]
        res[i] = CasHashResult(_ptr(bq), _ptr(bt) if bt is not None else None)
        outs.append((bq, bt))
    timg = CasHashImage(_ptr(train["desc"]), _ptr(train["ends"]), _ptr(train["buckets"]), None)
    check(lib.rsfm_cashash_find_best_match(ctypes.cast(qimg, c_void_p), ctypes.cast(qsize, c_void_p),
                                           ctypes.cast(ctypes.pointer(timg), c_void_p),
                                           train["n"], ctypes.cast(res, c_void_p), nq, None, 0, n_schemes,
                                           _stream(stream)), "find_best_match")
    return outs


def _rerank_image(im):
    if im["remap"] is None:
        raise ValueError("the re-rank mode needs the 192-channel hash (bucket codes and remap)")
    return CasHashRerankImage(_ptr(im["desc"]), _ptr(im["codes"]), _ptr(im["remap"]), _ptr(im["ends"]),
                              _ptr(im["buckets"]), im["n"])


def find_best_match_rerank(queries, train, n_schemes=6, keep=8, bidirectional=True, stream=None):
    """Config-1 mode: candidates from the first n_schemes bucket tables, the keep
    Hamming-nearest (128-bit remap) re-ranked by exact L2^2. queries: list of
    register_image() dicts built with 192-channel weights; train: one dict.
    Returns a list of (bestQ, bestT) int32 (n, 2) tensors."""
    nq = len(queries)
    qimg = (CasHashRerankImage * max(1, nq))()
    res = (CasHashResult * max(1, nq))()
    outs = []
    for i, q in enumerate(queries):
        bq = torch.empty((max(1, q["n"]), 2), dtype=torch.int32, device=q["desc"].device)
        bt = torch.empty((max(1, train["n"]), 2), dtype=torch.int32, device=q["desc"].device) if bidirectional else None
        qimg[i] = _rerank_image(q)
        res[i] = CasHashResult(_ptr(bq), _ptr(bt) if bt is not None else None)
        outs.append((bq[: q["n"]], bt[: train["n"]] if bt is not None else None))
    timg = _rerank_image(train)
    check(lib.rsfm_cashash_find_best_match_rerank(ctypes.cast(qimg, c_void_p), ctypes.cast(ctypes.pointer(timg),
                                                  c_void_p), ctypes.cast(res, c_void_p), nq, n_schemes, keep,
                                                  _stream(stream)), "find_best_match_rerank")
    return outs
