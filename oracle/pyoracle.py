"""numpy/ctypes front end of the CPU oracle (liborsfm.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py. The product (rapidsfm_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborsfm.so")
REF_SAMPLER = os.path.join(HERE, "_ref", "ref_sampler")


def build(ref_tree="/root/reference"):
    """Compile the oracle (and, when the reference tree is present, _ref/)."""
    subprocess.run(["make", "-s", "-C", HERE, "liborsfm.so"], check=True)
    if os.path.exists(os.path.join(ref_tree, "ransac.hpp")):
        subprocess.run(["make", "-s", "-C", HERE, "ref", "REF=" + ref_tree], check=True)


# a second instance of this module can bind another build of the same sources
# (the CPU baseline's -march=native library): oracle/cpu_baseline.py
_OTHER = os.environ.get("RSFM_ORACLE_LIB_OVERRIDE")
if _OTHER:
    LIB_PATH = _OTHER
elif not os.path.exists(LIB_PATH):
    build()
_lib = ctypes.CDLL(LIB_PATH)

P = ctypes.c_void_p
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64
I32 = ctypes.c_int
F32 = ctypes.c_float
F64 = ctypes.c_double


def _decl(name, res, args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = args
    return f


class MT64(ctypes.Structure):
    _fields_ = [("mt", ctypes.c_uint64 * 312), ("mti", ctypes.c_int)]


_decl("or_mt64_seed", None, [ctypes.POINTER(MT64), U64])
_decl("or_mt64_next", U64, [ctypes.POINTER(MT64)])
_decl("or_mt64_normal", F64, [ctypes.POINTER(MT64)])
_decl("or_lcg_step", I32, [ctypes.POINTER(ctypes.c_int32)])
_decl("or_gen_sift_pair", None, [U64, U32, U32, U32, F64, P, P, P])
_decl("or_gen_random_u8", None, [U64, U32, P])
_decl("or_gen_two_view", None, [U64, U32, U32, F64, F64, F64, F64, P, P, P])
_decl("or_sample_indices_hash_k", None, [U64, U32, U32, U32, P])
_decl("or_sample_indices", None, [U64, U32, U32, U32, P])
_decl("or_sqrnorm_u8", None, [P, U32, P])
_decl("or_desc8b_to4b", None, [P, U32, P])
_decl("or_sqrnorm_u4", None, [P, U32, P])
_decl("or_bf_match_u8", None, [P, U32, P, U32, P, P, I32])
_decl("or_bf_match_u4", None, [P, U32, P, U32, P, P, I32])
_decl("or_bf_match_top2_u8", None, [P, U32, P, U32, P, P, P, I32])
_decl("or_bf_match_dot_top2_u8", None, [P, U32, P, U32, P, P, P, P, I32])
_decl("or_cross_check", U32, [P, U32, P, U32, P])
_decl("or_remove_conflicts", U32, [P, U32, U32, P])
_decl("or_ratio_pass", I32, [U32, U32, U32, U32])
_decl("or_filter_matches", U32, [P, P, U32, P, U32, U32, U32, P])
_decl("or_cashash_accumulate", None, [P, U32, P])
_decl("or_cashash_compute_bias", None, [P, P, P, U64, U32, P])
_decl("or_cashash_compute_hash", None, [P, P, U32, P, U32, P, P])
_decl("or_cashash_histogram", None, [P, U32, P])
_decl("or_cashash_bucket_bounds", U32, [P, P, P])
_decl("or_cashash_build_buckets", None, [P, U32, P, P])
_decl("or_cashash_find_best", None, [P, P, U32, P, P, U32, U32, P, P])
_decl("or_cashash_find_best_rerank", None, [P, P, P, U32, P, P, P, U32, U32, U32, P])
_decl("or_cashash_gaussian_weights", None, [U64, U32, P, P])
_decl("or_rmf_bin", U32, [I32, I32, U32, P, U32, P, P, P, P, P, P])
_decl("or_rmf_filter", None, [P, P, P, U32, P, U32, U32, U32, F32, I32, U32])
_decl("or_solve_gauss3", None, [P, P, P])
_decl("or_f8_fit_f32", I32, [P, P])
_decl("or_f8_fit_f64", I32, [P, P])
_decl("or_f_inlier_f32", I32, [P, P, F32, I32])
_decl("or_f_inlier_f64", I32, [P, P, F64, I32])
_decl("or_f8_score", None, [P, U32, P, U32, F32, I32, F32, F32, P, P, I32])
_decl("or_f8_ransac_adaptive", U32, [P, U32, F32, I32, F32, U32, U64, F32, F32, P, P])
_decl("or_ransac_loop_hash", U32, [U32, F32, U32, U64, U32, P, P])
_decl("or_ransac_required_iters", U32, [U32, U32, U32, F64, U32])
_decl("or_sample_weighted", None, [P, U32, U32, U64, P])
_decl("or_f7_fit_f32", I32, [P, P])
_decl("or_f7_fit_f64", I32, [P, P])
_decl("or_f7_score", None, [P, U32, P, U32, F32, I32, F32, F32, P, P, I32])
_decl("or_pose_decompose", I32, [P, P, P, P])
_decl("or_pose_triangulate", I32, [P, P, P, P, P])
_decl("or_pose_from_f", None, [P, P, P, P, U32, P, F32, P, P])
_decl("or_pose_decompose_h", I32, [P, P, P, P, P])
_decl("or_pose_from_h", None, [P, P, P, P, U32, P, F32, P, P, P])
_decl("or_distort_f32", None, [P, F32, F32, P])
_decl("or_undistort_in_place", None, [P, P, U32, P, P, U32])
_decl("or_h4_fit_f32", I32, [P, P])
_decl("or_h4_score", None, [P, U32, P, U32, F32, F32, F32, P, P, I32])
_decl("or_fit_ls_f", I32, [P, P, U32, P])
_decl("or_fit_ls_h", I32, [P, P, U32, P])
_decl("or_h_from_f", I32, [P, P, P])
_decl("or_f_from_h", I32, [P, P, P, P])
_decl("or_h_inlier_f32", I32, [P, P, F32])
_decl("or_degensac_refine", None, [P, U32, P, P, U32, F32, F32, I32, U32, U32, U64, P, P, P, P, P])
_decl("or_det_log2", F64, [F64])
_decl("or_ransac_required_iters_det", U32, [U32, U32, F32, U32])
_decl("or_ransac_adaptive_cut", U32, [P, U32, U32, F32, U32, ctypes.POINTER(U32), ctypes.POINTER(U32)])
_decl("or_f8_finish", I32, [P, U32, P, P, P, U32, F32, I32, F32, U32, I32, P, P, P])
_decl("or_mlp_forward", None, [P, P, U32, P])
_decl("or_mlp_loss_grad", F64, [P, P, P, U32, P])
_decl("or_mlp_forward_bf16", None, [P, P, U32, P])
_decl("or_mlp_loss_grad_bf16", F64, [P, P, P, U32, P])
_decl("or_bf16_round", F64, [F64])
_decl("or_proposal_features", None, [P, U32, P, P, P, F32, F32, P])

MLP_NPARAM = 64 * 8 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1


def _p(a):
    return a.ctypes.data_as(P) if a is not None else None


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ---------------- RNG ----------------
def mt64_stream(seed, n):
    s = MT64()
    _lib.or_mt64_seed(ctypes.byref(s), seed)
    return np.array([_lib.or_mt64_next(ctypes.byref(s)) for _ in range(n)], dtype=np.uint64)


def mt64_normals(seed, n):
    s = MT64()
    _lib.or_mt64_seed(ctypes.byref(s), seed)
    return np.array([_lib.or_mt64_normal(ctypes.byref(s)) for _ in range(n)])


def lcg_stream(state, n):
    st = ctypes.c_int32(state)
    return np.array([_lib.or_lcg_step(ctypes.byref(st)) for _ in range(n)], dtype=np.int64)


# ---------------- generators ----------------
def gen_sift_pair(seed, n_train, n_query, n_true, noise_sigma=0.05):
    train = np.zeros((n_train, 128), np.uint8)
    query = np.zeros((n_query, 128), np.uint8)
    gt = np.zeros(n_query, np.int32)
    _lib.or_gen_sift_pair(seed, n_train, n_query, n_true, noise_sigma, _p(train), _p(query), _p(gt))
    return train, query, gt


def gen_random_u8(seed, n):
    d = np.zeros((n, 128), np.uint8)
    _lib.or_gen_random_u8(seed, n, _p(d))
    return d


def gen_two_view(seed, n, n_out, focal=2000.0, width=4000.0, height=3000.0, noise_px=0.5):
    pts = np.zeros((n, 4), np.float32)
    inl = np.zeros(n, np.uint8)
    F = np.zeros(9, np.float64)
    _lib.or_gen_two_view(seed, n, n_out, focal, width, height, noise_px, _p(pts), _p(inl), _p(F))
    return pts, inl.astype(bool), F.reshape(3, 3)


def sample_indices(seed, n, H, k=8):
    idx = np.zeros((H, k), np.uint32)
    _lib.or_sample_indices(seed, n, k, H, _p(idx))
    return idx


# ---------------- brute force ----------------
def sqrnorm_u8(desc):
    desc = _c(desc, np.uint8)
    out = np.zeros(desc.shape[0], np.uint32)
    _lib.or_sqrnorm_u8(_p(desc), desc.shape[0], _p(out))
    return out


def desc8b_to4b(desc):
    desc = _c(desc, np.uint8)
    out = np.zeros((desc.shape[0], 64), np.uint8)
    _lib.or_desc8b_to4b(_p(desc), desc.shape[0], _p(out))
    return out


def sqrnorm_u4(desc4):
    desc4 = _c(desc4, np.uint8)
    out = np.zeros(desc4.shape[0], np.uint32)
    _lib.or_sqrnorm_u4(_p(desc4), desc4.shape[0], _p(out))
    return out


def _bm(n):
    return np.zeros((n, 2), np.uint32)


def bf_match_u8(A, B, nthreads=0):
    A = _c(A, np.uint8)
    B = _c(B, np.uint8)
    ba, bb = _bm(A.shape[0]), _bm(B.shape[0])
    _lib.or_bf_match_u8(_p(A), A.shape[0], _p(B), B.shape[0], _p(ba), _p(bb), nthreads)
    return ba, bb


def bf_match_u4(A4, B4, nthreads=0):
    A4 = _c(A4, np.uint8)
    B4 = _c(B4, np.uint8)
    ba, bb = _bm(A4.shape[0]), _bm(B4.shape[0])
    _lib.or_bf_match_u4(_p(A4), A4.shape[0], _p(B4), B4.shape[0], _p(ba), _p(bb), nthreads)
    return ba, bb


def bf_match_top2_u8(A, B, nthreads=0):
    A = _c(A, np.uint8)
    B = _c(B, np.uint8)
    ba, sa, bb = _bm(A.shape[0]), _bm(A.shape[0]), _bm(B.shape[0])
    _lib.or_bf_match_top2_u8(_p(A), A.shape[0], _p(B), B.shape[0], _p(ba), _p(sa), _p(bb), nthreads)
    return ba, sa, bb


def cross_check(fwd, bwd):
    fwd = _c(fwd, np.uint32)
    bwd = _c(bwd, np.uint32)
    out = np.zeros((max(1, fwd.shape[0]), 3), np.uint32)
    n = _lib.or_cross_check(_p(fwd), fwd.shape[0], _p(bwd), bwd.shape[0], _p(out))
    return out[:n]


def remove_conflicts(fwd, n_train):
    fwd = _c(fwd, np.uint32)
    out = np.zeros((max(1, fwd.shape[0]), 3), np.uint32)
    n = _lib.or_remove_conflicts(_p(fwd), fwd.shape[0], n_train, _p(out))
    return out[:n]


def filter_matches(fwd, second, bwd, ratio=(4, 5)):
    fwd = _c(fwd, np.uint32)
    second = _c(second, np.uint32) if second is not None else None
    bwd = _c(bwd, np.uint32) if bwd is not None else None
    out = np.zeros((max(1, fwd.shape[0]), 3), np.uint32)
    nt = bwd.shape[0] if bwd is not None else 0xFFFFFFFF
    n = _lib.or_filter_matches(_p(fwd), _p(second), fwd.shape[0], _p(bwd), nt, ratio[0], ratio[1], _p(out))
    return out[:n]


def ratio_pass(d1, d2, num=4, den=5):
    return bool(_lib.or_ratio_pass(d1, d2, num, den))

_decl("or_sample_indices_hash", None, [U64, U32, U32, P])
_decl("or_gen_match_pair", None, [U64, U32, U32, U32, F64, F64, F64, F64, F64, P, P, P, P, P, P])
_decl("or_bf_match_top2_fast", None, [P, U32, P, U32, P, P, P, I32])


# ---------------- affine match filter ----------------
def rmf_bin(img_w, img_h, matches, cell_cols=16):
    m = _c(matches, np.float32).reshape(-1, 4)
    n = m.shape[0]
    cols, rows, cw = ctypes.c_uint32(0), ctypes.c_uint32(0), ctypes.c_uint32(0)
    mx = max(img_w, img_h)
    cwv = (mx + cell_cols - 1) // cell_cols
    ncell = ((img_w + cwv - 1) // cwv) * ((img_h + cwv - 1) // cwv)
    reordered = np.zeros_like(m)
    bounds = np.zeros(ncell + 1, np.uint32)
    order = np.zeros(n, np.uint32)
    _lib.or_rmf_bin(img_w, img_h, cell_cols, _p(m), n, ctypes.byref(cols), ctypes.byref(rows), ctypes.byref(cw),
                    _p(reordered), _p(bounds), _p(order))
    return reordered, bounds, order, cols.value, rows.value, cw.value


def rmf_filter(reordered, bounds, cols, rows, n_tests, threshold, try_other, seed):
    m = _c(reordered, np.float32).reshape(-1, 4)
    bounds = _c(bounds, np.uint32)
    votes = np.zeros(m.shape[0], np.uint8)
    aff = np.zeros((cols * rows, 6), np.float32)
    _lib.or_rmf_filter(_p(votes), _p(aff), _p(m), m.shape[0], _p(bounds), cols, rows, n_tests, threshold,
                       int(try_other), seed)
    return votes, aff


def solve_gauss3(A, b):
    A = _c(A, np.float32).reshape(9)
    b = _c(b, np.float32).reshape(6)
    x = np.zeros(6, np.float32)
    _lib.or_solve_gauss3(_p(A), _p(b), _p(x))
    return x.reshape(3, 2)


# ---------------- cascade hashing ----------------
def cashash_accumulate(desc, acc=None):
    desc = _c(desc, np.uint8)
    acc = np.zeros(128, np.uint64) if acc is None else acc.copy()
    _lib.or_cashash_accumulate(_p(desc), desc.shape[0], _p(acc))
    return acc


def cashash_compute_bias(w, bias_base, acc, n_desc):
    w = _c(w, np.int8).reshape(-1, 128)
    bias_base = _c(bias_base, np.int32)
    acc = _c(acc, np.uint64)
    out = np.zeros(w.shape[0], np.int32)
    _lib.or_cashash_compute_bias(_p(w), _p(bias_base), _p(acc), n_desc, w.shape[0], _p(out))
    return out


def cashash_compute_hash(w, bias, desc):
    w = _c(w, np.int8).reshape(-1, 128)
    bias = _c(bias, np.int32)
    desc = _c(desc, np.uint8)
    n = desc.shape[0]
    codes = np.zeros((n, 8), np.uint8)
    remap = np.zeros((n, 4), np.uint32) if w.shape[0] > 64 else None
    _lib.or_cashash_compute_hash(_p(w), _p(bias), w.shape[0], _p(desc), n, _p(codes), _p(remap))
    return codes, remap


def cashash_histogram(codes):
    codes = _c(codes, np.uint8)
    h = np.zeros((8, 256), np.uint32)
    _lib.or_cashash_histogram(_p(codes), codes.shape[0], _p(h))
    return h


def cashash_bucket_bounds(hist):
    hist = _c(hist, np.uint32)
    lo = np.zeros((8, 256), np.uint32)
    up = np.zeros((8, 256), np.uint32)
    mx = _lib.or_cashash_bucket_bounds(_p(hist), _p(lo), _p(up))
    return lo, up, mx


def cashash_build_buckets(codes, lower):
    codes = _c(codes, np.uint8)
    lower = _c(lower, np.uint32)
    n = codes.shape[0]
    b = np.zeros((8, n), np.uint32)
    _lib.or_cashash_build_buckets(_p(codes), n, _p(lower), _p(b))
    return b


def cashash_find_best(qdesc, qcodes, tdesc, tcodes, nscheme=8, both=True):
    qdesc, tdesc = _c(qdesc, np.uint8), _c(tdesc, np.uint8)
    qcodes, tcodes = _c(qcodes, np.uint8), _c(tcodes, np.uint8)
    bq = _bm(qdesc.shape[0])
    bt = _bm(tdesc.shape[0]) if both else None
    _lib.or_cashash_find_best(_p(qdesc), _p(qcodes), qdesc.shape[0], _p(tdesc), _p(tcodes), tdesc.shape[0], nscheme,
                              _p(bq), _p(bt))
    return bq, bt


def cashash_find_best_rerank(qdesc, qcodes, qremap, tdesc, tcodes, tremap, nscheme=6, keep=8):
    args = [_c(a, dt) for a, dt in ((qdesc, np.uint8), (qcodes, np.uint8), (qremap, np.uint32), (tdesc, np.uint8),
                                    (tcodes, np.uint8), (tremap, np.uint32))]
    bq = _bm(args[0].shape[0])
    _lib.or_cashash_find_best_rerank(_p(args[0]), _p(args[1]), _p(args[2]), args[0].shape[0], _p(args[3]),
                                     _p(args[4]), _p(args[5]), args[3].shape[0], nscheme, keep, _p(bq))
    return bq


def cashash_gaussian_weights(seed, nch):
    w = np.zeros((nch, 128), np.int8)
    b = np.zeros(nch, np.int32)
    _lib.or_cashash_gaussian_weights(seed, nch, _p(w), _p(b))
    return w, b


def gen_match_pair(seed, nA, nB, n_true, outlier_frac, focal=2000.0, width=4000.0, height=3000.0, noise_px=0.5):
    """descA (nA,128), descB (nB,128), kpA (nA,2), kpB (nB,2), gtA (nA,), geo_inlier (nA,)"""
    dA = np.zeros((nA, 128), np.uint8)
    dB = np.zeros((nB, 128), np.uint8)
    kA = np.zeros((nA, 2), np.float32)
    kB = np.zeros((nB, 2), np.float32)
    gt = np.zeros(nA, np.int32)
    geo = np.zeros(nA, np.uint8)
    _lib.or_gen_match_pair(seed, nA, nB, n_true, outlier_frac, focal, width, height, noise_px, _p(dA), _p(dB), _p(kA),
                           _p(kB), _p(gt), _p(geo))
    return dA, dB, kA, kB, gt, geo.astype(bool)


def bf_match_top2_fast(A, B, nthreads=0):
    A = _c(A, np.uint8)
    B = _c(B, np.uint8)
    ba, sa, bb = _bm(A.shape[0]), _bm(A.shape[0]), _bm(B.shape[0])
    _lib.or_bf_match_top2_fast(_p(A), A.shape[0], _p(B), B.shape[0], _p(ba), _p(sa), _p(bb), nthreads)
    return ba, sa, bb


def sample_weighted(w, H, seed):
    w = _c(w, np.float32)
    out = np.zeros((H, 8), np.uint32)
    _lib.or_sample_weighted(_p(w), len(w), H, seed, _p(out))
    return out


def sample_indices_hash(seed, n, H):
    idx = np.zeros((H, 8), np.uint32)
    _lib.or_sample_indices_hash(seed, n, H, _p(idx))
    return idx


def sample_indices_hash_k(seed, n, k, H):
    idx = np.zeros((H, k), np.uint32)
    _lib.or_sample_indices_hash_k(seed, n, k, H, _p(idx))
    return idx


# ---------------- F-RANSAC ----------------
def f8_fit_f32(p8):
    p8 = _c(p8, np.float32).reshape(8, 4)
    F = np.zeros(9, np.float32)
    ok = _lib.or_f8_fit_f32(_p(p8), _p(F))
    return bool(ok), F.reshape(3, 3)


def f8_fit_f64(p8):
    p8 = _c(p8, np.float64).reshape(8, 4)
    F = np.zeros(9, np.float64)
    ok = _lib.or_f8_fit_f64(_p(p8), _p(F))
    return bool(ok), F.reshape(3, 3)


def f_inlier_f32(F, pts, t, mode):
    F = _c(F, np.float32).reshape(9)
    pts = _c(pts, np.float32).reshape(-1, 4)
    return np.array([_lib.or_f_inlier_f32(_p(F), _p(pts[i]), t, mode) for i in range(pts.shape[0])], bool)


def f_inlier_f64(F, pts, t, mode):
    F = _c(F, np.float64).reshape(9)
    pts = _c(pts, np.float64).reshape(-1, 4)
    return np.array([_lib.or_f_inlier_f64(_p(F), _p(pts[i]), t, mode) for i in range(pts.shape[0])], bool)


def f8_score(pts, idx, t, mode, min_sep_sq0=0.0, min_sep_sq1=0.0, nthreads=0, with_models=True):
    pts = _c(pts, np.float32).reshape(-1, 4)
    idx = _c(idx, np.uint32).reshape(-1, 8)
    H = idx.shape[0]
    counts = np.zeros(H, np.uint32)
    models = np.zeros((H, 9), np.float32) if with_models else None
    _lib.or_f8_score(_p(pts), pts.shape[0], _p(idx), H, t, mode, min_sep_sq0, min_sep_sq1, _p(counts), _p(models),
                     nthreads)
    return counts, (models.reshape(H, 3, 3) if with_models else None)


def f8_ransac_adaptive(pts, t, mode, confidence, max_iter, seed, min_sep_sq0=0.0, min_sep_sq1=0.0):
    pts = _c(pts, np.float32).reshape(-1, 4)
    bi = np.zeros(8, np.uint32)
    draws = ctypes.c_uint32(0)
    best = _lib.or_f8_ransac_adaptive(_p(pts), pts.shape[0], t, mode, confidence, max_iter, seed, min_sep_sq0,
                                      min_sep_sq1, _p(bi), ctypes.byref(draws))
    return best, bi, draws.value


def required_iters(best, n, k, conf, max_iter):
    return _lib.or_ransac_required_iters(best, n, k, conf, max_iter)


def det_log2(x):
    return _lib.or_det_log2(float(x))


def required_iters_det(best, n, conf, max_iter):
    return _lib.or_ransac_required_iters_det(best, n, conf, max_iter)


def adaptive_cut(counts, n, conf, max_iter):
    """(walked, best hypothesis or ~0, its count) of the reference loop over a scored table"""
    counts = _c(counts, np.uint32)
    bh, bc = U32(0), U32(0)
    walked = _lib.or_ransac_adaptive_cut(_p(counts), counts.shape[0], n, conf, max_iter, ctypes.byref(bh),
                                         ctypes.byref(bc))
    return walked, bh.value, bc.value


def f8_finish(pts, idx, counts, models, t, mode, conf=0.0, max_iter=5000, refit=True):
    """rsfm_ransac_f8_finish for one pair: (F (3, 3), result (4,), mask (n,))"""
    pts = _c(pts, np.float32).reshape(-1, 4)
    idx = _c(idx, np.uint32).reshape(-1, 8)
    counts = _c(counts, np.uint32)
    mdl = _c(models, np.float32).reshape(-1, 9) if models is not None else None
    n = pts.shape[0]
    F = np.zeros(9, np.float32)
    res = np.zeros(4, np.uint32)
    mask = np.zeros(max(n, 1), np.uint8)
    _lib.or_f8_finish(_p(pts), n, _p(idx), _p(counts), _p(mdl) if mdl is not None else None, counts.shape[0], t,
                      mode, conf, max_iter, int(refit), _p(F), _p(res), _p(mask))
    return F.reshape(3, 3), res, mask[:n]


# ---------------- learned proposal MLP ----------------
def mlp_forward(params, x, bf16=False):
    """Sigmoid outputs of the 8-64-64-64-1 network for x (n x 8)."""
    params = _c(params, np.float64)
    x = _c(x, np.float32)
    n = x.shape[0]
    out = np.zeros(n, np.float64)
    (_lib.or_mlp_forward_bf16 if bf16 else _lib.or_mlp_forward)(_p(params), _p(x), n, _p(out))
    return out


def mlp_loss_grad(params, x, y, bf16=False, with_grad=True):
    """(mean BCE loss, gradient over the parameter vector)."""
    params = _c(params, np.float64)
    x = _c(x, np.float32)
    y = _c(y, np.float32)
    g = np.zeros(MLP_NPARAM, np.float64) if with_grad else None
    fn = _lib.or_mlp_loss_grad_bf16 if bf16 else _lib.or_mlp_loss_grad
    loss = fn(_p(params), _p(x), _p(y), x.shape[0], _p(g))
    return loss, g


def bf16_round(v):
    return _lib.or_bf16_round(float(v))


def proposal_features(matches, second, kpQ, kpT, width, height):
    """matches: (n, 3) [query, train, distance float bits] (uint32/int32 view);
    second: (nQ, 2) BestMatch; kp*: (n, 4) float32 -> (n, 8) float32."""
    m = _c(matches, np.uint32)
    s = _c(second, np.uint32)
    kq = _c(kpQ, np.float32)
    kt = _c(kpT, np.float32)
    out = np.zeros((m.shape[0], 8), np.float32)
    _lib.or_proposal_features(_p(m), m.shape[0], _p(s), _p(kq), _p(kt), width, height, _p(out))
    return out


# ---------------- 7-point hypotheses ----------------
def f7_fit_f32(p7):
    p = _c(p7, np.float32).reshape(7, 4)
    Fs = np.zeros((3, 9), np.float32)
    n = _lib.or_f7_fit_f32(_p(p), _p(Fs))
    return Fs[:n].reshape(n, 3, 3)


def f7_fit_f64(p7):
    p = _c(p7, np.float64).reshape(7, 4)
    Fs = np.zeros((3, 9), np.float64)
    n = _lib.or_f7_fit_f64(_p(p), _p(Fs))
    return Fs[:n].reshape(n, 3, 3)


def f7_score(pts, idx, t, mode, min_sep_sq0=0.0, min_sep_sq1=0.0, nthreads=0, with_models=True):
    pts = _c(pts, np.float32)
    idx = _c(idx, np.uint32)
    H = idx.shape[0]
    counts = np.zeros(H, np.uint32)
    models = np.zeros((H, 9), np.float32) if with_models else None
    _lib.or_f7_score(_p(pts), pts.shape[0], _p(idx), H, t, mode, min_sep_sq0, min_sep_sq1, _p(counts), _p(models),
                     nthreads)
    return counts, models


# ---------------- DEGENSAC finisher ----------------
def fit_ls_f(pts, sel):
    """least-squares F on pts[sel] (None when degenerate)"""
    pts = _c(pts, np.float32)
    sel = _c(sel, np.uint32)
    F = np.zeros(9, np.float32)
    return F.reshape(3, 3) if _lib.or_fit_ls_f(_p(pts), _p(sel), sel.shape[0], _p(F)) else None


def fit_ls_h(pts, sel):
    pts = _c(pts, np.float32)
    sel = _c(sel, np.uint32)
    H = np.zeros(9, np.float32)
    return H.reshape(3, 3) if _lib.or_fit_ls_h(_p(pts), _p(sel), sel.shape[0], _p(H)) else None


def h_from_f(F, pts3):
    F = _c(F, np.float32).reshape(9)
    p = _c(pts3, np.float32).reshape(3, 4)
    H = np.zeros(9, np.float32)
    return H.reshape(3, 3) if _lib.or_h_from_f(_p(F), _p(p), _p(H)) else None


def f_from_h(H, pa, pb):
    H = _c(H, np.float32).reshape(9)
    a = _c(pa, np.float32).reshape(4)
    b = _c(pb, np.float32).reshape(4)
    F = np.zeros(9, np.float32)
    return F.reshape(3, 3) if _lib.or_f_from_h(_p(H), _p(a), _p(b), _p(F)) else None


def h_inlier_f32(H, pts, t):
    H = _c(H, np.float32).reshape(9)
    pts = _c(pts, np.float32).reshape(-1, 4)
    return np.array([_lib.or_h_inlier_f32(_p(H), _p(pts[i]), t) for i in range(pts.shape[0])], bool)


def degensac_refine(pts, idx, counts, t_f, t_h, mode, n_cand, n_plane, seed):
    """rsfm_degensac_refine for one pair: dict of F, H, result, mask_f, mask_h"""
    pts = _c(pts, np.float32)
    idx = _c(idx, np.uint32).reshape(-1, 7)
    counts = _c(counts, np.uint32)
    n = pts.shape[0]
    F = np.zeros(9, np.float32)
    H = np.zeros(9, np.float32)
    res = np.zeros(4, np.uint32)
    mf = np.zeros(max(n, 1), np.uint8)
    mh = np.zeros(max(n, 1), np.uint8)
    _lib.or_degensac_refine(_p(pts), n, _p(idx), _p(counts), idx.shape[0], t_f, t_h, mode, n_cand, n_plane, seed,
                            _p(F), _p(H), _p(res), _p(mf), _p(mh))
    return {"F": F, "H": H, "result": res, "mask_f": mf[:n], "mask_h": mh[:n]}


# ---------------- 4-point homography hypotheses ----------------
def h4_fit_f32(p4):
    p = _c(p4, np.float32).reshape(4, 4)
    H = np.zeros(9, np.float32)
    return H.reshape(3, 3) if _lib.or_h4_fit_f32(_p(p), _p(H)) else None


def h4_score(pts, idx, t, min_sep_sq0=0.0, min_sep_sq1=0.0, nthreads=0, with_models=True):
    pts = _c(pts, np.float32).reshape(-1, 4)
    idx = _c(idx, np.uint32).reshape(-1, 4)
    H = idx.shape[0]
    counts = np.zeros(H, np.uint32)
    models = np.zeros((H, 9), np.float32) if with_models else None
    _lib.or_h4_score(_p(pts), pts.shape[0], _p(idx), H, t, min_sep_sq0, min_sep_sq1, _p(counts), _p(models), nthreads)
    return counts, models


# ---------------- relative pose from F ----------------
def pose_decompose(F, cam0, cam1):
    """four (R, t) candidates as (4, 12) float32 [R row-major, t], or None"""
    F = _c(F, np.float32).reshape(9)
    c0 = _c(cam0, np.float32).reshape(4)
    c1 = _c(cam1, np.float32).reshape(4)
    P = np.zeros(48, np.float32)
    return P.reshape(4, 12) if _lib.or_pose_decompose(_p(F), _p(c0), _p(c1), _p(P)) else None


def pose_triangulate(P, p, cam0, cam1):
    P = _c(P, np.float32).reshape(12)
    p = _c(p, np.float32).reshape(4)
    c0 = np.asarray(cam0, np.float64)
    c1 = np.asarray(cam1, np.float64)
    inv = []
    for c in (c0, c1):
        fx = np.float32(1.0) / np.float32(c[0])
        fy = np.float32(1.0) / np.float32(c[1])
        inv.append(np.array([fx, fy, -np.float32(c[2]) * fx, -np.float32(c[3]) * fy], np.float32))
    X = np.zeros(3, np.float32)
    front = _lib.or_pose_triangulate(_p(P), _p(p), _p(inv[0]), _p(inv[1]), _p(X))
    return X, bool(front)


def pose_from_f(F, cam0, cam1, pts, mask=None, ratio=0.95):
    F = _c(F, np.float32).reshape(9)
    c0 = _c(cam0, np.float32).reshape(4)
    c1 = _c(cam1, np.float32).reshape(4)
    pts = _c(pts, np.float32).reshape(-1, 4)
    m = _c(mask, np.uint8) if mask is not None else None
    poses = np.zeros(48, np.float32)
    res = np.zeros(10, np.uint32)
    _lib.or_pose_from_f(_p(F), _p(c0), _p(c1), _p(pts), pts.shape[0], _p(m), ratio, _p(poses), _p(res))
    return poses.reshape(4, 12), res


# ---------------- keypoint undistortion ----------------
def distort_f32(k, x0, x1):
    k = _c(k, np.float32).reshape(5)
    out = np.zeros(2, np.float32)
    _lib.or_distort_f32(_p(k), x0, x1, _p(out))
    return out


def undistort(px, py, cam, k, iters=2):
    """undistorted copies of the pixel arrays px, py"""
    x = _c(px, np.float32).copy()
    y = _c(py, np.float32).copy()
    c = _c(cam, np.float32).reshape(4)
    kk = _c(k, np.float32).reshape(5)
    _lib.or_undistort_in_place(_p(x), _p(y), x.shape[0], _p(c), _p(kk), iters)
    return x, y


# ---------------- relative pose from H ----------------
def pose_decompose_h(H, cam0, cam1):
    """(poses [n][12], normals [n][3]) of the valid homography candidates"""
    Hm = _c(H, np.float32).reshape(9)
    c0 = _c(cam0, np.float32).reshape(4)
    c1 = _c(cam1, np.float32).reshape(4)
    P = np.zeros(96, np.float32)
    N = np.zeros(24, np.float32)
    n = _lib.or_pose_decompose_h(_p(Hm), _p(c0), _p(c1), _p(P), _p(N))
    return P.reshape(8, 12)[:n], N.reshape(8, 3)[:n]


def pose_from_h(H, cam0, cam1, pts, mask=None, ratio=0.95):
    """(poses [8][12], normals [8][3], result [19]) as rsfm_pose_from_h writes them"""
    Hm = _c(H, np.float32).reshape(9)
    c0 = _c(cam0, np.float32).reshape(4)
    c1 = _c(cam1, np.float32).reshape(4)
    pts = _c(pts, np.float32).reshape(-1, 4)
    m = None if mask is None else _c(mask, np.uint8)
    poses = np.zeros(96, np.float32)
    normals = np.zeros(24, np.float32)
    res = np.zeros(19, np.uint32)
    _lib.or_pose_from_h(_p(Hm), _p(c0), _p(c1), _p(pts), pts.shape[0], _p(m), ratio, _p(poses), _p(normals),
                        _p(res))
    return poses.reshape(8, 12), normals.reshape(8, 3), res
