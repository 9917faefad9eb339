"""Torch-tensor helpers over the C ABI (device memory and streams only).

Descriptors are uint8 CUDA tensors of shape (n, 128) (or (n, 64) for 4-bit);
best-match outputs are int32 tensors of shape (n, 2) holding the reference's
BestMatch {index, distance} (as uint32 bit patterns; no match = (-1, -1)).
"""
import ctypes

import torch

from ._abi import BfTask, BfTop2Task, Camera, DegensacTask, FinishTask, PoseHTask, PoseTask, RansacTask, RealCamera, UndistortTask, check, lib


def _stream(stream):
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def _side(stream, *tensors):
    """Work enqueued on an explicit stream other than the current one: that stream
    first waits for the current stream (where the inputs and temporaries were
    made), and the temporaries stay allocated until the stream's work is done."""
    if stream is None or stream == torch.cuda.current_stream():
        return
    stream.wait_stream(torch.cuda.current_stream())
    for t in tensors:
        if t is not None:
            t.record_stream(stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else ctypes.c_void_p(0)


def _check_desc(d, width=128):
    if d.dtype != torch.uint8 or d.dim() != 2 or d.shape[1] != width or not d.is_cuda or not d.is_contiguous():
        raise ValueError("descriptors must be contiguous CUDA uint8 tensors of shape (n, %d)" % width)


def sqrnorm_u8(desc, stream=None):
    _check_desc(desc)
    out = torch.empty(desc.shape[0], dtype=torch.int32, device=desc.device)
    check(lib.rsfm_sqrnorm_u8(_ptr(out), _ptr(desc), desc.shape[0], _stream(stream)), "rsfm_sqrnorm_u8")
    return out


def sqrnorm_u4(desc4, stream=None):
    _check_desc(desc4, 64)
    out = torch.empty(desc4.shape[0], dtype=torch.int32, device=desc4.device)
    check(lib.rsfm_sqrnorm_u4(_ptr(out), _ptr(desc4), desc4.shape[0], _stream(stream)), "rsfm_sqrnorm_u4")
    return out


def desc8b_to_4b(desc, stream=None):
    _check_desc(desc)
    out = torch.empty((desc.shape[0], 64), dtype=torch.uint8, device=desc.device)
    check(lib.rsfm_desc8b_to_4b(_ptr(out), _ptr(desc), desc.shape[0], _stream(stream)), "rsfm_desc8b_to_4b")
    return out


def bf_match(pairs, impl="mfma", bidirectional=True, stream=None):
    """Top-1 matching of a list of (descA, descB) pairs in one call.

    impl: "mfma" (rsfm_bf_match_u8), "idp" (rsfm_bf_match_idp_u8) or "u4"
    (rsfm_bf_match_u4 on 64-byte 4-bit descriptors). Returns a list of
    (bestA, bestB) int32 (n, 2) tensors (bestB None if not bidirectional).
    """
    width = 64 if impl == "u4" else 128
    tasks = (BfTask * max(1, len(pairs)))()
    outs = []
    for i, (a, b) in enumerate(pairs):
        _check_desc(a, width)
        _check_desc(b, width)
        ba = torch.empty((a.shape[0], 2), dtype=torch.int32, device=a.device)
        bb = torch.empty((b.shape[0], 2), dtype=torch.int32, device=a.device) if bidirectional else None
        tasks[i] = BfTask(a.shape[0], b.shape[0], _ptr(a), _ptr(b), _ptr(ba), _ptr(bb) if bb is not None else None, None, None)
        outs.append((ba, bb))
    fn = {"mfma": lib.rsfm_bf_match_u8, "idp": lib.rsfm_bf_match_idp_u8, "u4": lib.rsfm_bf_match_u4}[impl]
    check(fn(ctypes.cast(tasks, ctypes.c_void_p), len(pairs), _stream(stream)), "bf_match(%s)" % impl)
    return outs


def bf_match_top2(pairs, stream=None, workspace=None):
    """Top-2 query side + top-1 train side for a list of (descA, descB).

    Returns a list of (bestA, secondA, bestB) int32 (n, 2) tensors.
    """
    tasks = (BfTop2Task * max(1, len(pairs)))()
    outs = []
    for i, (a, b) in enumerate(pairs):
        _check_desc(a)
        _check_desc(b)
        ba = torch.empty((a.shape[0], 2), dtype=torch.int32, device=a.device)
        sa = torch.empty((a.shape[0], 2), dtype=torch.int32, device=a.device)
        bb = torch.empty((b.shape[0], 2), dtype=torch.int32, device=a.device)
        tasks[i] = BfTop2Task(BfTask(a.shape[0], b.shape[0], _ptr(a), _ptr(b), _ptr(ba), _ptr(bb), None, None), _ptr(sa))
        outs.append((ba, sa, bb))
    tp = ctypes.cast(tasks, ctypes.c_void_p)
    need = ctypes.c_size_t(0)
    check(lib.rsfm_bf_match_top2_workspace_size(tp, len(pairs), ctypes.byref(need)), "top2 workspace size")
    ws_ptr = 0
    if need.value > 0:
        # the library wants a 256-B aligned pointer with need bytes behind it
        fits = False
        if workspace is not None:
            ws_ptr = (workspace.data_ptr() + 255) // 256 * 256
            fits = workspace.numel() * workspace.element_size() - (ws_ptr - workspace.data_ptr()) >= need.value
        if not fits:
            workspace = torch.empty(need.value + 256, dtype=torch.uint8, device=pairs[0][0].device)
            ws_ptr = (workspace.data_ptr() + 255) // 256 * 256
            if stream is not None:
                # the temporary is freed on return: keep it until the work on stream is done
                workspace.record_stream(stream)
    check(lib.rsfm_bf_match_top2(tp, len(pairs), ctypes.c_void_p(ws_ptr), need.value, _stream(stream)), "bf_match_top2")
    return outs


def ransac_f7(problems, threshold, mode=1, min_sep_sq=(0.0, 0.0), want_models=False, stream=None):
    """7-point hypotheses: like ransac_f8 with sampleIdx of shape (H, 7); each
    hypothesis reports its best of up to three F."""
    return _ransac(lib.rsfm_ransac_f7, "rsfm_ransac_f7", problems, threshold, mode, min_sep_sq, want_models, stream)


def degensac_refine(problems, threshold_f, threshold_h, mode=0, nb_candidates=8, nb_plane_samples=256, seed=0,
                    want_masks=True, stream=None):
    """DEGENSAC finisher over 7-point samples already scored by ransac_f7.

    problems: list of (pts, sampleIdx (H, 7), counts (H,)[, nDev]) CUDA tensors.
    Returns a list of dicts with F (3, 3), H (3, 3), result (4,) = [F inliers,
    H inliers, 1 if F came from plane and parallax, best sample], and the inlier
    masks mask_f / mask_h (n,) when asked.
    """
    tasks = (DegensacTask * max(1, len(problems)))()
    outs = []
    for i, prob in enumerate(problems):
        pts, idx, counts = prob[0], prob[1], prob[2]
        ndev = prob[3] if len(prob) > 3 else None
        if pts.dtype != torch.float32 or pts.dim() != 2 or pts.shape[1] != 4 or not pts.is_contiguous():
            raise ValueError("pts must be contiguous float32 (n, 4)")
        if idx.dim() != 2 or idx.shape[1] != 7 or not idx.is_contiguous() or counts.shape[0] != idx.shape[0]:
            raise ValueError("sampleIdx must be contiguous (H, 7) with counts (H,)")
        n, dev = pts.shape[0], pts.device
        o = {
            "F": torch.empty((3, 3), dtype=torch.float32, device=dev),
            "H": torch.empty((3, 3), dtype=torch.float32, device=dev),
            "result": torch.empty(4, dtype=torch.int32, device=dev),
            "mask_f": torch.empty(n, dtype=torch.uint8, device=dev) if want_masks else None,
            "mask_h": torch.empty(n, dtype=torch.uint8, device=dev) if want_masks else None,
        }
        tasks[i] = DegensacTask(_ptr(pts), n, _ptr(ndev) if ndev is not None else None, _ptr(idx), _ptr(counts),
                                idx.shape[0], _ptr(o["F"]), _ptr(o["H"]), _ptr(o["result"]),
                                _ptr(o["mask_f"]) if want_masks else None, _ptr(o["mask_h"]) if want_masks else None)
        outs.append(o)
    check(lib.rsfm_degensac_refine(ctypes.cast(tasks, ctypes.c_void_p), len(problems), float(threshold_f),
                                   float(threshold_h), int(mode), int(nb_candidates), int(nb_plane_samples),
                                   int(seed) & 0xFFFFFFFFFFFFFFFF, _stream(stream)), "rsfm_degensac_refine")
    return outs


def undistort_in_place(px, py, cam, distortion, n_iters=2, stream=None):
    """Undistort pixel coordinates in place (px, py: contiguous float32 CUDA tensors);
    cam = (fx, fy, cx, cy), distortion = (k1, k2, p1, p2, k3)."""
    if px.dtype != torch.float32 or py.dtype != torch.float32 or px.shape != py.shape:
        raise ValueError("px, py must be float32 tensors of one shape")
    if px.numel() and not (px.is_cuda and py.is_cuda and px.is_contiguous() and py.is_contiguous()):
        raise ValueError("px, py must be contiguous device tensors")
    rc = RealCamera(Camera(*[float(v) for v in cam]), (ctypes.c_float * 5)(*[float(v) for v in distortion]))
    check(lib.rsfm_undistort_in_place(_ptr(px), _ptr(py), px.numel(), ctypes.byref(rc), int(n_iters), _stream(stream)),
          "rsfm_undistort_in_place")


def undistort_pt_pairs(problems, n_iters=2, stream=None):
    """Undistort gathered keypoint pairs in place: problems of (pts float32 CUDA (n, 4), cam0, cam1, dist0,
    dist1[, nDev]) with cameras (fx, fy, cx, cy) and distortions (k1, k2, p1, p2, k3)."""
    tasks = (UndistortTask * max(1, len(problems)))()
    for i, prob in enumerate(problems):
        pts, c0, c1, k0, k1 = prob[:5]
        ndev = prob[5] if len(prob) > 5 else None
        if pts.dtype != torch.float32 or pts.dim() != 2 or pts.shape[1] != 4 or not pts.is_contiguous():
            raise ValueError("pts must be contiguous float32 (n, 4)")
        mk = lambda c, k: RealCamera(Camera(*[float(v) for v in c]), (ctypes.c_float * 5)(*[float(v) for v in k]))  # noqa: E731
        tasks[i] = UndistortTask(_ptr(pts), pts.shape[0], _ptr(ndev) if ndev is not None else None, mk(c0, k0),
                                 mk(c1, k1))
    check(lib.rsfm_undistort_pt_pairs(ctypes.cast(tasks, ctypes.c_void_p), len(problems), int(n_iters),
                                      _stream(stream)), "rsfm_undistort_pt_pairs")


def pose_from_f(problems, front_ratio=0.95, stream=None):
    """Relative pose from F for a list of (F (3, 3) float32, cam0, cam1, pts (n, 4)[, mask (n,) u8[, nDev]])
    with cameras (fx, fy, cx, cy). Returns dicts with poses (4, 12) = [R row-major, t] in the reference's
    candidate order and result (10,) = front counts [4], order by count [4], kept, points considered."""
    tasks = (PoseTask * max(1, len(problems)))()
    outs = []
    for i, prob in enumerate(problems):
        F, c0, c1, pts = prob[:4]
        mask = prob[4] if len(prob) > 4 else None
        ndev = prob[5] if len(prob) > 5 else None
        if pts.dtype != torch.float32 or pts.dim() != 2 or pts.shape[1] != 4 or not pts.is_contiguous():
            raise ValueError("pts must be contiguous float32 (n, 4)")
        Fd = F.to(torch.float32).contiguous()
        o = {"poses": torch.empty((4, 12), dtype=torch.float32, device=pts.device),
             "result": torch.empty(10, dtype=torch.int32, device=pts.device), "_F": Fd}
        _side(stream, Fd)
        tasks[i] = PoseTask(_ptr(Fd), Camera(*[float(v) for v in c0]), Camera(*[float(v) for v in c1]), _ptr(pts),
                            pts.shape[0], _ptr(ndev) if ndev is not None else None,
                            _ptr(mask) if mask is not None else None, _ptr(o["poses"]), _ptr(o["result"]))
        outs.append(o)
    check(lib.rsfm_pose_from_f(ctypes.cast(tasks, ctypes.c_void_p), len(problems), float(front_ratio), _stream(stream)),
          "rsfm_pose_from_f")
    return outs


def pose_from_h(problems, front_ratio=0.95, stream=None):
    """Relative pose from a pixel homography for a list of (H (3, 3) float32, cam0, cam1, pts (n, 4)[, mask
    (n,) u8[, nDev]]). Returns dicts with poses (8, 12) (the valid candidates first, in the reference's
    order), normals (8, 3) and result (19,) = front counts [8], order by count [8], kept, points
    considered, number of valid candidates."""
    tasks = (PoseHTask * max(1, len(problems)))()
    outs = []
    for i, prob in enumerate(problems):
        H, c0, c1, pts = prob[:4]
        mask = prob[4] if len(prob) > 4 else None
        ndev = prob[5] if len(prob) > 5 else None
        if pts.dtype != torch.float32 or pts.dim() != 2 or pts.shape[1] != 4 or not pts.is_contiguous():
            raise ValueError("pts must be contiguous float32 (n, 4)")
        Hd = H.to(torch.float32).contiguous()
        o = {"poses": torch.empty((8, 12), dtype=torch.float32, device=pts.device),
             "normals": torch.empty((8, 3), dtype=torch.float32, device=pts.device),
             "result": torch.empty(19, dtype=torch.int32, device=pts.device), "_H": Hd}
        _side(stream, Hd)
        tasks[i] = PoseHTask(_ptr(Hd), Camera(*[float(v) for v in c0]), Camera(*[float(v) for v in c1]), _ptr(pts),
                             pts.shape[0], _ptr(ndev) if ndev is not None else None,
                             _ptr(mask) if mask is not None else None, _ptr(o["poses"]), _ptr(o["normals"]),
                             _ptr(o["result"]))
        outs.append(o)
    check(lib.rsfm_pose_from_h(ctypes.cast(tasks, ctypes.c_void_p), len(problems), float(front_ratio), _stream(stream)),
          "rsfm_pose_from_h")
    return outs


def ransac_h4(problems, threshold, min_sep_sq=(0.0, 0.0), want_models=False, stream=None):
    """4-point homography hypotheses: problems of (pts, sampleIdx (H, 4)[, nDev]);
    the inlier test is checkProjectiveTransformation's with threshold t pixels."""
    tasks, outs = _ransac_tasks(problems, want_models)
    check(lib.rsfm_ransac_h4(ctypes.cast(tasks, ctypes.c_void_p), len(problems), float(threshold),
                             float(min_sep_sq[0]), float(min_sep_sq[1]), _stream(stream)), "rsfm_ransac_h4")
    return outs


def ransac_f8(problems, threshold, mode=1, min_sep_sq=(0.0, 0.0), want_models=False, stream=None):
    """Score 8-point hypotheses for a list of (pts, sampleIdx[, nDev]) problems.

    pts: float32 CUDA (n, 4) [x0, y0, x1, y1]; sampleIdx: int32/uint32 CUDA (H, 8);
    nDev: optional int32 CUDA (1,) count. mode 0 = reference two-sided line test,
    1 = Sampson. Returns a list of dicts with counts (H,), best (2,), mask (n,) and
    models (H, 3, 3) when asked.
    """
    return _ransac(lib.rsfm_ransac_f8, "rsfm_ransac_f8", problems, threshold, mode, min_sep_sq, want_models, stream)


def ransac_f8_finish(problems, threshold, mode=1, confidence=0.0, max_iter=5000, refit=True, stream=None):
    """End of the 8-point driver for problems of (pts, sampleIdx (H, 8), counts (H,)[, models (H, 3, 3)[, nDev]])
    scored by ransac_f8 with the same threshold and mode: optional adaptive cut (confidence > 0), refit on the
    best F's inliers (kept when it has more). Returns dicts with F (3, 3), result (4,) = [best hypothesis, its
    inliers, hypotheses walked, final inliers] and mask (n,)."""
    tasks = (FinishTask * max(1, len(problems)))()
    outs = []
    for i, prob in enumerate(problems):
        pts, idx, counts = prob[0], prob[1], prob[2]
        models = prob[3] if len(prob) > 3 else None
        ndev = prob[4] if len(prob) > 4 else None
        if pts.dtype != torch.float32 or pts.dim() != 2 or pts.shape[1] != 4 or not pts.is_contiguous():
            raise ValueError("pts must be contiguous float32 (n, 4)")
        if idx.dim() != 2 or idx.shape[1] != 8 or not idx.is_contiguous() or counts.shape[0] != idx.shape[0]:
            raise ValueError("sampleIdx must be contiguous (H, 8) with counts (H,)")
        n, dev = pts.shape[0], pts.device
        o = {"F": torch.empty((3, 3), dtype=torch.float32, device=dev),
             "result": torch.empty(4, dtype=torch.int32, device=dev),
             "mask": torch.empty(n, dtype=torch.uint8, device=dev)}
        tasks[i] = FinishTask(_ptr(pts), n, _ptr(ndev) if ndev is not None else None, _ptr(idx), _ptr(counts),
                              _ptr(models) if models is not None else None, idx.shape[0], _ptr(o["F"]),
                              _ptr(o["result"]), _ptr(o["mask"]))
        outs.append(o)
    check(lib.rsfm_ransac_f8_finish(ctypes.cast(tasks, ctypes.c_void_p), len(problems), float(threshold), int(mode),
                                    float(confidence), int(max_iter), int(bool(refit)), _stream(stream)),
          "rsfm_ransac_f8_finish")
    return outs


def _ransac_tasks(problems, want_models):
    tasks = (RansacTask * max(1, len(problems)))()
    outs = []
    for i, prob in enumerate(problems):
        pts, idx = prob[0], prob[1]
        ndev = prob[2] if len(prob) > 2 else None
        n, H = pts.shape[0], idx.shape[0]
        if pts.dtype != torch.float32 or pts.dim() != 2 or pts.shape[1] != 4 or not pts.is_contiguous():
            raise ValueError("pts must be contiguous float32 (n, 4)")
        dev = pts.device
        o = {
            "counts": torch.empty(H, dtype=torch.int32, device=dev),
            "best": torch.empty(2, dtype=torch.int32, device=dev),
            "mask": torch.empty(n, dtype=torch.uint8, device=dev),
            "models": torch.empty((H, 3, 3), dtype=torch.float32, device=dev) if want_models else None,
        }
        tasks[i] = RansacTask(_ptr(pts), n, _ptr(ndev) if ndev is not None else None, _ptr(idx), H, _ptr(o["counts"]),
                              _ptr(o["models"]) if want_models else None, _ptr(o["best"]), _ptr(o["mask"]))
        outs.append(o)
    return tasks, outs


def _ransac(fn, name, problems, threshold, mode, min_sep_sq, want_models, stream):
    tasks, outs = _ransac_tasks(problems, want_models)
    check(fn(ctypes.cast(tasks, ctypes.c_void_p), len(problems), float(threshold), int(mode), float(min_sep_sq[0]),
             float(min_sep_sq[1]), _stream(stream)), name)
    return outs


def ransac_sample_indices(n, H, seed, n_dev=None, device=None, stream=None):
    idx = torch.empty((H, 8), dtype=torch.int32, device=device or "cuda")
    check(lib.rsfm_ransac_sample_indices(_ptr(idx), H, n, _ptr(n_dev) if n_dev is not None else None, seed,
                                         _stream(stream)), "rsfm_ransac_sample_indices")
    return idx


def filter_matches(fwd, second, bwd, n_train, ratio=(4, 5), stream=None):
    """Device cross-check (+ ratio test when second is given).

    Returns (matches int32 (nQ, 3) [query, train, distance-float-bits], count tensor).
    """
    nq = fwd.shape[0]
    out = torch.empty((max(1, nq), 3), dtype=torch.int32, device=fwd.device)
    count = torch.zeros(1, dtype=torch.int32, device=fwd.device)
    num, den = ratio
    check(
        lib.rsfm_filter_matches(_ptr(fwd), _ptr(second) if second is not None else None, nq,
                                _ptr(bwd) if bwd is not None else None, n_train, num, den,
                                _ptr(out), _ptr(count), None, 0, _stream(stream)),
        "rsfm_filter_matches",
    )
    return out, count
