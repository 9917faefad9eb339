"""Batched two-view pipeline over the C ABI (rsfm_pipeline_*): many image pairs of
uneven size matched and verified by one HIP graph launch. Torch tensors provide
the device buffers; the library owns only its task tables and partials."""
import ctypes

import torch

from ._abi import Camera, check, lib

c_u32, c_f32, c_vp = ctypes.c_uint32, ctypes.c_float, ctypes.c_void_p


# wg_per_xcd value: one matcher workgroup per tile instead of a persistent grid
WG_PER_TILE = 0xFFFFFFFF
SOLVERS = {"f8": 0, "degensac": 1}


class PipelinePair(ctypes.Structure):
    _fields_ = [("nA", c_u32), ("nB", c_u32), ("descA", c_vp), ("descB", c_vp), ("kpA", c_vp), ("kpB", c_vp),
                ("bestA", c_vp), ("secondA", c_vp), ("bestB", c_vp), ("matches", c_vp), ("count", c_vp),
                ("pts", c_vp), ("sampleIdx", c_vp), ("counts", c_vp), ("best", c_vp), ("inlierMask", c_vp),
                ("F", c_vp), ("H", c_vp), ("solverResult", c_vp), ("inlierMaskH", c_vp), ("cam0", Camera),
                ("cam1", Camera), ("posesF", c_vp), ("poseResultF", c_vp), ("posesH", c_vp), ("normalsH", c_vp),
                ("poseResultH", c_vp), ("distortion", c_vp), ("width0", ctypes.c_int), ("height0", ctypes.c_int)]


class PipelineConfig(ctypes.Structure):
    _fields_ = [("nbHyp", c_u32), ("threshold", c_f32), ("mode", ctypes.c_int), ("ratioNum", c_u32),
                ("ratioDen", c_u32), ("seed", ctypes.c_uint64), ("minSepSq0", c_f32), ("minSepSq1", c_f32),
                ("segCols", c_u32), ("wgPerXcd", c_u32), ("solver", c_u32), ("thresholdH", c_f32),
                ("nbCandidates", c_u32), ("nbPlaneSamples", c_u32), ("frontRatio", c_f32), ("confidence", c_f32),
                ("maxIter", c_u32), ("refit", c_u32), ("matchFilter", c_u32), ("mfCellCols", c_u32),
                ("mfTests", c_u32), ("mfThreshold", c_f32), ("mfMinVotes", c_u32), ("mfTryOther", c_u32)]


class TuningEntry(ctypes.Structure):
    _fields_ = [("segCols", c_u32), ("wgPerXcd", c_u32), ("matchMs", c_f32)]


class PipelineInfo(ctypes.Structure):
    _fields_ = [("segCols", c_u32), ("wgPerXcd", c_u32), ("nbTiles", c_u32), ("tunedMatchMs", c_f32)]


lib.rsfm_pipeline_create.restype = ctypes.c_int
lib.rsfm_pipeline_create.argtypes = [c_vp, ctypes.c_size_t, ctypes.POINTER(PipelineConfig), ctypes.POINTER(c_vp)]
lib.rsfm_pipeline_run.restype = ctypes.c_int
lib.rsfm_pipeline_run.argtypes = [c_vp, c_vp]
lib.rsfm_pipeline_info.restype = ctypes.c_int
lib.rsfm_pipeline_info.argtypes = [c_vp, ctypes.POINTER(PipelineInfo), c_vp, ctypes.c_size_t,
                                   ctypes.POINTER(ctypes.c_size_t)]
lib.rsfm_pipeline_destroy.restype = ctypes.c_int
lib.rsfm_pipeline_destroy.argtypes = [c_vp]


class Pipeline:
    """pairs: list of (descA uint8 (nA,128), descB uint8 (nB,128), kpA float32 (nA,2), kpB float32 (nB,2)).

    After run(), self.out[p] holds the device results of pair p: bestA, secA,
    bestB (int32 (n,2) BestMatch bit patterns), matches (nA,3), count (1,),
    pts (nA,4), idx (H,8), counts (H,), best (2,), mask (nA,).

    refit / confidence (8-point solver): the end of the reference's driver as a graph
    stage (rsfm_ransac_f8_finish): F (3,3) and result (4,) = [best hypothesis, its
    inliers, hypotheses walked, final inliers] are added and mask holds the final F's
    inliers.

    solver="degensac" runs the DEGENSAC pair solver instead of 8-point F RANSAC:
    idx is then (H,7), mask holds F's inliers, and F (3,3), H (3,3), result (4,),
    maskH (nA,) are added; with cams (a list of (cam0, cam1) as (fx, fy, cx, cy)
    per pair) also posesF (4,12), poseF (10,), posesH (8,12), normalsH (8,3),
    poseH (19,).

    match_filter: optional dict with img_size (a (width, height) of the first image per
    pair) and cell_cols / n_tests / threshold / min_votes / try_other (defaults 16 / 256 /
    0.15 / 2 / True, the reference's): the affine match filter stage after the ratio and
    cross-check filter; matches, pts and count then hold the filtered matches.

    distortion: optional list (one entry per pair, None for none) of 10 floats,
    k1, k2, p1, p2, k3 of camera 0 then camera 1; those pairs' gathered keypoint
    pairs are undistorted (with cams, required then) before sampling.
    """

    def __init__(self, pairs, n_hyp=4096, threshold=1.0, mode=1, ratio=(4, 5), seed=77, min_sep_sq=(0.0, 0.0),
                 seg_cols=0, wg_per_xcd=0, solver="f8", threshold_h=3.0, n_candidates=8, n_plane_samples=256,
                 front_ratio=0.95, cams=None, distortion=None, refit=False, confidence=0.0, max_iter=5000,
                 match_filter=None):
        dev = pairs[0][0].device
        dg = solver == "degensac"
        fin = not dg and (refit or confidence > 0)
        mf = dict(cell_cols=16, n_tests=256, threshold=0.15, min_votes=2, try_other=True)
        img_size = None
        if match_filter is not None:
            mf.update({k: v for k, v in match_filter.items() if k != "img_size"})
            img_size = match_filter["img_size"]
        k = 7 if dg else 8
        self.inputs = pairs
        self.out = []
        arr = (PipelinePair * len(pairs))()
        keep = []  # host distortion arrays, read by the creation call
        for i, (dA, dB, kA, kB) in enumerate(pairs):
            nA, nB = dA.shape[0], dB.shape[0]
            o = {
                "bestA": torch.empty((max(1, nA), 2), dtype=torch.int32, device=dev),
                "secA": torch.empty((max(1, nA), 2), dtype=torch.int32, device=dev),
                "bestB": torch.empty((max(1, nB), 2), dtype=torch.int32, device=dev),
                "matches": torch.empty((max(1, nA), 3), dtype=torch.int32, device=dev),
                "count": torch.zeros(1, dtype=torch.int32, device=dev),
                "pts": torch.empty((max(1, nA), 4), dtype=torch.float32, device=dev),
                "idx": torch.empty((max(1, n_hyp), k), dtype=torch.int32, device=dev),
                "counts": torch.empty(max(1, n_hyp), dtype=torch.int32, device=dev),
                "best": torch.empty(2, dtype=torch.int32, device=dev),
                "mask": torch.empty(max(1, nA), dtype=torch.uint8, device=dev),
            }
            if fin:
                o.update(F=torch.zeros((3, 3), dtype=torch.float32, device=dev),
                         result=torch.zeros(4, dtype=torch.int32, device=dev))
            if dg:
                o.update(F=torch.zeros((3, 3), dtype=torch.float32, device=dev),
                         H=torch.zeros((3, 3), dtype=torch.float32, device=dev),
                         result=torch.zeros(4, dtype=torch.int32, device=dev),
                         maskH=torch.empty(max(1, nA), dtype=torch.uint8, device=dev))
                if cams is not None:
                    o.update(posesF=torch.zeros((4, 12), dtype=torch.float32, device=dev),
                             poseF=torch.zeros(10, dtype=torch.int32, device=dev),
                             posesH=torch.zeros((8, 12), dtype=torch.float32, device=dev),
                             normalsH=torch.zeros((8, 3), dtype=torch.float32, device=dev),
                             poseH=torch.zeros(19, dtype=torch.int32, device=dev))
            p = lambda t: t.data_ptr()  # noqa: E731
            q = lambda key: o[key].data_ptr() if key in o else None  # noqa: E731
            c0, c1 = cams[i] if cams is not None else ((1.0, 1.0, 0.0, 0.0), (1.0, 1.0, 0.0, 0.0))
            dist = None
            if distortion is not None and distortion[i] is not None:
                if cams is None:
                    raise ValueError("distortion needs cams")
                dist = (ctypes.c_float * 10)(*[float(v) for v in distortion[i]])
                keep.append(dist)
            arr[i] = PipelinePair(nA, nB, p(dA), p(dB), p(kA), p(kB), p(o["bestA"]), p(o["secA"]), p(o["bestB"]),
                                  p(o["matches"]), p(o["count"]), p(o["pts"]), p(o["idx"]), p(o["counts"]),
                                  p(o["best"]), p(o["mask"]), q("F"), q("H"), q("result"), q("maskH"),
                                  Camera(*[float(v) for v in c0]), Camera(*[float(v) for v in c1]), q("posesF"),
                                  q("poseF"), q("posesH"), q("normalsH"), q("poseH"),
                                  ctypes.cast(dist, c_vp) if dist is not None else None,
                                  *(img_size[i] if img_size is not None else (0, 0)))
            self.out.append(o)
        cfg = PipelineConfig(n_hyp, threshold, mode, ratio[0], ratio[1], seed, min_sep_sq[0], min_sep_sq[1], seg_cols,
                             wg_per_xcd, SOLVERS[solver], threshold_h, n_candidates, n_plane_samples, front_ratio,
                             confidence, max_iter, int(bool(refit)), int(match_filter is not None), mf["cell_cols"],
                             mf["n_tests"], mf["threshold"], mf["min_votes"], int(mf["try_other"]))
        torch.cuda.synchronize()  # the inputs must be complete: creation times candidate tilings on them
        self.handle = c_vp()
        check(lib.rsfm_pipeline_create(ctypes.cast(arr, c_vp), len(pairs), ctypes.byref(cfg), ctypes.byref(self.handle)),
              "rsfm_pipeline_create")
        self.n_hyp = n_hyp

    def run(self, stream=None):
        s = stream or torch.cuda.current_stream()
        check(lib.rsfm_pipeline_run(self.handle, c_vp(s.cuda_stream)), "rsfm_pipeline_run")

    def info(self):
        inf = PipelineInfo()
        n = ctypes.c_size_t(0)
        ents = (TuningEntry * 64)()
        check(lib.rsfm_pipeline_info(self.handle, ctypes.byref(inf), ctypes.cast(ents, c_vp), 64, ctypes.byref(n)),
              "rsfm_pipeline_info")
        return {"segCols": inf.segCols, "wgPerXcd": inf.wgPerXcd, "tiles": inf.nbTiles,
                "tuned_match_ms": inf.tunedMatchMs,
                "candidates": [(ents[i].segCols, ents[i].wgPerXcd, round(ents[i].matchMs, 4)) for i in range(n.value)]}

    def close(self):
        if self.handle:
            check(lib.rsfm_pipeline_destroy(self.handle), "rsfm_pipeline_destroy")
            self.handle = c_vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
