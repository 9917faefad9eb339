"""Batched two-view pipeline over the C ABI (rsfm_pipeline_*): many image pairs of
uneven size matched and verified by one HIP graph launch. Torch tensors provide
the device buffers; the library owns only its task tables and partials."""
import ctypes

import torch

from ._abi import check, lib

c_u32, c_f32, c_vp = ctypes.c_uint32, ctypes.c_float, ctypes.c_void_p


# wg_per_xcd value: one matcher workgroup per tile instead of a persistent grid
WG_PER_TILE = 0xFFFFFFFF


class PipelinePair(ctypes.Structure):
    _fields_ = [("nA", c_u32), ("nB", c_u32), ("descA", c_vp), ("descB", c_vp), ("kpA", c_vp), ("kpB", c_vp),
                ("bestA", c_vp), ("secondA", c_vp), ("bestB", c_vp), ("matches", c_vp), ("count", c_vp),
                ("pts", c_vp), ("sampleIdx", c_vp), ("counts", c_vp), ("best", c_vp), ("inlierMask", c_vp)]


class PipelineConfig(ctypes.Structure):
    _fields_ = [("nbHyp", c_u32), ("threshold", c_f32), ("mode", ctypes.c_int), ("ratioNum", c_u32),
                ("ratioDen", c_u32), ("seed", ctypes.c_uint64), ("minSepSq0", c_f32), ("minSepSq1", c_f32),
                ("segCols", c_u32), ("wgPerXcd", c_u32)]


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
    """

    def __init__(self, pairs, n_hyp=4096, threshold=1.0, mode=1, ratio=(4, 5), seed=77, min_sep_sq=(0.0, 0.0),
                 seg_cols=0, wg_per_xcd=0):
        dev = pairs[0][0].device
        self.inputs = pairs
        self.out = []
        arr = (PipelinePair * len(pairs))()
        for i, (dA, dB, kA, kB) in enumerate(pairs):
            nA, nB = dA.shape[0], dB.shape[0]
            o = {
                "bestA": torch.empty((max(1, nA), 2), dtype=torch.int32, device=dev),
                "secA": torch.empty((max(1, nA), 2), dtype=torch.int32, device=dev),
                "bestB": torch.empty((max(1, nB), 2), dtype=torch.int32, device=dev),
                "matches": torch.empty((max(1, nA), 3), dtype=torch.int32, device=dev),
                "count": torch.zeros(1, dtype=torch.int32, device=dev),
                "pts": torch.empty((max(1, nA), 4), dtype=torch.float32, device=dev),
                "idx": torch.empty((max(1, n_hyp), 8), dtype=torch.int32, device=dev),
                "counts": torch.empty(max(1, n_hyp), dtype=torch.int32, device=dev),
                "best": torch.empty(2, dtype=torch.int32, device=dev),
                "mask": torch.empty(max(1, nA), dtype=torch.uint8, device=dev),
            }
            p = lambda t: t.data_ptr()  # noqa: E731
            arr[i] = PipelinePair(nA, nB, p(dA), p(dB), p(kA), p(kB), p(o["bestA"]), p(o["secA"]), p(o["bestB"]),
                                  p(o["matches"]), p(o["count"]), p(o["pts"]), p(o["idx"]), p(o["counts"]),
                                  p(o["best"]), p(o["mask"]))
            self.out.append(o)
        cfg = PipelineConfig(n_hyp, threshold, mode, ratio[0], ratio[1], seed, min_sep_sq[0], min_sep_sq[1], seg_cols,
                             wg_per_xcd)
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
