"""Batched HIP-graph pipeline: bit-identical to the individual entry points on a batch of
uneven pairs (including multi-segment and empty ones), stable across graph replays,
and independent of the tiling the autotuner picks."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _batch(dev, sizes, seed=0):
    from rapidsfm_amd import synth

    pairs = []
    for i, (nA, nB) in enumerate(sizes):
        if nA == 0 or nB == 0:
            dA = torch.zeros((nA, 128), dtype=torch.uint8, device=dev)
            dB = torch.zeros((nB, 128), dtype=torch.uint8, device=dev)
            kA = torch.zeros((nA, 2), dtype=torch.float32, device=dev)
            kB = torch.zeros((nB, 2), dtype=torch.float32, device=dev)
        else:
            pr = synth.make_pair(seed + i, nA, nB, 0.4, 0.6, device=dev)
            dA, dB = pr["descA"], pr["descB"]
            kA = torch.from_numpy(np.ascontiguousarray(pr["kpA"][:, :2])).to(dev)
            kB = torch.from_numpy(np.ascontiguousarray(pr["kpB"][:, :2])).to(dev)
        pairs.append((dA, dB, kA, kB))
    return pairs


def _matched(pairs):
    """match + ratio test + cross-check + keypoint gather through the individual entry points"""
    from rapidsfm_amd import ops
    from rapidsfm_amd._abi import FilterTask, check, lib

    dev = pairs[0][0].device
    res = ops.bf_match_top2([(a, b) for a, b, _, _ in pairs])
    outs = []
    filt = (FilterTask * len(pairs))()
    for i, ((dA, dB, kA, kB), (ba, sa, bb)) in enumerate(zip(pairs, res)):
        nA = dA.shape[0]
        o = {"bestA": ba, "secA": sa, "bestB": bb,
             "matches": torch.empty((max(1, nA), 3), dtype=torch.int32, device=dev),
             "count": torch.zeros(1, dtype=torch.int32, device=dev),
             "pts": torch.empty((max(1, nA), 4), dtype=torch.float32, device=dev)}
        filt[i] = FilterTask(ba.data_ptr() if nA else None, sa.data_ptr() if nA else None, nA,
                             bb.data_ptr() if dB.shape[0] else None, dB.shape[0], o["matches"].data_ptr(),
                             o["count"].data_ptr(), kA.data_ptr() if nA else 16, kB.data_ptr() if dB.shape[0] else 16,
                             o["pts"].data_ptr())
        outs.append(o)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(lib.rsfm_filter_matches_batch(ctypes.cast(filt, ctypes.c_void_p), len(pairs), 4, 5, s), "filter")
    return outs


def _sample(pairs, outs, H, seed, k):
    from rapidsfm_amd._abi import SampleTask, check, lib

    dev = pairs[0][0].device
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    samp = (SampleTask * len(pairs))()
    for i, o in enumerate(outs):
        o["idx"] = torch.empty((H, k), dtype=torch.int32, device=dev)
        samp[i] = SampleTask(o["idx"].data_ptr(), H, max(1, pairs[i][0].shape[0]), o["count"].data_ptr(), seed + i)
    check(lib.rsfm_ransac_sample_indices_batch_k(ctypes.cast(samp, ctypes.c_void_p), len(pairs), k, s), "sample")


def _reference(pairs, H, seed):
    """The same pass through the individual entry points."""
    from rapidsfm_amd import ops
    from rapidsfm_amd._abi import SampleTask, check, lib

    dev = pairs[0][0].device
    outs = _matched(pairs)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    samp = (SampleTask * len(pairs))()
    for i, o in enumerate(outs):
        o["idx"] = torch.empty((H, 8), dtype=torch.int32, device=dev)
        samp[i] = SampleTask(o["idx"].data_ptr(), H, max(1, pairs[i][0].shape[0]), o["count"].data_ptr(), seed + i)
    check(lib.rsfm_ransac_sample_indices_batch(ctypes.cast(samp, ctypes.c_void_p), len(pairs), s), "sample")
    r = ops.ransac_f8([(o["pts"], o["idx"], o["count"]) for o in outs], 1.0, mode=1)
    for o, x in zip(outs, r):
        o.update(counts=x["counts"], best=x["best"], mask=x["mask"])
    return outs


def _compare(pl, ref, pairs):
    for i, (o, r) in enumerate(zip(pl.out, ref)):
        nA, nB = pairs[i][0].shape[0], pairs[i][1].shape[0]
        for k, n in (("bestA", nA), ("secA", nA), ("bestB", nB)):
            assert torch.equal(o[k][:n], r[k][:n]), (i, k)
        c = int(r["count"].item())
        assert int(o["count"].item()) == c, i
        assert torch.equal(o["matches"][:c], r["matches"][:c]), i
        assert torch.equal(o["pts"][:c], r["pts"][:c]), i
        assert torch.equal(o["idx"], r["idx"]), i
        assert torch.equal(o["counts"], r["counts"]), i
        assert torch.equal(o["best"], r["best"]), i
        assert torch.equal(o["mask"][:nA], r["mask"][:nA]), i


SIZES = [(1000, 1200), (300, 9000), (8800, 700), (2048, 2048), (0, 500), (400, 0), (4100, 16384), (77, 64),
         (5000, 4000), (1500, 8193), (256, 257), (3000, 100), (1200, 1100), (12000, 3000)]


def test_pipeline_matches_entry_points(dev):
    from rapidsfm_amd.pipeline import Pipeline

    _say("pipeline vs entry points, autotuned", len(SIZES))
    pairs = _batch(dev, SIZES, seed=40)
    H = 512
    pl = Pipeline(pairs, n_hyp=H, seed=500)
    info = pl.info()
    print(info, flush=True)
    assert info["segCols"] in (1024, 2048, 4096) and len(info["candidates"]) == 9
    pl.run()
    ref = _reference(pairs, H, 500)
    torch.cuda.synchronize()
    _compare(pl, ref, pairs)
    # replay after clobbering every output: the graph restores all of them
    for o in pl.out:
        for t in o.values():
            t.fill_(-7)
    pl.run()
    pl.run()
    torch.cuda.synchronize()
    _compare(pl, ref, pairs)
    pl.close()


@pytest.mark.parametrize("seg,wg", [(1024, 32), (4096, 64), (2048, 40), (4096, 0xFFFFFFFF)])
def test_pipeline_fixed_tilings_agree(dev, seg, wg):
    from rapidsfm_amd.pipeline import Pipeline

    _say("pipeline fixed tiling", seg, wg)
    pairs = _batch(dev, SIZES[:8], seed=90)
    pl = Pipeline(pairs, n_hyp=256, seed=3, seg_cols=seg, wg_per_xcd=wg)
    assert pl.info()["candidates"] == []
    pl.run()
    ref = _reference(pairs, 256, 3)
    torch.cuda.synchronize()
    _compare(pl, ref, pairs)
    pl.close()


def test_pipeline_pair_against_oracle(orc, dev):
    from rapidsfm_amd.pipeline import Pipeline

    _say("pipeline pair vs oracle")
    pairs = _batch(dev, [(1800, 2500), (900, 700)], seed=7)
    pl = Pipeline(pairs, n_hyp=128, seed=11)
    pl.run()
    torch.cuda.synchronize()
    dA, dB, kA, kB = (t.cpu().numpy() for t in pairs[0])
    ra, sa, rb = orc.bf_match_top2_u8(dA, dB)
    u32 = lambda t: t.cpu().numpy().view(np.uint32)  # noqa: E731
    o = pl.out[0]
    assert np.array_equal(u32(o["bestA"]), ra) and np.array_equal(u32(o["secA"]), sa)
    assert np.array_equal(u32(o["bestB"]), rb)
    m = orc.filter_matches(ra, sa, rb, (4, 5))
    c = int(o["count"].item())
    assert c == len(m)
    pts = np.concatenate([kA[m[:, 0]], kB[m[:, 1]]], 1).astype(np.float32)
    assert np.array_equal(o["pts"][:c].cpu().numpy(), pts)
    idx = orc.sample_indices_hash(11, c, 128)
    counts, _ = orc.f8_score(pts, idx, 1.0, 1, with_models=False)
    assert np.array_equal(u32(o["counts"]), counts)
    pl.close()


CAM = (2000.0, 2000.0, 2000.0, 1500.0)


def _reference_degensac(pairs, H, seed, thr, thr_h, mode, ncand, nplane, front):
    """The DEGENSAC solver chain through the individual entry points."""
    from rapidsfm_amd import ops

    outs = _matched(pairs)
    _sample(pairs, outs, H, seed, 7)
    r7 = ops.ransac_f7([(o["pts"], o["idx"], o["count"]) for o in outs], thr, mode=mode)
    dg = ops.degensac_refine([(o["pts"], o["idx"], r["counts"], o["count"]) for o, r in zip(outs, r7)], thr, thr_h,
                             mode=mode, nb_candidates=ncand, nb_plane_samples=nplane, seed=seed)
    pf = ops.pose_from_f([(d["F"], CAM, CAM, o["pts"], d["mask_f"], o["count"]) for o, d in zip(outs, dg)], front)
    ph = ops.pose_from_h([(d["H"], CAM, CAM, o["pts"], d["mask_h"], o["count"]) for o, d in zip(outs, dg)], front)
    for o, r, d, a, b in zip(outs, r7, dg, pf, ph):
        o.update(counts=r["counts"], best=r["best"], mask=d["mask_f"], maskH=d["mask_h"], F=d["F"], H=d["H"],
                 result=d["result"], posesF=a["poses"], poseF=a["result"], posesH=b["poses"],
                 normalsH=b["normals"], poseH=b["result"])
    return outs


def _compare_degensac(pl, ref, pairs):
    for i, (o, r) in enumerate(zip(pl.out, ref)):
        nA = pairs[i][0].shape[0]
        c = int(r["count"].item())
        assert int(o["count"].item()) == c, i
        assert torch.equal(o["pts"][:c], r["pts"][:c]), i
        for k in ("idx", "counts", "best", "F", "H", "result", "posesF", "poseF", "posesH", "normalsH", "poseH"):
            assert torch.equal(o[k].view(torch.int32), r[k].view(torch.int32)), (i, k)
        assert torch.equal(o["mask"][:nA], r["mask"][:nA]), i
        assert torch.equal(o["maskH"][:nA], r["maskH"][:nA]), i


@pytest.mark.parametrize("mode", [0, 1])
def test_pipeline_degensac_matches_entry_points(dev, mode):
    from rapidsfm_amd.pipeline import Pipeline

    rng = np.random.default_rng(5 + mode)
    sizes = SIZES[:8] + [(int(a), int(b)) for a, b in rng.integers(150, 3000, size=(28, 2))]
    _say("pipeline DEGENSAC vs entry points", len(sizes), "mode", mode)
    pairs = _batch(dev, sizes, seed=300 + 50 * mode)
    H = 384
    args = dict(thr=1.5, thr_h=3.0, mode=mode, ncand=6, nplane=128, front=0.95)
    pl = Pipeline(pairs, n_hyp=H, seed=21, threshold=args["thr"], mode=mode, solver="degensac",
                  threshold_h=args["thr_h"], n_candidates=args["ncand"], n_plane_samples=args["nplane"],
                  front_ratio=args["front"], cams=[(CAM, CAM)] * len(pairs), seg_cols=2048, wg_per_xcd=32)
    pl.run()
    ref = _reference_degensac(pairs, H, 21, **args)
    torch.cuda.synchronize()
    _compare_degensac(pl, ref, pairs)
    # geometry: on pairs with many matches the F pose keeps a candidate
    good = [o for o in pl.out if int(o["result"][0].item()) >= 100]
    assert good and sum(int(o["poseF"][8].item()) >= 1 for o in good) >= 0.8 * len(good)
    for o in pl.out:
        for t in o.values():
            t.fill_(-7)
    pl.run()
    torch.cuda.synchronize()
    _compare_degensac(pl, ref, pairs)
    pl.close()


def test_pipeline_degensac_sampler_matches_oracle(orc, dev):
    from rapidsfm_amd.pipeline import Pipeline

    _say("pipeline DEGENSAC samples vs oracle")
    pairs = _batch(dev, [(1500, 1700)], seed=8)
    pl = Pipeline(pairs, n_hyp=200, seed=4, solver="degensac", seg_cols=2048, wg_per_xcd=32)
    pl.run()
    torch.cuda.synchronize()
    o = pl.out[0]
    c = int(o["count"].item())
    idx = orc.sample_indices_hash_k(4, c, 7, 200)
    assert np.array_equal(o["idx"].cpu().numpy().view(np.uint32), idx)
    pts = o["pts"][:c].cpu().numpy()
    counts, _ = orc.f7_score(pts, idx, 1.0, 1, nthreads=8)
    assert np.array_equal(o["counts"].cpu().numpy().view(np.uint32), counts)
    pl.close()


def test_pipeline_solver_arguments(dev):
    from rapidsfm_amd.pipeline import Pipeline

    _say("pipeline solver arguments")
    pairs = _batch(dev, [(300, 300)], seed=9)
    with pytest.raises(RuntimeError):  # pose stage without the DEGENSAC solver
        from rapidsfm_amd import pipeline as P

        arr = (P.PipelinePair * 1)()
        cfg = P.PipelineConfig(64, 1.0, 1, 4, 5, 1, 0.0, 0.0, 2048, 32, 0, 3.0, 8, 256, 0.95)
        o = [torch.empty(max(1, 300 * 4), dtype=torch.int32, device=dev) for _ in range(12)]
        a, b, ka, kb = pairs[0]
        p = [t.data_ptr() for t in o]
        arr[0] = P.PipelinePair(300, 300, a.data_ptr(), b.data_ptr(), ka.data_ptr(), kb.data_ptr(), *p[:10],
                                None, None, None, None, P.Camera(1, 1, 0, 0), P.Camera(1, 1, 0, 0), p[10], p[11],
                                None, None, None)
        h = ctypes.c_void_p()
        P.check(P.lib.rsfm_pipeline_create(ctypes.cast(arr, ctypes.c_void_p), 1, ctypes.byref(cfg), ctypes.byref(h)),
                "create")
    with pytest.raises(KeyError):
        Pipeline(pairs, n_hyp=64, solver="lo-ransac")


def test_pipeline_undistorts_pairs(orc, dev):
    from rapidsfm_amd import ops
    from rapidsfm_amd.pipeline import Pipeline

    _say("pipeline with lens distortion")
    sizes = [(1200, 1300), (800, 900), (2000, 1500), (300, 0)]
    pairs = _batch(dev, sizes, seed=61)
    dist = [(-0.05, 0.01, 0.0, 0.0005, 0.0, 0.03, 0.0, 0.0002, 0.0, 0.001), None,
            (0.0,) * 5 + (-0.02, 0.0, 0.0, 0.0, 0.0), (-0.05,) + (0.0,) * 9]
    H = 256
    pl = Pipeline(pairs, n_hyp=H, seed=5, cams=[(CAM, CAM)] * len(pairs), distortion=dist, seg_cols=2048,
                  wg_per_xcd=32)
    pl.run()
    outs = _matched(pairs)
    probs = []
    for o, d in zip(outs, dist):
        if d is not None:
            probs.append((o["pts"], CAM, CAM, d[:5], d[5:], o["count"]))
    ops.undistort_pt_pairs(probs, 2)
    _sample(pairs, outs, H, 5, 8)
    r = ops.ransac_f8([(o["pts"], o["idx"], o["count"]) for o in outs], 1.0, mode=1)
    torch.cuda.synchronize()
    for i, (o, ref, x) in enumerate(zip(pl.out, outs, r)):
        c = int(ref["count"].item())
        assert int(o["count"].item()) == c, i
        assert torch.equal(o["pts"][:c], ref["pts"][:c]), i
        assert torch.equal(o["counts"], x["counts"]) and torch.equal(o["best"], x["best"]), i
    # pair 0 against the oracle: the gathered raw keypoints, undistorted on the CPU
    dA, dB, kA, kB = (t.cpu().numpy() for t in pairs[0])
    ra, sa, rb = orc.bf_match_top2_u8(dA, dB)
    m = orc.filter_matches(ra, sa, rb, (4, 5))
    raw = np.concatenate([kA[m[:, 0]], kB[m[:, 1]]], 1).astype(np.float32)
    cam = np.array(CAM, np.float32)
    e0 = orc.undistort(raw[:, 0], raw[:, 1], cam, np.array(dist[0][:5], np.float32), 2)
    e1 = orc.undistort(raw[:, 2], raw[:, 3], cam, np.array(dist[0][5:], np.float32), 2)
    want = np.column_stack([e0[0], e0[1], e1[0], e1[1]]).astype(np.float32)
    got = pl.out[0]["pts"][:len(m)].cpu().numpy()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.abs(got - raw).max() > 1.0
    pl.close()


@pytest.mark.parametrize("conf,refit", [(0.0, True), (0.99, True), (0.995, False)])
def test_pipeline_finish_stage(orc, dev, conf, refit):
    """The 8-point graph with the end of the reference's driver as a stage:
    identical to rsfm_ransac_f8 + rsfm_ransac_f8_finish over the same matches,
    and pair 0 identical to the oracle's restatement."""
    from rapidsfm_amd import ops
    from rapidsfm_amd.pipeline import Pipeline

    _say("pipeline finish stage conf", conf, "refit", refit)
    sizes = SIZES[:4] + [(2500, 2600), (6000, 5000)]
    pairs = _batch(dev, sizes, seed=700)
    H = 1024
    pl = Pipeline(pairs, n_hyp=H, seed=9, seg_cols=2048, wg_per_xcd=32, refit=refit, confidence=conf, max_iter=600)
    pl.run()
    outs = _matched(pairs)
    _sample(pairs, outs, H, 9, 8)
    sc = ops.ransac_f8([(o["pts"], o["idx"], o["count"]) for o in outs], 1.0, mode=1, want_models=True)
    fin = ops.ransac_f8_finish([(o["pts"], o["idx"], s["counts"], s["models"], o["count"]) for o, s in zip(outs, sc)],
                               1.0, 1, confidence=conf, max_iter=600, refit=refit)
    torch.cuda.synchronize()
    for i, (o, r, s, f) in enumerate(zip(pl.out, outs, sc, fin)):
        nA = pairs[i][0].shape[0]
        c = int(r["count"].item())
        assert int(o["count"].item()) == c and torch.equal(o["pts"][:c], r["pts"][:c]), i
        assert torch.equal(o["counts"], s["counts"]), i
        assert torch.equal(o["F"].view(torch.int32), f["F"].view(torch.int32)), i
        assert torch.equal(o["result"], f["result"]), i
        assert torch.equal(o["mask"][:nA], f["mask"][:nA]), i
        res = o["result"].cpu().numpy().view(np.uint32)
        assert o["best"].cpu().numpy().view(np.uint32)[0] == res[0], i
    # pair 0 against the oracle
    o = pl.out[0]
    c = int(o["count"].item())
    pts = o["pts"][:c].cpu().numpy()
    idx = orc.sample_indices_hash(9, c, H)
    counts, models = orc.f8_score(pts, idx, 1.0, 1)
    F, res, mask = orc.f8_finish(pts, idx, counts, models, 1.0, 1, conf, 600, refit)
    assert np.array_equal(o["F"].cpu().numpy().view(np.uint32), F.view(np.uint32))
    assert np.array_equal(o["result"].cpu().numpy().view(np.uint32), res)
    assert np.array_equal(o["mask"][:c].cpu().numpy(), mask)
    pl.close()


def test_pipeline_match_filter_stage(orc, dev):
    """The affine match filter as a graph stage: identical to the entry-point
    chain (filter, rsfm_rmf_filter_batch with seed + pair, samples, scorer),
    and pair 0's kept pairs identical to the oracle's binning + filter."""
    from rapidsfm_amd import ops, rmf
    from rapidsfm_amd.pipeline import Pipeline

    _say("pipeline with the affine match filter stage")
    sizes = SIZES[:4] + [(2500, 2600), (6000, 5000), (30, 40)]
    pairs = _batch(dev, sizes, seed=900)
    H = 512
    img = [(4000, 3000)] * len(pairs)
    pl = Pipeline(pairs, n_hyp=H, seed=13, seg_cols=2048, wg_per_xcd=32, match_filter={"img_size": img})
    pl.run()
    outs = _matched(pairs)
    torch.cuda.synchronize()
    raw = [(int(o["count"].item()), o["pts"].clone()) for o in outs]
    rmf.filter_batch([(o["pts"], o["count"], img[i], 13 + i, o["matches"]) for i, o in enumerate(outs)])
    _sample(pairs, outs, H, 13, 8)
    sc = ops.ransac_f8([(o["pts"], o["idx"], o["count"]) for o in outs], 1.0, mode=1)
    torch.cuda.synchronize()
    for i, (o, r, s) in enumerate(zip(pl.out, outs, sc)):
        c = int(r["count"].item())
        assert int(o["count"].item()) == c, i
        assert torch.equal(o["matches"][:c], r["matches"][:c]) and torch.equal(o["pts"][:c], r["pts"][:c]), i
        assert torch.equal(o["idx"], r["idx"]) and torch.equal(o["counts"], s["counts"]), i
        assert torch.equal(o["best"], s["best"]), i
    n0, p0 = raw[0]
    m = p0[:n0].cpu().numpy()
    reordered, bounds, order, cols, rows, cw = orc.rmf_bin(4000, 3000, m, 16)
    votes, _ = orc.rmf_filter(reordered, bounds, cols, rows, 256, cw * 0.15, True, 13)
    keep = np.zeros(n0, bool)
    keep[order] = votes >= 2
    c0 = int(pl.out[0]["count"].item())
    assert c0 == int(keep.sum()) and np.array_equal(pl.out[0]["pts"][:c0].cpu().numpy(), m[keep])
    assert int(pl.out[-1]["count"].item()) == 0 or raw[-1][0] >= 16
    pl.close()
