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


def _reference(pairs, H, seed):
    """The same pass through the individual entry points."""
    from rapidsfm_amd import ops
    from rapidsfm_amd._abi import FilterTask, SampleTask, check, lib

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
    assert info["segCols"] in (1024, 2048, 4096, 8192) and len(info["candidates"]) == 12
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


@pytest.mark.parametrize("seg,wg", [(1024, 32), (8192, 64), (2048, 40), (4096, 0xFFFFFFFF)])
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
