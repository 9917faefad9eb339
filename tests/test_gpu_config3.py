"""The batched pipeline at its full workload: 256 image pairs with keypoint counts
uniform in [1024, 16384] per image and geometric outlier rates uniform in
[30%, 80%] among the true descriptor matches, 4096 8-point hypotheses per pair
with the refit on the inliers, one graph launch (the bench's batched workload).

- every pair: identical to the individual entry points (top-2 match, filter and
  gather, seeded samples, scorer, finisher);
- five pairs (the largest, nA x nB at 16K, among them): identical to the CPU
  oracle bit for bit (matches, distances, kept pairs, counts, F, mask);
- every pair: properties (kept matches pass the ratio test and cross-check,
  best count = max of the counts, mask population = final count).
"""
import numpy as np
import pytest
import torch

from test_gpu_pipeline import _matched, _sample

pytestmark = pytest.mark.gpu

P, H = 256, 4096


def _say(*a):
    print("[case]", *a, flush=True)


def _config3(dev):
    from rapidsfm_amd import synth

    rng = np.random.default_rng(2024)
    pairs = []
    for i in range(P):
        nA, nB = (int(v) for v in rng.integers(1024, 16385, 2))
        rate = float(rng.uniform(0.3, 0.8))
        pr = synth.make_pair(5000 + i, nA, nB, copy_frac=0.5, geo_frac=1.0 - rate, device=dev)
        kA = torch.from_numpy(np.ascontiguousarray(pr["kpA"][:, :2])).to(dev)
        kB = torch.from_numpy(np.ascontiguousarray(pr["kpB"][:, :2])).to(dev)
        pairs.append((pr["descA"], pr["descB"], kA, kB))
    return pairs


def test_config3_full_batch(orc, dev):
    from rapidsfm_amd import ops
    from rapidsfm_amd.pipeline import Pipeline

    _say("config 3: 256 pairs U[1K,16K], outliers U[30%,80%], 4096 hypotheses, refit")
    pairs = _config3(dev)
    pl = Pipeline(pairs, n_hyp=H, threshold=1.0, mode=1, seed=77, refit=True)
    pl.run()
    _say("entry points")
    outs = _matched(pairs)
    _sample(pairs, outs, H, 77, 8)
    sc = ops.ransac_f8([(o["pts"], o["idx"], o["count"]) for o in outs], 1.0, mode=1, want_models=True)
    fin = ops.ransac_f8_finish([(o["pts"], o["idx"], s["counts"], s["models"], o["count"]) for o, s in zip(outs, sc)],
                               1.0, 1, refit=True)
    torch.cuda.synchronize()
    for i, (o, r, s, f) in enumerate(zip(pl.out, outs, sc, fin)):
        nA, nB = pairs[i][0].shape[0], pairs[i][1].shape[0]
        for k, n in (("bestA", nA), ("secA", nA), ("bestB", nB)):
            assert torch.equal(o[k][:n], r[k][:n]), (i, k)
        c = int(r["count"].item())
        assert int(o["count"].item()) == c, i
        assert torch.equal(o["matches"][:c], r["matches"][:c]) and torch.equal(o["pts"][:c], r["pts"][:c]), i
        assert torch.equal(o["idx"], r["idx"]) and torch.equal(o["counts"], s["counts"]), i
        assert torch.equal(o["F"].view(torch.int32), f["F"].view(torch.int32)), i
        assert torch.equal(o["result"], f["result"]) and torch.equal(o["mask"][:nA], f["mask"][:nA]), i

    _say("properties on every pair")
    u32 = lambda t: t.cpu().numpy().view(np.uint32)  # noqa: E731
    kept = 0
    for i, o in enumerate(pl.out):
        nA = pairs[i][0].shape[0]
        c = int(o["count"].item())
        m = u32(o["matches"][:c])
        b1, b2, bb = u32(o["bestA"]), u32(o["secA"]), u32(o["bestB"])
        q, t = m[:, 0], m[:, 1]
        assert np.all(np.diff(q.astype(np.int64)) > 0)  # query order
        assert np.array_equal(b1[q, 0], t) and np.array_equal(bb[t, 0], q)  # cross-check
        d1, d2 = b1[q, 1].astype(np.uint64), b2[q, 1].astype(np.uint64)
        assert np.all(25 * d1 * d1 < 16 * d2 * d2)  # ratio 0.8 on L2
        counts = u32(o["counts"])
        res = u32(o["result"])
        if c >= 8:
            assert res[1] == counts.max() and counts[res[0]] == res[1]
            assert res[3] >= res[1]
        assert int(o["mask"][:nA].sum().item()) == res[3]
        kept += c

    # five pairs against the oracle, the largest one among them
    big = int(np.argmax([a.shape[0] * b.shape[0] for a, b, _, _ in pairs]))
    for i in sorted({big, 0, 1, 100, 255}):
        dA, dB, kA, kB = (x.cpu().numpy() for x in pairs[i])
        _say("oracle pair", i, dA.shape[0], dB.shape[0])
        ra, sa, rb = orc.bf_match_top2_fast(dA, dB)
        o = pl.out[i]
        assert np.array_equal(u32(o["bestA"])[: dA.shape[0]], ra) and np.array_equal(u32(o["secA"])[: dA.shape[0]], sa)
        assert np.array_equal(u32(o["bestB"])[: dB.shape[0]], rb)
        m = orc.filter_matches(ra, sa, rb, (4, 5))
        c = int(o["count"].item())
        assert c == len(m)
        pts = np.concatenate([kA[m[:, 0]], kB[m[:, 1]]], 1).astype(np.float32)
        assert np.array_equal(o["pts"][:c].cpu().numpy(), pts)
        idx = orc.sample_indices_hash(77 + i, c, H)
        assert np.array_equal(u32(o["idx"]), idx)
        counts, models = orc.f8_score(pts, idx, 1.0, 1, nthreads=16)
        assert np.array_equal(u32(o["counts"]), counts)
        F, res, mask = orc.f8_finish(pts, idx, counts, models, 1.0, 1, 0.0, 5000, True)
        assert np.array_equal(o["F"].cpu().numpy().view(np.uint32), F.view(np.uint32))
        assert np.array_equal(u32(o["result"]), res) and np.array_equal(o["mask"][:c].cpu().numpy(), mask)
    assert max(pairs[big][0].shape[0], pairs[big][1].shape[0]) >= 15000
    print("kept matches over the batch:", kept, flush=True)
    pl.close()
