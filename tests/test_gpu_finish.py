"""GPU parity of the end of the 8-point F-RANSAC driver (rsfm_ransac_f8_finish)
against its CPU restatement: adaptive cut, best sample, refit on the inliers,
final F (bit-exact FP32), result and mask (ImagePair.cpp:185-203,
ransac.hpp:139-160)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _d(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _u32(t):
    return t.cpu().numpy().view(np.uint32)


def _problem(orc, seed, n, n_out, H, t, mode, noise=0.5):
    pts, inl, _ = orc.gen_two_view(seed, n, n_out, noise_px=noise)
    idx = orc.sample_indices_hash(seed + 17, n, H)
    counts, models = orc.f8_score(pts, idx, t, mode)
    return pts, idx, counts, models


def _check(o, pts, ref, n_host=None):
    F, res, mask = ref
    n = pts.shape[0]
    assert np.array_equal(o["F"].cpu().numpy().view(np.uint32), F.view(np.uint32)), (o["F"].cpu().numpy(), F)
    assert np.array_equal(_u32(o["result"]), res), (_u32(o["result"]), res)
    gm = o["mask"].cpu().numpy()
    assert np.array_equal(gm[:n], mask)
    if n_host is not None:
        assert not gm[n:].any()


@pytest.mark.parametrize("mode,t", [(1, 1.0), (0, 24.0)])
@pytest.mark.parametrize("conf,max_iter", [(0.0, 5000), (0.99, 5000), (0.995, 300), (0.999, 20)])
def test_finish_bit_exact(rs, orc, dev, mode, t, conf, max_iter):
    _say("finish mode", mode, "conf", conf, "maxIter", max_iter)
    probs, refs = [], []
    for k, (n, nout, H) in enumerate([(600, 240, 256), (2048, 1200, 1024), (4096, 1638, 4096), (5000, 4000, 2048),
                                      (16384, 6000, 1500), (900, 100, 3000)]):
        pts, idx, counts, models = _problem(orc, 40 + k, n, nout, H, t, mode)
        use_models = k % 2 == 0
        probs.append((_d(pts, dev), _d(idx.view(np.int32), dev), _d(counts.view(np.int32), dev),
                      _d(models, dev) if use_models else None))
        refs.append((pts, orc.f8_finish(pts, idx, counts, models if use_models else None, t, mode, conf, max_iter)))
    outs = rs.ransac_f8_finish(probs, t, mode, confidence=conf, max_iter=max_iter, refit=True)
    kept = 0
    for o, (pts, ref) in zip(outs, refs):
        _check(o, pts, ref)
        kept += int(ref[1][3] > ref[1][1])
    assert kept >= 3  # the refit wins on most noisy pairs


def test_finish_without_refit_equals_best_model(rs, orc, dev):
    _say("finish, refit off")
    pts, idx, counts, models = _problem(orc, 5, 3000, 1200, 1024, 1.0, 1)
    o, = rs.ransac_f8_finish([(_d(pts, dev), _d(idx.view(np.int32), dev), _d(counts.view(np.int32), dev),
                               _d(models, dev))], 1.0, 1, refit=False)
    ref = orc.f8_finish(pts, idx, counts, models, 1.0, 1, refit=False)
    _check(o, pts, ref)
    b = int(np.argmax(counts))
    assert np.array_equal(o["F"].cpu().numpy(), models[b]) and _u32(o["result"])[3] == counts[b]


def test_finish_chain_from_scorer(rs, orc, dev):
    """Scorer and finisher on the device, nDev below the capacity: config-2 size
    (4096 correspondences, 4096 hypotheses)."""
    _say("scorer -> finish, nDev")
    pts, inl, _ = orc.gen_two_view(0, 4096, 1638)
    idx = orc.sample_indices(1, 3900, 4096)
    pd = _d(pts, dev)
    idd = _d(idx.view(np.int32), dev)
    ndev = torch.tensor([3900], dtype=torch.int32, device=dev)
    sc, = rs.ransac_f8([(pd, idd, ndev)], 1.0, 1, want_models=True)
    o, = rs.ransac_f8_finish([(pd, idd, sc["counts"], sc["models"], ndev)], 1.0, 1, confidence=0.995)
    counts, models = orc.f8_score(pts[:3900], idx, 1.0, 1)
    assert np.array_equal(_u32(sc["counts"]), counts)
    ref = orc.f8_finish(pts[:3900], idx, counts, models, 1.0, 1, 0.995, 5000)
    F, res, mask = ref
    assert np.array_equal(o["F"].cpu().numpy().view(np.uint32), F.view(np.uint32))
    assert np.array_equal(_u32(o["result"]), res)
    gm = o["mask"].cpu().numpy()
    assert np.array_equal(gm[:3900], mask) and not gm[3900:].any()
    assert res[3] >= res[1] and (mask.astype(bool) & inl[:3900]).sum() > 0.8 * inl[:3900].sum()


def test_finish_edge_cases(rs, orc, dev):
    _say("finish edge cases: n < 8, every count zero, few inliers")
    pts, idx, counts, models = _problem(orc, 9, 400, 100, 128, 1.0, 1)
    zeros = np.zeros(128, np.uint32)
    few = np.zeros(128, np.uint32)
    few[5] = 3
    cases = [
        (pts[:7], idx % 7, zeros),       # fewer than 8 points
        (pts, idx, zeros),               # fallback sample 0..7
        (pts, idx, few),                 # best count < 8 after its own recount: no refit
    ]
    probs = [(_d(p, dev), _d(i.astype(np.uint32).view(np.int32), dev), _d(c.view(np.int32), dev)) for p, i, c in cases]
    outs = rs.ransac_f8_finish(probs, 1.0, 1, confidence=0.99)
    for o, (p, i, c) in zip(outs, cases):
        _check(o, p, orc.f8_finish(p, i, c, None, 1.0, 1, 0.99, 5000))
    assert _u32(outs[0]["result"]).tolist() == [0xFFFFFFFF, 0, 0, 0]
    assert _u32(outs[1]["result"])[0] == 0xFFFFFFFF


def test_finish_many_tasks(rs, orc, dev):
    """40 problems in one call: one launch of 32 workgroups and one of 8."""
    _say("finish, 40 problems in one call")
    rng = np.random.default_rng(4)
    probs, refs = [], []
    for k in range(40):
        n = int(rng.integers(300, 1500))
        pts, idx, counts, models = _problem(orc, 200 + k, n, n // 3, int(rng.integers(64, 512)), 1.0, 1)
        probs.append((_d(pts, dev), _d(idx.view(np.int32), dev), _d(counts.view(np.int32), dev), _d(models, dev)))
        refs.append((pts, orc.f8_finish(pts, idx, counts, models, 1.0, 1, 0.99, 5000)))
    outs = rs.ransac_f8_finish(probs, 1.0, 1, confidence=0.99)
    for o, (pts, ref) in zip(outs, refs):
        _check(o, pts, ref)
