"""GPU parity of the batched F-RANSAC scorer against the CPU oracle.

Inlier counts, models, best hypothesis and inlier mask must be identical
(bit-exact FP32) for the same sample indices; model error is stated against
the FP64 oracle fit.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _d(a, dev, dt=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dt is not None:
        t = t.to(dt)
    return t.to(dev)


def _u32(t):
    return t.cpu().numpy().view(np.uint32)


def _norm(F):
    F = np.asarray(F, np.float64).reshape(3, 3)
    n = np.linalg.norm(F)
    if n == 0:
        return F
    F = F / n
    return F * np.sign(F.flat[np.argmax(np.abs(F))])


@pytest.mark.parametrize("mode,t,sep", [(1, 1.0, 0.0), (0, 24.0, 16384.0)])
def test_config2_counts_models_bit_exact(rs, orc, dev, mode, t, sep):
    _say("config2: 4096 pts, 40% outliers, 4096 hypotheses, mode", mode)
    pts, inl, Ft = orc.gen_two_view(0, 4096, 1638)
    idx = orc.sample_indices(1, 4096, 4096)
    out, = rs.ransac_f8([(_d(pts, dev), _d(idx.view(np.int32), dev))], t, mode, (sep, sep), want_models=True)
    rc, rm = orc.f8_score(pts, idx, t, mode, sep, sep)
    gc = _u32(out["counts"])
    gm = out["models"].cpu().numpy()
    assert np.array_equal(gc, rc), "inlier counts differ at %s" % np.nonzero(gc != rc)[0][:10]
    assert np.array_equal(gm.view(np.uint32), rm.view(np.uint32))
    best = _u32(out["best"])
    b = int(np.argmax(rc))  # first maximum
    assert best[0] == b and best[1] == rc[b]
    mask = out["mask"].cpu().numpy().astype(bool)
    assert np.array_equal(mask, orc.f_inlier_f32(rm[b], pts, t, mode))
    # the best model explains the synthetic geometry (an 8-point model from
    # noisy samples keeps ~85% of the true inliers at 1 px Sampson)
    tp = (mask & inl).sum()
    assert tp > 0.75 * inl.sum() and (mask & ~inl).sum() < 0.05 * (~inl).sum()


def test_model_error_vs_fp64(rs, orc, dev):
    _say("FP32 models vs FP64 oracle fit")
    pts, inl, _ = orc.gen_two_view(3, 2048, 600)
    idx = orc.sample_indices(5, 2048, 512)
    out, = rs.ransac_f8([(_d(pts, dev), _d(idx.view(np.int32), dev))], 1.0, 1, want_models=True)
    gm = out["models"].cpu().numpy()
    gc = _u32(out["counts"])
    errs = []
    for h in range(512):
        ok, F64 = orc.f8_fit_f64(pts[idx[h]].astype(np.float64))
        if not ok or gc[h] == 0:
            continue
        errs.append(np.abs(_norm(gm[h]) - _norm(F64)).max())
    errs = np.array(errs)
    print("model error (max |F32 - F64|, unit Frobenius): median %.3g p99 %.3g" % (np.median(errs), np.quantile(errs, 0.99)))
    assert np.median(errs) < 1e-4
    # inlier decisions of the best hypothesis: FP32 vs FP64 differ only at the threshold
    b = int(np.argmax(gc))
    _, F64 = orc.f8_fit_f64(pts[idx[b]].astype(np.float64))
    m64 = orc.f_inlier_f64(F64, pts, 1.0, 1)
    m32 = out["mask"].cpu().numpy().astype(bool)
    assert (m64 != m32).sum() <= 0.01 * len(pts)


def test_device_sampler_and_edges(rs, orc, dev):
    _say("device sampler vs restatement; small n; out-of-range indices; device counts")
    for n, H, seed in [(4096, 1000, 7), (9, 64, 1), (8, 4, 2)]:
        g = rs.ransac_sample_indices(n, H, seed, device=dev)
        assert np.array_equal(_u32(g), orc.sample_indices_hash(seed, n, H))
    g = rs.ransac_sample_indices(5, 3, 1, device=dev)
    assert (_u32(g) == 0xFFFFFFFF).all()
    pts, _, _ = orc.gen_two_view(8, 600, 200)
    idx = orc.sample_indices(2, 600, 300)
    bad = idx.copy()
    bad[10, 3] = 600  # out of range -> rejected
    ndev = torch.tensor([450], dtype=torch.int32, device=dev)  # only the first 450 points count
    outs = rs.ransac_f8([(_d(pts, dev), _d(bad.view(np.int32), dev)),
                         (_d(pts, dev), _d(idx.view(np.int32), dev), ndev),
                         (_d(pts[:5], dev), _d(idx[:4].view(np.int32), dev))], 1.0, 1)
    c0 = _u32(outs[0]["counts"])
    r0, _ = orc.f8_score(pts, bad, 1.0, 1)
    assert c0[10] == 0 and np.array_equal(c0, r0)
    c1 = _u32(outs[1]["counts"])
    sub = idx.copy()
    r1 = np.zeros(300, np.uint32)
    keep = (sub < 450).all(1)
    rr, _ = orc.f8_score(pts[:450], np.where(keep[:, None], sub, 0), 1.0, 1)
    r1[keep] = rr[keep]
    assert np.array_equal(c1, r1)
    assert (_u32(outs[2]["counts"]) == 0).all()
    assert (outs[2]["mask"].cpu().numpy() == 0).all()


def test_batched_pairs_uneven(rs, orc, dev):
    _say("20 pairs of uneven size in one call")
    probs, refs = [], []
    rng = np.random.default_rng(3)
    for p in range(20):
        n = int(rng.integers(20, 3000))
        H = int(rng.integers(1, 700))
        pts, _, _ = orc.gen_two_view(100 + p, n, n // 2)
        idx = orc.sample_indices(200 + p, n, H)
        probs.append((_d(pts, dev), _d(idx.view(np.int32), dev)))
        refs.append(orc.f8_score(pts, idx, 2.0, 0)[0])
    outs = rs.ransac_f8(probs, 2.0, 0)
    for o, r in zip(outs, refs):
        assert np.array_equal(_u32(o["counts"]), r)


@pytest.mark.parametrize("k", [4, 5, 7, 8])
def test_device_sampler_k(orc, dev, k):
    import ctypes

    from rapidsfm_amd._abi import SampleTask, check, lib

    print("[case] sampler k", k, flush=True)
    ns = [3, k - 1, k, 50, 3000]
    idx = [torch.full((333, k), -5, dtype=torch.int32, device=dev) for _ in ns]
    tasks = (SampleTask * len(ns))()
    for i, n in enumerate(ns):
        tasks[i] = SampleTask(idx[i].data_ptr(), 333, n, None, 1000 + i)
    check(lib.rsfm_ransac_sample_indices_batch_k(ctypes.cast(tasks, ctypes.c_void_p), len(ns), k,
                                                 ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "sample k")
    torch.cuda.synchronize()
    for i, n in enumerate(ns):
        assert np.array_equal(_u32(idx[i]), orc.sample_indices_hash_k(1000 + i, n, k, 333)), (k, n)
    with pytest.raises(RuntimeError):
        check(lib.rsfm_ransac_sample_indices_batch_k(ctypes.cast(tasks, ctypes.c_void_p), 1, 9, None), "sample k")
