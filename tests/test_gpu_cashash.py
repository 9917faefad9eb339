"""GPU parity of cascade hashing against the CPU oracle (bit-exact)."""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "cashash_weights_seed11.npz")


def _say(*a):
    print("[case]", *a, flush=True)


def _d(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _h(t, dt=None):
    a = t.cpu().numpy()
    return a.view(dt) if dt is not None else a


@pytest.fixture(scope="module")
def ch():
    from rapidsfm_amd import cashash

    return cashash


def test_accumulate_bias_hash_primitives(ch, orc, dev):
    _say("accumulate / bias / hash (64 and 192 channels) / histogram / bounds / buckets, n=897 and 5000")
    z = np.load(GOLDEN)
    w64, base64 = z["weights"], z["bias_base"]
    w192, base192 = orc.cashash_gaussian_weights(5, 192)
    for n in (897, 5000):
        d = orc.gen_random_u8(n, n)
        dd = _d(d, dev)
        acc = ch.accumulate(dd)
        racc = orc.cashash_accumulate(d)
        assert np.array_equal(_h(acc, np.uint64), racc)
        acc2 = ch.accumulate(dd[:100], acc=acc.clone(), init=False)
        assert np.array_equal(_h(acc2, np.uint64), orc.cashash_accumulate(d[:100], racc))
        for w, base in ((w64, base64), (w192, base192)):
            bias = ch.compute_bias(_d(w, dev), _d(base, dev), acc, n)
            rbias = orc.cashash_compute_bias(w, base, racc, n)
            assert np.array_equal(_h(bias), rbias)
            codes, remap, hist = ch.compute_hash(_d(w, dev), bias, dd, with_histogram=True)
            rcodes, rremap = orc.cashash_compute_hash(w, rbias, d)
            assert np.array_equal(_h(codes), rcodes)
            if rremap is not None:
                assert np.array_equal(_h(remap, np.uint32), rremap)
            rhist = orc.cashash_histogram(rcodes)
            assert np.array_equal(_h(hist, np.uint32), rhist)
        h2 = ch.build_histogram(codes)
        assert np.array_equal(_h(h2, np.uint32), rhist)
        lo, up, mx = ch.bucket_bounds(h2)
        rlo, rup, rmx = orc.cashash_bucket_bounds(rhist)
        assert np.array_equal(_h(lo, np.uint32), rlo) and np.array_equal(_h(up, np.uint32), rup)
        assert int(mx.item()) == rmx
        b = _h(ch.build_buckets(codes, lo), np.uint32)
        rb = orc.cashash_build_buckets(rcodes, rlo)
        for s in range(8):
            for k in range(256):
                assert sorted(b[s, rlo[s, k]:rup[s, k]]) == sorted(rb[s, rlo[s, k]:rup[s, k]])


def _register(ch, orc, desc, dev, w, bias_np):
    dd = _d(desc, dev)
    img = ch.register_image(dd, _d(w, dev), _d(bias_np, dev))
    codes, remap = orc.cashash_compute_hash(w, bias_np, desc)
    assert np.array_equal(_h(img["codes"]), codes)
    return img, codes, remap


def test_config1_reference_mode_8k(ch, orc, dev):
    """8192 vs 8192, 30% noisy true matches, golden reference weights, 8 tables,
    bias from the running descriptor mean in registration order (query first)."""
    _say("config1 reference mode 8192x8192, 8 schemes")
    z = np.load(GOLDEN)
    w, base = z["weights"], z["bias_base"]
    train, query, gt = orc.gen_sift_pair(1, 8192, 8192, 2458, 0.05)
    acc_q = orc.cashash_accumulate(query)
    bias_q = orc.cashash_compute_bias(w, base, acc_q, 8192)
    acc_t = orc.cashash_accumulate(train, acc_q)
    bias_t = orc.cashash_compute_bias(w, base, acc_t, 16384)
    qi, qc, _ = _register(ch, orc, query, dev, w, bias_q)
    ti, tc, _ = _register(ch, orc, train, dev, w, bias_t)
    (gq, gt_), = ch.find_best_match([qi], ti, 8)
    rq, rt = orc.cashash_find_best(query, qc, train, tc, 8)
    assert np.array_equal(_h(gq, np.uint32), rq)
    assert np.array_equal(_h(gt_, np.uint32), rt)
    hit = (rq[:, 0].astype(np.int64) == gt) & (gt >= 0)
    print("true matches recovered: %d / %d" % (hit.sum(), (gt >= 0).sum()))
    assert hit.sum() > 0.8 * (gt >= 0).sum()


def test_six_tables_and_many_tasks(ch, orc, dev):
    _say("6 tables with 192-channel Gaussian weights; 20 queries vs one train")
    w, base = orc.cashash_gaussian_weights(7, 192)
    train, _, _ = orc.gen_sift_pair(2, 3000, 10, 0, 0.05)
    acc = orc.cashash_accumulate(train)
    bias = orc.cashash_compute_bias(w, base, acc, 3000)
    ti, tc, _ = _register(ch, orc, train, dev, w, bias)
    qimgs, refs = [], []
    for k in range(20):
        _, q, _ = orc.gen_sift_pair(2, 3000, 100 + 37 * k, 30 + k, 0.05)
        qi, qc, _ = _register(ch, orc, q, dev, w, bias)
        qimgs.append(qi)
        refs.append(orc.cashash_find_best(q, qc, train, tc, 6))
    outs = ch.find_best_match(qimgs, ti, 6)
    for (gq, gt_), (rq, rt) in zip(outs, refs):
        assert np.array_equal(_h(gq, np.uint32), rq)
        assert np.array_equal(_h(gt_, np.uint32), rt)


def test_planted_ties_both_directions(ch, orc, dev):
    """Duplicated descriptors on both sides: a duplicate lands in the same
    buckets as its original, so every query that meets one meets both at the
    same distance and the lower index must win, in both directions."""
    _say("planted duplicate descriptors, 3000 x 2000, 6 tables")
    w, base = orc.cashash_gaussian_weights(7, 192)
    train, query, _ = orc.gen_sift_pair(4, 3000, 2000, 600, 0.05)
    train, query = train.copy(), query.copy()
    train[2000:2100] = train[0:100]
    train[500:600] = train[2900:3000]
    query[1500:1600] = query[10:110]
    query[300:340] = query[1960:2000]
    acc = orc.cashash_accumulate(train)
    bias = orc.cashash_compute_bias(w, base, acc, 3000)
    ti, tc, _ = _register(ch, orc, train, dev, w, bias)
    qi, qc, _ = _register(ch, orc, query, dev, w, bias)
    (gq, gt_), = ch.find_best_match([qi], ti, 6)
    rq, rt = orc.cashash_find_best(query, qc, train, tc, 6)
    assert np.array_equal(_h(gq, np.uint32), rq)
    assert np.array_equal(_h(gt_, np.uint32), rt)
    # the planted copies exist in the oracle's answer (ties were exercised)
    assert np.isin(rq[:, 0], np.arange(0, 100)).any() and np.isin(rt[:, 0], np.arange(10, 110)).any()


@pytest.mark.parametrize("keep", [8, 16])
def test_config1_rerank_mode_8k(ch, orc, dev, keep):
    """Config 1: 8192 vs 8192, 30% noisy true matches (2458), 192-channel seeded
    Gaussian projections, 6 bucket tables, the `keep` Hamming-nearest of the
    union by 128-bit remap, then exact L2^2; both directions bit-exact against
    the oracle's restatement (the train direction as the oracle with the roles
    swapped)."""
    _say("config1 re-rank mode 8192x8192, 6 schemes, keep", keep)
    w, base = orc.cashash_gaussian_weights(3, 192)
    train, query, gt = orc.gen_sift_pair(1, 8192, 8192, 2458, 0.05)
    acc_q = orc.cashash_accumulate(query)
    bias_q = orc.cashash_compute_bias(w, base, acc_q, 8192)
    acc_t = orc.cashash_accumulate(train, acc_q)
    bias_t = orc.cashash_compute_bias(w, base, acc_t, 16384)
    qi, qc, qr = _register(ch, orc, query, dev, w, bias_q)
    ti, tc, tr = _register(ch, orc, train, dev, w, bias_t)
    assert np.array_equal(_h(qi["remap"], np.uint32), qr) and np.array_equal(_h(ti["remap"], np.uint32), tr)
    (gq, gt_), = ch.find_best_match_rerank([qi], ti, 6, keep)
    rq = orc.cashash_find_best_rerank(query, qc, qr, train, tc, tr, 6, keep)
    rt = orc.cashash_find_best_rerank(train, tc, tr, query, qc, qr, 6, keep)
    assert np.array_equal(_h(gq, np.uint32), rq)
    assert np.array_equal(_h(gt_, np.uint32), rt)
    hit = (rq[:, 0].astype(np.int64) == gt) & (gt >= 0)
    print("true matches recovered (keep %d): %d / %d" % (keep, hit.sum(), (gt >= 0).sum()))
    assert hit.sum() > 0.6 * (gt >= 0).sum()


def test_rerank_many_tasks_and_edges(ch, orc, dev):
    _say("re-rank: 12 queries of uneven size vs one train, keep 32, empty query")
    w, base = orc.cashash_gaussian_weights(7, 192)
    train, _, _ = orc.gen_sift_pair(2, 3000, 10, 0, 0.05)
    acc = orc.cashash_accumulate(train)
    bias = orc.cashash_compute_bias(w, base, acc, 3000)
    ti, tc, tr = _register(ch, orc, train, dev, w, bias)
    qimgs, refs = [], []
    for k in range(12):
        _, q, _ = orc.gen_sift_pair(2, 3000, 1 + 251 * k, 40 + k, 0.05)
        qi, qc, qr = _register(ch, orc, q, dev, w, bias)
        qimgs.append(qi)
        refs.append((orc.cashash_find_best_rerank(q, qc, qr, train, tc, tr, 5, 32),
                     orc.cashash_find_best_rerank(train, tc, tr, q, qc, qr, 5, 32)))
    outs = ch.find_best_match_rerank(qimgs, ti, 5, 32)
    for (gq, gt_), (rq, rt) in zip(outs, refs):
        assert np.array_equal(_h(gq, np.uint32), rq)
        assert np.array_equal(_h(gt_, np.uint32), rt)


def test_rerank_batches_of_directions(ch, orc, dev):
    """20 query images, both directions: 40 searches in one call, more than one
    launch batch holds."""
    _say("re-rank: 20 queries vs one train, both directions, keep 8")
    w, base = orc.cashash_gaussian_weights(5, 192)
    train, _, _ = orc.gen_sift_pair(6, 1500, 10, 0, 0.05)
    acc = orc.cashash_accumulate(train)
    bias = orc.cashash_compute_bias(w, base, acc, 1500)
    ti, tc, tr = _register(ch, orc, train, dev, w, bias)
    qimgs, refs = [], []
    for k in range(20):
        _, q, _ = orc.gen_sift_pair(6, 1500, 200 + 37 * k, 60 + k, 0.05)
        qi, qc, qr = _register(ch, orc, q, dev, w, bias)
        qimgs.append(qi)
        refs.append((orc.cashash_find_best_rerank(q, qc, qr, train, tc, tr, 6, 8),
                     orc.cashash_find_best_rerank(train, tc, tr, q, qc, qr, 6, 8)))
    outs = ch.find_best_match_rerank(qimgs, ti, 6, 8)
    for (gq, gt_), (rq, rt) in zip(outs, refs):
        assert np.array_equal(_h(gq, np.uint32), rq)
        assert np.array_equal(_h(gt_, np.uint32), rt)
