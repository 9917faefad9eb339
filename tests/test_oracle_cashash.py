"""Pins of the oracle's cascade hashing (CPU).

Closed forms via numpy (sums, integer matrix products, bincount/cumsum,
candidate-masked brute force), the bit mapping of the reference's own test
(test/casHashComputeHashTest.cpp:120-138) on its data recipe, and the golden
weights fixture generated from the reference constructor's recipe.
"""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "cashash_weights_seed11.npz")


def _np_bits(w, bias, desc):
    return ((desc.astype(np.int64) >> 1) @ w.astype(np.int64).T + bias[None, :].astype(np.int64)) >= 0


def _np_codes(bits):
    # code byte s = channels 8s..8s+7, bit b = channel 8s+b
    return np.packbits(bits[:, :64].reshape(-1, 8, 8), axis=2, bitorder="little").reshape(-1, 8)


def test_accumulate_closed_form(orc):
    d = orc.gen_random_u8(3, 1000)
    acc = orc.cashash_accumulate(d)
    assert np.array_equal(acc, (d.astype(np.uint64) >> 1).sum(0))
    acc2 = orc.cashash_accumulate(d[:10], acc)
    assert np.array_equal(acc2, acc + (d[:10].astype(np.uint64) >> 1).sum(0))


def test_bias_exact_on_integer_centres_and_close_otherwise(orc):
    z = np.load(GOLDEN)
    w, base = z["weights"], z["bias_base"]
    # acc = n * integer centre -> every product and partial sum is an exact float
    n = 1000
    centre = np.arange(128) % 60
    acc = (centre * n).astype(np.uint64)
    got = orc.cashash_compute_bias(w, base, acc, n)
    exp = -(w.astype(np.int64) @ centre.astype(np.int64)) + base
    assert np.array_equal(got, exp)
    # general centres: within +-1 of the float64 value (the reference test's tolerance)
    d = orc.gen_random_u8(4, 777)
    acc = orc.cashash_accumulate(d)
    got = orc.cashash_compute_bias(w, base, acc, 777)
    ref = np.round(w.astype(np.float64) @ (-acc.astype(np.float64) / 777)) + base
    assert np.abs(got - ref).max() <= 1


def test_hash_bits_match_reference_test_recipe(orc):
    # test/casHashComputeHashTest.cpp:38-75: random weight/descriptor words,
    # bias = roundf(sum w * -63.5); exact bit check (:120-138)
    rng = np.random.default_rng(11)
    w = rng.integers(-128, 128, (192, 128)).astype(np.int8)
    bias = np.round(w.astype(np.float64).sum(1) * -63.5).astype(np.int32)
    d = rng.integers(0, 256, (897, 128)).astype(np.uint8)
    codes, remap = orc.cashash_compute_hash(w, bias, d)
    bits = _np_bits(w, bias, d)
    assert np.array_equal(codes, _np_codes(bits))
    rb = np.packbits(bits[:, 64:].reshape(-1, 4, 32), axis=2, bitorder="little").view("<u4").reshape(-1, 4)
    assert np.array_equal(remap, rb)
    codes64, none = orc.cashash_compute_hash(w[:64], bias[:64], d)
    assert none is None and np.array_equal(codes64, codes)


def test_one_hot_weights_plant_known_bits(orc):
    # channel c looks at dimension c only: bit = (byte_c >> 1) >= t_c
    w = np.zeros((64, 128), np.int8)
    w[np.arange(64), np.arange(64)] = 1
    t = np.arange(64) % 128
    bias = -t.astype(np.int32)
    d = np.zeros((3, 128), np.uint8)
    d[0, :64] = 255
    d[1, :64] = (2 * t).astype(np.uint8)
    d[2, :64] = np.maximum(2 * t - 2, 0).astype(np.uint8)
    codes, _ = orc.cashash_compute_hash(w, bias, d)
    bits = np.unpackbits(codes, axis=1, bitorder="little").astype(bool)
    assert bits[0].all()
    assert bits[1].all()
    assert np.array_equal(bits[2], t == 0)


def test_histogram_bounds_buckets(orc):
    rng = np.random.default_rng(2)
    codes = rng.integers(0, 256, (897, 8)).astype(np.uint8)
    h = orc.cashash_histogram(codes)
    assert np.array_equal(h, np.stack([np.bincount(codes[:, s], minlength=256) for s in range(8)]))
    lo, up, mx = orc.cashash_bucket_bounds(h)
    assert np.array_equal(up, np.cumsum(h, 1)) and np.array_equal(lo, up - h) and mx == h.max()
    b = orc.cashash_build_buckets(codes, lo)
    for s in range(8):
        for k in range(256):
            members = sorted(b[s, lo[s, k]:up[s, k]].tolist())
            assert members == np.nonzero(codes[:, s] == k)[0].tolist()


@pytest.mark.parametrize("nscheme", [8, 6, 1])
def test_find_best_against_masked_brute_force(orc, nscheme):
    rng = np.random.default_rng(nscheme)
    q = rng.integers(0, 256, (300, 128)).astype(np.uint8)
    t = np.concatenate([q[::4], rng.integers(0, 256, (200, 128)).astype(np.uint8)])
    qc = rng.integers(0, 4, (300, 8)).astype(np.uint8)
    tc = np.concatenate([qc[::4], rng.integers(0, 4, (200, 8)).astype(np.uint8)])
    bq, bt = orc.cashash_find_best(q, qc, t, tc, nscheme)
    D = ((q.astype(np.int64)[:, None, :] - t.astype(np.int64)[None, :, :]) ** 2).sum(2)
    mask = (qc[:, None, :nscheme] == tc[None, :, :nscheme]).any(2)
    BIG = np.iinfo(np.int64).max
    Dm = np.where(mask, D, BIG)
    for i in range(300):
        if not mask[i].any():
            assert tuple(bq[i]) == (0xFFFFFFFF, 0xFFFFFFFF)
            continue
        j = np.lexsort((np.arange(Dm.shape[1]), Dm[i]))[0]
        assert tuple(bq[i]) == (j, D[i, j])
    for j in range(t.shape[0]):
        if not mask[:, j].any():
            assert tuple(bt[j]) == (0xFFFFFFFF, 0xFFFFFFFF)
            continue
        i = np.lexsort((np.arange(Dm.shape[0]), Dm[:, j]))[0]
        assert tuple(bt[j]) == (i, D[i, j])


def test_rerank_mode_keeps_hamming_nearest(orc):
    rng = np.random.default_rng(9)
    q = rng.integers(0, 256, (50, 128)).astype(np.uint8)
    t = rng.integers(0, 256, (80, 128)).astype(np.uint8)
    qc = np.zeros((50, 8), np.uint8)
    tc = np.zeros((80, 8), np.uint8)  # everything shares every bucket
    qr = rng.integers(0, 2**32, (50, 4), dtype=np.uint64).astype(np.uint32)
    tr = rng.integers(0, 2**32, (80, 4), dtype=np.uint64).astype(np.uint32)
    bq = orc.cashash_find_best_rerank(q, qc, qr, t, tc, tr, 6, 5)
    D = ((q.astype(np.int64)[:, None, :] - t.astype(np.int64)[None, :, :]) ** 2).sum(2)
    ham = np.array([[sum(bin(int(a) ^ int(b)).count("1") for a, b in zip(qr[i], tr[j])) for j in range(80)]
                    for i in range(50)])
    for i in range(50):
        keep = np.lexsort((np.arange(80), ham[i]))[:5]
        j = min(keep, key=lambda jj: (D[i, jj], jj))
        assert tuple(bq[i]) == (j, D[i, j])


def test_golden_weights_fixture():
    z = np.load(GOLDEN)
    w, b = z["weights"], z["bias_base"]
    assert w.shape == (64, 128) and w.dtype == np.int8 and b.shape == (64,)
    assert abs(float(w.mean())) < 2 and 28 < float(w.std()) < 36
    assert w.max() <= 127 and w.min() >= -128
