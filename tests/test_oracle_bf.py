"""Pins of the oracle's brute-force matcher and match filters (CPU).

Distances are checked against numpy's own integer products (np.einsum / broadcasting),
argmin ties against np.lexsort, and the filters against hand-built cases with
known answers.
"""
import numpy as np
import pytest


def _np_dist(A, B):
    A = A.astype(np.int64)
    B = B.astype(np.int64)
    return (A * A).sum(1)[:, None] + (B * B).sum(1)[None, :] - 2 * np.einsum("ik,jk->ij", A, B)


def _lex_top2(D):
    """per row: (index, dist) of best and second by (dist, index) via lexsort"""
    n, m = D.shape
    best = np.zeros((n, 2), np.uint64)
    second = np.full((n, 2), 0xFFFFFFFF, np.uint64)
    for i in range(n):
        order = np.lexsort((np.arange(m), D[i]))
        best[i] = (order[0], D[i, order[0]])
        if m > 1:
            second[i] = (order[1], D[i, order[1]])
    return best.astype(np.uint32), second.astype(np.uint32)


@pytest.mark.parametrize("nA,nB", [(1, 1), (17, 33), (120, 77)])
def test_bf_top2_against_numpy(orc, nA, nB):
    A = orc.gen_random_u8(nA, nA)
    B = orc.gen_random_u8(nB + 1000, nB)
    D = _np_dist(A, B)
    eb, es = _lex_top2(D)
    ebB, _ = _lex_top2(D.T)
    ba, sa, bb = orc.bf_match_top2_u8(A, B)
    assert np.array_equal(ba, eb)
    assert np.array_equal(sa, es)
    assert np.array_equal(bb, ebB)
    b1, b2 = orc.bf_match_u8(A, B)
    assert np.array_equal(b1, eb) and np.array_equal(b2, ebB)


def test_bf_ties_lowest_index(orc):
    rng = np.random.default_rng(1)
    A = rng.integers(0, 256, (20, 128), dtype=np.uint8)
    B = np.concatenate([A[5:10], A[0:10], A[5:8]])
    ba, sa, bb = orc.bf_match_top2_u8(A, B)
    D = _np_dist(A, B)
    eb, es = _lex_top2(D)
    assert np.array_equal(ba, eb) and np.array_equal(sa, es)
    assert ba[5, 0] == 0 and sa[5, 0] == 10 and ba[5, 1] == 0 and sa[5, 1] == 0
    # train side: B[0] == B[10] == A[5]; lowest query index wins
    assert bb[0, 0] == 5


@pytest.mark.parametrize("nA,nB", [(1, 1), (7, 300), (300, 9), (2048, 2048)])
def test_bf_top2_fast_matches_plain(orc, nA, nB):
    """The cache-blocked oracle path (eight rows per pass, per-thread column
    minima merged at the end) against the plain one, including planted ties
    (duplicated train and query rows) across the 8-row blocks."""
    train, query, _ = orc.gen_sift_pair(5, max(nB, 1), max(nA, 1), min(nA, nB) // 3, 0.05)
    A, B = query[:nA].copy(), train[:nB].copy()
    if nA >= 16 and nB >= 16:
        B[nB // 2] = B[3]
        A[nA - 1] = A[2]
        A[9] = B[5]
    r_plain = orc.bf_match_top2_u8(A, B)
    for nthreads in (1, 3):
        r_fast = orc.bf_match_top2_fast(A, B, nthreads)
        for x, y in zip(r_fast, r_plain):
            assert np.array_equal(x, y)


def test_empty_inputs(orc):
    A = np.zeros((0, 128), np.uint8)
    B = orc.gen_random_u8(1, 5)
    ba, bb = orc.bf_match_u8(A, B)
    assert ba.shape == (0, 2)
    assert (bb == 0xFFFFFFFF).all()
    ba, bb = orc.bf_match_u8(B, A)
    assert (ba == 0xFFFFFFFF).all()


def test_sqrnorm_and_nibble_pack(orc):
    d = orc.gen_random_u8(4, 50)
    assert np.array_equal(orc.sqrnorm_u8(d), (d.astype(np.int64) ** 2).sum(1))
    d4 = orc.desc8b_to4b(d)
    # reference word formula: dst word i = ((src[2i] & 0xF0F0F0F0) >> 4) | (src[2i+1] & 0xF0F0F0F0)
    w = d.view("<u4").reshape(50, 32)
    exp = ((w[:, 0::2] & 0xF0F0F0F0) >> 4) | (w[:, 1::2] & 0xF0F0F0F0)
    assert np.array_equal(d4.view("<u4").reshape(50, 16), exp)
    nib = np.stack([d4 & 15, d4 >> 4], -1).reshape(50, 128).astype(np.int64)
    assert np.array_equal(orc.sqrnorm_u4(d4), (nib ** 2).sum(1))
    # 4-bit match = L2 on the high nibbles of the 8-bit descriptors
    e = orc.gen_random_u8(5, 40)
    e4 = orc.desc8b_to4b(e)
    ba, bb = orc.bf_match_u4(d4, e4)
    D = _np_dist(d >> 4, e >> 4)
    eb, _ = _lex_top2(D)
    assert np.array_equal(ba, eb)


def _bm(pairs):
    return np.array(pairs, np.uint32).reshape(-1, 2)


def test_cross_check_known_answer(orc):
    # queries 0..4, train 0..3
    fwd = _bm([(2, 10), (0, 5), (0xFFFFFFFF, 0xFFFFFFFF), (3, 7), (2, 9)])
    bwd = _bm([(1, 5), (0, 1), (4, 9), (0, 3)])
    out = orc.cross_check(fwd, bwd)
    # q1 <-> t0 ; q4 <-> t2 ; q0 -> t2 but t2 -> q4 ; q3 -> t3 but t3 -> q0
    assert out[:, :2].tolist() == [[1, 0], [4, 2]]
    assert out[:, 2].view(np.float32).tolist() == [5.0, 9.0]


def test_remove_conflicts_known_answer(orc):
    fwd = _bm([(2, 10), (0, 5), (0xFFFFFFFF, 0xFFFFFFFF), (2, 10), (2, 9), (1, 4)])
    out = orc.remove_conflicts(fwd, 4)
    # t0 <- q1 ; t1 <- q5 ; t2 <- q4 (9 < 10; the first of equal 10s would win otherwise)
    assert out[:, :2].tolist() == [[1, 0], [5, 1], [4, 2]]


def test_ratio_test_exact_boundary(orc):
    # sqrt(d1)/sqrt(d2) < 0.8  <=>  25 d1 < 16 d2 ; at equality it fails
    assert not orc.ratio_pass(16, 25)
    assert orc.ratio_pass(15, 25)
    assert orc.ratio_pass(64, 101) and not orc.ratio_pass(64, 100)
    assert orc.ratio_pass(0, 0) is False
    assert orc.ratio_pass(8323200, 0xFFFFFFFF)
    fwd = _bm([(0, 15), (1, 16)])
    sec = _bm([(1, 25), (0, 25)])
    bwd = _bm([(0, 15), (1, 16)])
    out = orc.filter_matches(fwd, sec, bwd, (4, 5))
    assert out[:, :2].tolist() == [[0, 0]]
