"""GPU parity of the brute-force matcher against the CPU oracle (bit-exact).

Every case prints its name and sizes before touching the device.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _np(t):
    return t.cpu().numpy().view(np.uint32)


def _to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def test_sqrnorm_and_4bit_precompute(rs, orc, dev):
    _say("sqrnorm/8b->4b n=1037 full-range u8")
    d = orc.gen_random_u8(11, 1037)
    dd = _to_dev(d, dev)
    assert np.array_equal(_np(rs.sqrnorm_u8(dd)), orc.sqrnorm_u8(d))
    d4 = rs.desc8b_to_4b(dd)
    ref4 = orc.desc8b_to4b(d)
    assert np.array_equal(d4.cpu().numpy(), ref4)
    assert np.array_equal(_np(rs.sqrnorm_u4(d4)), orc.sqrnorm_u4(ref4))


@pytest.mark.parametrize("impl", ["mfma", "idp"])
def test_config0_top1_bidirectional(rs, orc, dev, impl):
    _say("config0 2048x2048 top1", impl)
    train, query, gt = orc.gen_sift_pair(0, 2048, 2048, 819, 0.05)
    ra, rb = orc.bf_match_u8(query, train)
    (ga, gb), = rs.bf_match([(_to_dev(query, dev), _to_dev(train, dev))], impl=impl)
    torch.cuda.synchronize()
    assert np.array_equal(_np(ga), ra)
    assert np.array_equal(_np(gb), rb)
    # the noisy copies are found
    hit = (ra[:, 0].astype(np.int64) == gt) & (gt >= 0)
    assert hit.sum() > 0.95 * (gt >= 0).sum()


def test_config0_top2_ratio_crosscheck(rs, orc, dev):
    _say("config0 2048x2048 top2 + ratio 0.8 + cross-check")
    train, query, _ = orc.gen_sift_pair(0, 2048, 2048, 819, 0.05)
    ra, sa, rb = orc.bf_match_top2_u8(query, train)
    (ga, gs, gb), = rs.bf_match_top2([(_to_dev(query, dev), _to_dev(train, dev))])
    assert np.array_equal(_np(ga), ra)
    assert np.array_equal(_np(gs), sa)
    assert np.array_equal(_np(gb), rb)
    ref = orc.filter_matches(ra, sa, rb, (4, 5))
    out, cnt = rs.filter_matches(ga, gs, gb, 2048, (4, 5))
    n = int(cnt.item())
    assert n == ref.shape[0]
    assert np.array_equal(out[:n].cpu().numpy().view(np.uint32), ref)
    # cross-check only
    ref_cc = orc.cross_check(ra, rb)
    out2, cnt2 = rs.filter_matches(ga, None, gb, 2048)
    n2 = int(cnt2.item())
    assert n2 == ref_cc.shape[0]
    assert np.array_equal(out2[:n2].cpu().numpy().view(np.uint32), ref_cc)


@pytest.mark.parametrize("nB", [50, 100, 150, 300, 4160, 8257])
def test_top2_chunk_counts(rs, orc, dev, nB):
    """Top-2 over train ranges of 1, 2, 3 and 5 chunks of 64 columns (odd
    chunk counts end with a lone chunk after the last trip) and over several
    4096-column segments: 4160 = 4096 + 64 and 8257 = 2 x 4096 + 65 columns."""
    _say("top2 chunk counts, nB", nB)
    train, query, _ = orc.gen_sift_pair(12, nB, 700, min(nB, 700) // 3, 0.05)
    ra, sa, rb = orc.bf_match_top2_u8(query, train)
    (ga, gs, gb), = rs.bf_match_top2([(_to_dev(query, dev), _to_dev(train, dev))])
    assert np.array_equal(_np(ga), ra)
    assert np.array_equal(_np(gs), sa)
    assert np.array_equal(_np(gb), rb)


@pytest.mark.parametrize("nA,nB", [(1, 1), (31, 65), (897, 1000), (257, 63), (300, 2049), (5, 9000)])
@pytest.mark.parametrize("impl", ["mfma", "idp"])
def test_ragged_full_range(rs, orc, dev, nA, nB, impl):
    _say("ragged full-range u8", nA, nB, impl)
    A = orc.gen_random_u8(100 + nA, nA)
    B = orc.gen_random_u8(200 + nB, nB)
    ra, rb = orc.bf_match_u8(A, B)
    (ga, gb), = rs.bf_match([(_to_dev(A, dev), _to_dev(B, dev))], impl=impl)
    assert np.array_equal(_np(ga), ra)
    assert np.array_equal(_np(gb), rb)
    if impl == "mfma":
        qa, qs, qb = orc.bf_match_top2_u8(A, B)
        (ha, hs, hb), = rs.bf_match_top2([(_to_dev(A, dev), _to_dev(B, dev))])
        assert np.array_equal(_np(ha), qa)
        assert np.array_equal(_np(hs), qs)
        assert np.array_equal(_np(hb), qb)


def test_extreme_values_and_ties(rs, orc, dev):
    _say("extreme bytes 0/255 and planted duplicate descriptors")
    rng = np.random.default_rng(5)
    A = rng.choice(np.array([0, 255], np.uint8), size=(300, 128))
    B = np.concatenate([A[::3], A[::3], rng.integers(0, 256, (50, 128), dtype=np.uint8), A[::5]])
    A[7] = 255
    B[3] = 255
    B[40] = 255
    ra, sa, rb = orc.bf_match_top2_u8(A, B)
    # duplicates: the lowest index must win
    assert ra[7, 0] == 3 and ra[7, 1] == 0 and sa[7, 0] == 40 and sa[7, 1] == 0
    (ga, gs, gb), = rs.bf_match_top2([(_to_dev(A, dev), _to_dev(B, dev))])
    assert np.array_equal(_np(ga), ra)
    assert np.array_equal(_np(gs), sa)
    assert np.array_equal(_np(gb), rb)
    for impl in ("mfma", "idp"):
        (ta, tb), = rs.bf_match([(_to_dev(A, dev), _to_dev(B, dev))], impl=impl)
        assert np.array_equal(_np(ta), ra), impl
        assert np.array_equal(_np(tb), rb), impl


def test_empty_sides(rs, orc, dev):
    _say("empty query/train sides")
    A = _to_dev(orc.gen_random_u8(1, 40), dev)
    E = torch.empty((0, 128), dtype=torch.uint8, device=dev)
    (ga, gb), (ha, hb) = rs.bf_match([(A, E), (E, A)])
    assert (_np(ga) == 0xFFFFFFFF).all() and gb.numel() == 0
    assert ha.numel() == 0 and (_np(hb) == 0xFFFFFFFF).all()
    (ta, ts, tb), = rs.bf_match_top2([(A, E)])
    assert (_np(ta) == 0xFFFFFFFF).all() and (_np(ts) == 0xFFFFFFFF).all()
    # cross-check skips index ~0u (bruteForceMatch.cpp:29-31)
    out, cnt = rs.filter_matches(ga, None, tb, 0)
    assert int(cnt.item()) == 0


def test_many_uneven_tasks(rs, orc, dev):
    _say("40 tasks of uneven sizes in one call (crosses the 32-task batch)")
    rng = np.random.default_rng(7)
    pairs, refs = [], []
    for t in range(40):
        nA, nB = int(rng.integers(1, 700)), int(rng.integers(1, 900))
        A = orc.gen_random_u8(1000 + t, nA)
        B = orc.gen_random_u8(2000 + t, nB)
        pairs.append((_to_dev(A, dev), _to_dev(B, dev)))
        refs.append(orc.bf_match_top2_u8(A, B))
    outs = rs.bf_match([(a, b) for a, b in pairs])
    outs2 = rs.bf_match_top2(pairs)
    for (ga, gb), (ha, hs, hb), (ra, sa, rb) in zip(outs, outs2, refs):
        assert np.array_equal(_np(ga), ra) and np.array_equal(_np(gb), rb)
        assert np.array_equal(_np(ha), ra) and np.array_equal(_np(hs), sa) and np.array_equal(_np(hb), rb)


def test_4bit_mode(rs, orc, dev):
    _say("4-bit mode 700x900")
    A = orc.gen_random_u8(31, 700)
    B = orc.gen_random_u8(32, 900)
    A4, B4 = orc.desc8b_to4b(A), orc.desc8b_to4b(B)
    ra, rb = orc.bf_match_u4(A4, B4)
    (ga, gb), = rs.bf_match([(_to_dev(A4, dev), _to_dev(B4, dev))], impl="u4")
    assert np.array_equal(_np(ga), ra)
    assert np.array_equal(_np(gb), rb)


def test_segmented_train_range(rs, orc, dev):
    _say("one query block vs 20000 train (several segments)")
    train, query, _ = orc.gen_sift_pair(3, 20000, 600, 240, 0.05)
    ra, sa, rb = orc.bf_match_top2_u8(query, train)
    (ga, gs, gb), = rs.bf_match_top2([(_to_dev(query, dev), _to_dev(train, dev))])
    assert np.array_equal(_np(ga), ra)
    assert np.array_equal(_np(gs), sa)
    assert np.array_equal(_np(gb), rb)
    (ta, tb), = rs.bf_match([(_to_dev(query, dev), _to_dev(train, dev))])
    assert np.array_equal(_np(ta), ra)
    assert np.array_equal(_np(tb), rb)


def test_8k_properties(rs, orc, dev):
    """8K x 8K (the benchmark size): every row and column against the oracle,
    plus size-independent properties."""
    _say("8192x8192 exact, all rows and columns")
    train, query, gt = orc.gen_sift_pair(9, 8192, 8192, 2458, 0.05)
    qd, td = _to_dev(query, dev), _to_dev(train, dev)
    (ga, gs, gb), = rs.bf_match_top2([(qd, td)])
    (ta, tb), = rs.bf_match([(qd, td)])
    a, s, b = _np(ga), _np(gs), _np(gb)
    ra, sa, rb = orc.bf_match_top2_fast(query, train)
    assert np.array_equal(a, ra) and np.array_equal(s, sa) and np.array_equal(b, rb)
    assert np.array_equal(a, _np(ta)) and np.array_equal(b, _np(tb))
    (pa, pb), = rs.bf_match([(qd, td)], impl="idp")
    assert np.array_equal(a, _np(pa)) and np.array_equal(b, _np(pb))
    # determinism
    (ga2, gs2, gb2), = rs.bf_match_top2([(qd, td)])
    assert np.array_equal(a, _np(ga2)) and np.array_equal(s, _np(gs2)) and np.array_equal(b, _np(gb2))
    # best <= second, distances recomputed exactly
    assert ((a[:, 1] < s[:, 1]) | ((a[:, 1] == s[:, 1]) & (a[:, 0] < s[:, 0]))).all()
    qi = query.astype(np.int64)
    ti = train.astype(np.int64)
    d_best = ((qi - ti[a[:, 0]]) ** 2).sum(1)
    assert np.array_equal(d_best, a[:, 1].astype(np.int64))
    d_b = ((ti - qi[b[:, 0]]) ** 2).sum(1)
    assert np.array_equal(d_b, b[:, 1].astype(np.int64))


@pytest.mark.parametrize("nq", [1, 1025, 8192, 8193, 20000])
def test_filter_trip_boundaries(rs, orc, dev, nq):
    """The device filter takes 1024 queries per round and 8192 per trip; match
    tables of these sizes (about half the forward matches mutual, 5% "no match")
    must give the oracle's ordered output with and without each input."""
    _say("filter: synthetic match tables, nQ", nq)
    rng = np.random.default_rng(nq)
    nt = 12000
    dist1 = rng.integers(0, 1 << 20, nq).astype(np.uint64)
    idx = rng.integers(0, nt, nq).astype(np.uint64)
    idx[rng.random(nq) < 0.05] = 0xFFFFFFFF
    fwd = (dist1 << np.uint64(32)) | idx
    dist2 = dist1 + rng.integers(0, 1 << 19, nq).astype(np.uint64)
    second = (dist2 << np.uint64(32)) | rng.integers(0, nt, nq).astype(np.uint64)
    bq = rng.integers(0, nq, nt).astype(np.uint64)
    sel = np.nonzero((idx != 0xFFFFFFFF) & (rng.random(nq) < 0.6))[0]
    bq[idx[sel].astype(np.int64)] = sel.astype(np.uint64)
    bwd = (rng.integers(0, 1 << 20, nt).astype(np.uint64) << np.uint64(32)) | bq

    def host(a):
        return None if a is None else a.view(np.uint32).reshape(-1, 2)

    def device(a):
        return None if a is None else torch.from_numpy(a.view(np.int32).reshape(-1, 2).copy()).to(dev)

    for sec, bw in [(second, bwd), (None, bwd), (second, None)]:
        ref = orc.filter_matches(host(fwd), host(sec), host(bw), (4, 5))
        out, cnt = rs.filter_matches(device(fwd), device(sec), device(bw), nt, (4, 5))
        n = int(cnt.item())
        assert n == ref.shape[0]
        assert np.array_equal(out[:n].cpu().numpy().view(np.uint32), ref)


def test_top2_caller_workspace_and_stream(rs, orc, dev):
    """A caller workspace at an unaligned offset, too small after alignment,
    and a side stream: the wrapper must realign or replace it and the results
    must not change."""
    _say("top2 with caller workspaces and a side stream")
    A = orc.gen_random_u8(301, 700)
    B = orc.gen_random_u8(302, 1500)
    qa, qs, qb = orc.bf_match_top2_u8(A, B)
    pair = [(_to_dev(A, dev), _to_dev(B, dev))]
    big = torch.empty(4 << 20, dtype=torch.uint8, device=dev)
    side = torch.cuda.Stream()
    for ws, st in [(big[3:], None), (big[: 1 << 10], None), (None, side), (big[5:], side)]:
        with torch.cuda.stream(side if st is not None else torch.cuda.current_stream()):
            (ha, hs, hb), = rs.bf_match_top2(pair, stream=st, workspace=ws)
        torch.cuda.synchronize()
        assert np.array_equal(_np(ha), qa)
        assert np.array_equal(_np(hs), qs)
        assert np.array_equal(_np(hb), qb)
