"""GPU parity of the affine RANSAC match filter against the CPU oracle (votes bit-exact)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _scene(seed, n_in, n_out, w=4000, h=3000, clustered=False):
    rng = np.random.default_rng(seed)
    A = np.array([[1.05, 0.1], [-0.08, 0.97]])
    t = np.array([30.0, -20.0])
    if clustered:
        centres = rng.uniform([0, 0], [w, h], (6, 2))
        left = centres[rng.integers(0, 6, n_in)] + rng.normal(0, 120, (n_in, 2))
        left = np.clip(left, 0, [w - 1, h - 1])
    else:
        left = rng.uniform([0, 0], [w, h], (n_in, 2))
    right = left @ A.T + t + rng.normal(0, 1.0, (n_in, 2))
    ol = rng.uniform([0, 0], [w, h], (n_out, 2))
    orr = rng.uniform([0, 0], [w, h], (n_out, 2))
    m = np.concatenate([np.concatenate([left, right], 1), np.concatenate([ol, orr], 1)]).astype(np.float32)
    return m[rng.permutation(len(m))]


@pytest.mark.parametrize("try_other", [False, True])
@pytest.mark.parametrize("seed,n_in,n_out,clustered", [(1, 3000, 1000, False), (2, 800, 1600, True), (3, 40, 10, False)])
def test_votes_bit_exact(orc, dev, try_other, seed, n_in, n_out, clustered):
    from rapidsfm_amd import rmf

    _say("rmf scene", seed, n_in, n_out, clustered, "tryOther", try_other)
    m = _scene(seed, n_in, n_out, clustered=clustered)
    reordered, bounds, order, cols, rows, cw = orc.rmf_bin(4000, 3000, m, 16)
    r2, b2, o2, c2, rw2, cw2 = rmf.bin_matches(4000, 3000, m, 16)
    assert np.array_equal(r2, reordered) and np.array_equal(b2, bounds) and np.array_equal(o2, order)
    assert (c2, rw2, cw2) == (cols, rows, cw)
    for s in (0, 987654321):
        ref, raff = orc.rmf_filter(reordered, bounds, cols, rows, 256, cw * 0.15, try_other, s)
        votes, aff = rmf.filter_votes(torch.from_numpy(reordered).to(dev),
                                      torch.from_numpy(bounds.view(np.int32)).to(dev), cols, rows, 256, cw * 0.15,
                                      try_other, s)
        assert np.array_equal(votes.cpu().numpy(), ref)
        ga = aff.cpu().numpy()
        assert np.array_equal(np.isnan(ga), np.isnan(raff))
        assert np.array_equal(np.nan_to_num(ga).view(np.uint32), np.nan_to_num(raff).view(np.uint32))
    mask = rmf.get_inlier_mask((4000, 3000), m, 2, 256, 0.15, 16, try_other, seed=0, device=dev)
    ref0, _ = orc.rmf_filter(reordered, bounds, cols, rows, 256, cw * 0.15, try_other, 0)
    exp = np.zeros(len(m), bool)
    exp[order] = ref0 >= 2
    assert np.array_equal(mask, exp)


def test_small_inputs(dev):
    from rapidsfm_amd import rmf

    _say("rmf fewer than 16 matches")
    m = _scene(9, 10, 3)
    mask = rmf.get_inlier_mask((4000, 3000), m, 2, 256, 0.15, 16, True, seed=1, device=dev)
    assert not mask.any()


def _batch_ref(orc, m, w, h, seed, min_votes=2, try_other=True):
    """the oracle's binning + filter, then the caller's selection: kept pairs in order"""
    reordered, bounds, order, cols, rows, cw = orc.rmf_bin(w, h, m, 16)
    votes, _ = orc.rmf_filter(reordered, bounds, cols, rows, 256, cw * 0.15, try_other, seed)
    mask = np.zeros(len(m), bool)
    mask[order] = votes >= min_votes
    return mask


@pytest.mark.parametrize("try_other", [True, False])
def test_batch_filter_matches_oracle(orc, dev, try_other):
    """rsfm_rmf_filter_batch over 10 pairs of uneven size and image shape (one
    below 16 points, one empty), with a device count below the capacity on
    some: masks, compacted pairs and matches, kept counts identical to the
    oracle's binning + filter + selection."""
    from rapidsfm_amd import rmf

    _say("rmf batch, 10 pairs, tryOther", try_other)
    specs = [(3000, 1000, 4000, 3000), (800, 1600, 4000, 3000), (40, 10, 4000, 3000), (5, 4, 4000, 3000),
             (0, 0, 4000, 3000), (1500, 500, 3000, 4000), (2500, 2500, 1920, 1080), (6000, 3000, 4000, 3000),
             (700, 300, 640, 480), (9000, 7000, 4000, 3000)]
    probs, refs = [], []
    for k, (ni, no, w, h) in enumerate(specs):
        m = _scene(100 + k, ni, no, w=w, h=h) if ni + no else np.zeros((0, 4), np.float32)
        cap = len(m) + (37 if k % 3 == 0 else 0)
        buf = np.zeros((max(1, cap), 4), np.float32)
        buf[: len(m)] = m
        matches = np.zeros((max(1, cap), 3), np.int32)
        matches[: len(m), 0] = np.arange(len(m))
        matches[: len(m), 1] = np.arange(len(m)) * 7
        pd = torch.from_numpy(buf).to(dev)
        md = torch.from_numpy(matches).to(dev)
        cnt = torch.tensor([len(m)], dtype=torch.int32, device=dev)
        probs.append((pd, cnt, (w, h), 50 + k, md))
        refs.append((m, _batch_ref(orc, m, w, h, 50 + k, try_other=try_other) if len(m) else np.zeros(0, bool),
                     matches[: len(m)]))
    masks = rmf.filter_batch(probs, try_other=try_other)
    torch.cuda.synchronize()
    for k, ((pd, cnt, _, _, md), (m, mask, mt), gm) in enumerate(zip(probs, refs, masks)):
        n = len(m)
        got = gm.cpu().numpy().astype(bool)
        assert np.array_equal(got[:n], mask), k
        assert not got[n:].any(), k
        kept = int(mask.sum())
        assert int(cnt.item()) == kept, k
        assert np.array_equal(pd[:kept].cpu().numpy(), m[mask]), k
        assert np.array_equal(md[:kept].cpu().numpy(), mt[mask]), k
    # the first pair keeps its inliers and sheds most outliers
    assert int(probs[0][1].item()) > 2500
