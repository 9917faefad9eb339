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
