"""Pins of the oracle's affine match filter (CPU)."""
import numpy as np
import pytest


def test_gauss_elimination_against_numpy(orc):
    rng = np.random.default_rng(1)
    for _ in range(50):
        A = rng.normal(size=(3, 3)).astype(np.float32) + 3 * np.eye(3, dtype=np.float32)
        b = rng.normal(size=(3, 2)).astype(np.float32)
        x = orc.solve_gauss3(A, b)
        ref = np.linalg.solve(A.astype(np.float64), b.astype(np.float64))
        assert np.allclose(x, ref, rtol=1e-4, atol=1e-5)
    # exactly representable system: every operation is exact in float32
    A = np.array([[2, 0, 0], [4, 4, 0], [8, 8, 8]], np.float32)
    b = np.array([[2, 4], [12, 8], [40, 24]], np.float32)
    assert np.array_equal(orc.solve_gauss3(A, b), np.array([[1, 2], [2, 0], [2, 1]], np.float32))


def test_binning_known_answer(orc):
    m = np.array([[0, 0, 1, 1], [250, 250, 0, 0], [251, 0, 0, 0], [3999, 2999, 0, 0], [1000, 1, 0, 0]], np.float32)
    reordered, bounds, order, cols, rows, cw = orc.rmf_bin(4000, 3000, m, 16)
    assert (cols, rows, cw) == (16, 12, 250)
    # cells: (0,0)->0 ; (ceil(1),ceil(1)) = (1,1)->17 ; (0, ceil(1.004)=2)->2 ; clamp (11,15)->191 ; (ceil(.004)=1, 4)->20
    cells = [0, 17, 2, 191, 20]
    exp_order = sorted(range(5), key=lambda k: (cells[k], k))
    assert order.tolist() == exp_order
    assert np.array_equal(reordered, m[exp_order])
    counts = np.bincount(cells, minlength=192)
    assert np.array_equal(bounds, np.concatenate([[0], np.cumsum(counts)]))


def _scene(seed, n_in, n_out, w=4000, h=3000):
    rng = np.random.default_rng(seed)
    A = np.array([[1.05, 0.1], [-0.08, 0.97]], np.float64)
    t = np.array([30.0, -20.0])
    left = rng.uniform([0, 0], [w, h], (n_in, 2))
    right = left @ A.T + t
    ol = rng.uniform([0, 0], [w, h], (n_out, 2))
    orr = rng.uniform([0, 0], [w, h], (n_out, 2))
    m = np.concatenate([np.concatenate([left, right], 1), np.concatenate([ol, orr], 1)]).astype(np.float32)
    perm = rng.permutation(len(m))
    inl = np.concatenate([np.ones(n_in, bool), np.zeros(n_out, bool)])[perm]
    return m[perm], inl


@pytest.mark.parametrize("try_other", [False, True])
def test_votes_on_affine_scene(orc, try_other):
    m, inl = _scene(3, 1500, 500)
    reordered, bounds, order, cols, rows, cw = orc.rmf_bin(4000, 3000, m, 16)
    votes, aff = orc.rmf_filter(reordered, bounds, cols, rows, 256, cw * 0.15, try_other, 12345)
    mask = np.zeros(len(m), bool)
    mask[order] = votes >= 2
    assert mask[inl].mean() > 0.95
    assert mask[~inl].mean() < 0.02
    assert votes.max() <= 9
    v2, _ = orc.rmf_filter(reordered, bounds, cols, rows, 256, cw * 0.15, try_other, 12345)
    assert np.array_equal(votes, v2)


def test_exact_affine_votes_closed_form(orc):
    """Every match follows one exact affine: each hypothesis recovers it, so a
    cell votes iff its 3x3 neighbourhood holds >= 6 matches (and, without
    trying other cells' models, iff its own model was fitted: >= 9 matches).
    The expected votes per match follow from the cell counts alone."""
    rng = np.random.default_rng(4)
    left = rng.uniform([0, 0], [4000, 3000], (600, 2)).round()
    right = left * 2 + np.array([4.0, -8.0])  # exactly representable
    m = np.concatenate([left, right], 1).astype(np.float32)
    reordered, bounds, order, cols, rows, cw = orc.rmf_bin(4000, 3000, m, 16)
    cnt = np.diff(bounds.astype(np.int64)).reshape(rows, cols)
    nb = np.zeros_like(cnt)
    for r in range(rows):
        for c in range(cols):
            nb[r, c] = cnt[max(0, r - 1):r + 2, max(0, c - 1):c + 2].sum()
    cell_of = np.repeat(np.arange(rows * cols), cnt.ravel())
    for try_other in (False, True):
        votes, _ = orc.rmf_filter(reordered, bounds, cols, rows, 64, cw * 0.15, try_other, 7)
        voting = (nb >= 6) if try_other else (nb >= 9)
        exp = np.zeros(len(m), np.int64)
        for k in range(len(m)):
            r, c = divmod(cell_of[k], cols)
            exp[k] = voting[max(0, r - 1):r + 2, max(0, c - 1):c + 2].sum()
        assert np.array_equal(votes.astype(np.int64), exp)


def test_few_matches_no_votes(orc):
    m, _ = _scene(5, 10, 5)
    reordered, bounds, order, cols, rows, cw = orc.rmf_bin(4000, 3000, m, 16)
    votes, _ = orc.rmf_filter(reordered, bounds, cols, rows, 256, cw * 0.15, True, 1)
    assert (votes == 0).all()
