"""Pins of the oracle's end of the 8-point F-RANSAC driver (CPU).

Reference: ransac.hpp:139-160 (adaptive stop), ImagePair.cpp:185-203 (refit on
the best sample's inliers, kept when it has more inliers).

- the library's own log2 (frexp + atanh series) against libm, and the
  required iteration count it gives against the one libm gives (the
  reference's formula) over a grid of (best, n, confidence);
- the walk over a scored table against the reference loop drawing the same
  samples (or_ransac_loop_hash, itself pinned to the reference's ransac.hpp
  compiled in oracle/_ref) and against ref_sampler directly;
- the refit: noise-free geometry gives the closed-form F; on noisy data the
  refit equals an independent numpy SVD least-squares fit on the same inliers;
  it uses the inliers, not the first k points (the reference's mask2pts slip);
  a refit is kept only when it has more inliers; edge cases.
"""
import math
import os
import subprocess

import numpy as np
import pytest

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "ref_sampler")


def _norm(F):
    F = np.asarray(F, np.float64).reshape(3, 3)
    F = F / np.linalg.norm(F)
    i = np.argmax(np.abs(F))
    return F * np.sign(F.flat[i])


def _numpy_ls(pts):
    """least-squares F on pts (m x 4): Hartley normalisation, SVD null vector of
    the m x 9 system, rank 2 by SVD, denormalised"""
    p = np.asarray(pts, np.float64)
    c = p.mean(0)
    d = p - c
    s0 = np.sqrt(2) / np.linalg.norm(d[:, :2], axis=1).mean()
    s1 = np.sqrt(2) / np.linalg.norm(d[:, 2:], axis=1).mean()
    x0, y0 = d[:, 0] * s0, d[:, 1] * s0
    x1, y1 = d[:, 2] * s1, d[:, 3] * s1
    P = np.stack([x1 * x0, y1 * x0, x0, x1 * y0, y1 * y0, y0, x1, y1, np.ones(len(p))], 1)
    f = np.linalg.svd(P)[2][-1]
    Fh = f.reshape(3, 3).T
    U, S, Vt = np.linalg.svd(Fh)
    Fh = U @ np.diag([S[0], S[1], 0]) @ Vt
    T0 = np.array([[s0, 0, -s0 * c[0]], [0, s0, -s0 * c[1]], [0, 0, 1]])
    T1 = np.array([[s1, 0, -s1 * c[2]], [0, s1, -s1 * c[3]], [0, 0, 1]])
    return T1.T @ Fh @ T0


def test_det_log2_against_libm(orc):
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(1e-12, 1.0, 20000), 1.0 - np.logspace(-15, -0.01, 2000),
                         np.logspace(-15, 0, 2000), [0.5, 0.25, 2.0 ** -40, 1.0]])
    for x in xs:
        a, b = orc.det_log2(x), math.log2(x)
        assert abs(a - b) <= 4 * np.spacing(abs(b)) + 1e-300, x
    assert orc.det_log2(1.0) == 0.0
    assert orc.det_log2(0.5) == -1.0
    assert orc.det_log2(0.0) == -math.inf


def test_required_iters_det_matches_libm(orc):
    """The formula of ransac.hpp:155-157 with libm (or_ransac_required_iters)
    and with the library's log2 agree on every point of the grid."""
    for n in list(range(8, 200)) + list(range(200, 16385, 331)):
        for best in range(8, n + 1, max(1, n // 150)):
            for conf in (0.9, 0.99, 0.995, 0.999):
                assert orc.required_iters_det(best, n, conf, 100000) == orc.required_iters(best, n, 8, conf, 100000)
    # the closed form on a round case, the caps and the degenerate quotients
    assert orc.required_iters_det(700, 1000, 0.995, 100000) == int(np.ceil(np.log2(0.005) / np.log2(1 - 0.7 ** 8)))
    assert orc.required_iters_det(1, 100000, 0.99, 5000) == 5000  # (1/n)^8 vanishes: log2(1) = 0
    assert orc.required_iters_det(50, 50, 0.99, 5000) == 0  # all inliers: log2(0) = -inf
    assert orc.required_iters_det(60, 100, 0.99, 7) == 7


def _hash_counts(idx, modulus):
    s = np.zeros(idx.shape[0], np.uint64)
    with np.errstate(over="ignore"):
        for i in range(8):
            s = s * np.uint64(1000003) + idx[:, i].astype(np.uint64)
        return ((s * np.uint64(2654435761)) % np.uint64(modulus)).astype(np.uint32)


@pytest.mark.parametrize("n,conf,maxit,seed,mod", [(1000, 0.995, 5000, 3, 50), (300, 0.99, 2000, 9, 7),
                                                   (50, 0.9, 100, 1, 2), (64, 0.999, 5000, 4, 1000),
                                                   (200, 0.99, 300, 11, 3)])
def test_cut_walk_equals_reference_loop(orc, n, conf, maxit, seed, mod):
    """The scored table of the reference's own sample stream, walked, stops
    where the drawing loop stops and picks the same best sample."""
    import ctypes as ct

    bi = np.zeros(8, np.uint32)
    draws = ct.c_uint32(0)
    best = orc._lib.or_ransac_loop_hash(n, conf, maxit, seed, mod, bi.ctypes.data_as(ct.c_void_p), ct.byref(draws))
    idx = orc.sample_indices(seed, n, draws.value + 50)
    counts = _hash_counts(idx, mod)
    walked, bh, bc = orc.adaptive_cut(counts, n, conf, maxit)
    assert walked == draws.value
    assert bc == best
    assert list(idx[bh]) == list(bi)
    if os.path.exists(REF):
        out = subprocess.run([REF, "loop", str(n), str(conf), str(maxit), str(seed), str(mod)], capture_output=True,
                             text=True, check=True).stdout.split()
        assert int(out[0]) == bc and int(out[1]) == walked and [int(x) for x in out[2:10]] == list(idx[bh])


def test_cut_insane_run_and_table_end(orc):
    # more than min(128, maxIter) zero counts in a row end the walk
    counts = np.zeros(400, np.uint32)
    counts[3] = 5
    walked, bh, bc = orc.adaptive_cut(counts, 1000, 0.99, 5000)
    assert (walked, bh, bc) == (3 + 1 + 129, 3, 5)
    walked, bh, bc = orc.adaptive_cut(counts, 1000, 0.99, 10)
    assert (walked, bh, bc) == (4 + 11, 3, 5)
    # no sane hypothesis at all
    assert orc.adaptive_cut(np.zeros(50, np.uint32), 100, 0.99, 5000) == (50, 0xFFFFFFFF, 0)
    # the table ends before the required count: all walked, first of the largest kept
    counts = np.array([3, 9, 9, 4], np.uint32)
    assert orc.adaptive_cut(counts, 1000, 0.99, 5000) == (4, 1, 9)
    # everything an inlier: required = 0 after the first sane draw
    assert orc.adaptive_cut(np.array([0, 100, 100], np.uint32), 100, 0.99, 5000) == (2, 1, 100)


def test_refit_noise_free_gives_true_F(orc):
    pts, inl, Ft = orc.gen_two_view(5, 600, 200, noise_px=0.0)
    idx = orc.sample_indices_hash(3, 600, 256)
    counts, models = orc.f8_score(pts, idx, 1.0, 1)
    F, res, mask = orc.f8_finish(pts, idx, counts, models, 1.0, 1, refit=True)
    assert res[3] >= res[1] >= 8
    assert np.abs(_norm(F) - _norm(Ft)).max() < 2e-5
    assert mask.sum() == res[3]
    assert mask[inl].all()


@pytest.mark.parametrize("seed,mode,t", [(1, 1, 1.0), (2, 0, 3.0), (7, 1, 2.0)])
def test_refit_equals_numpy_least_squares(orc, seed, mode, t):
    pts, inl, Ft = orc.gen_two_view(seed, 1500, 600, noise_px=0.5)
    idx = orc.sample_indices_hash(seed, 1500, 512)
    counts, models = orc.f8_score(pts, idx, t, mode)
    bh = int(np.argmax(counts))
    F0 = models[bh]
    m0 = np.array([orc._lib.or_f_inlier_f32(orc._p(np.ascontiguousarray(F0.reshape(9))),
                                            orc._p(np.ascontiguousarray(pts[i])), t, mode) for i in range(len(pts))], bool)
    F, res, mask = orc.f8_finish(pts, idx, counts, models, t, mode, refit=True)
    assert res[0] == bh and res[1] == counts[bh] == m0.sum()
    Fn = _numpy_ls(pts[m0])
    if res[3] > res[1]:
        # kept: the refit is the least-squares F of the best sample's inliers
        assert np.abs(_norm(F) - _norm(Fn)).max() < 1e-5
    # the refit is at least as good; with noise it usually wins
    assert res[3] >= res[1]
    assert mask.sum() == res[3]


def test_refit_wins_on_noisy_data(orc):
    kept = 0
    for seed in range(6):
        pts, _, _ = orc.gen_two_view(100 + seed, 2000, 800, noise_px=0.5)
        idx = orc.sample_indices_hash(seed, 2000, 256)
        counts, models = orc.f8_score(pts, idx, 1.0, 1)
        F, res, mask = orc.f8_finish(pts, idx, counts, models, 1.0, 1, refit=True)
        kept += int(res[3] > res[1])
        F2, res2, _ = orc.f8_finish(pts, idx, counts, models, 1.0, 1, refit=False)
        assert res2[3] == res2[1] == res[1]
        assert np.array_equal(F2, models[res[0]])
    assert kept >= 4


def test_refit_uses_inliers_not_first_points(orc):
    """The reference's mask2pts gathers pts[*][k] for the k-th inlier
    (legacy/geometry.cpp:341-354), i.e. the first k points. Here the first
    points are all outliers: a refit on them explains almost nothing, the
    refit on the inliers explains at least the sample's count."""
    pts, inl, Ft = orc.gen_two_view(9, 1200, 0, noise_px=0.3)
    rng = np.random.default_rng(9)
    nout = 400
    out = np.stack([rng.uniform(0, 4000, nout), rng.uniform(0, 3000, nout), rng.uniform(0, 4000, nout),
                    rng.uniform(0, 3000, nout)], 1).astype(np.float32)
    allp = np.concatenate([out, pts]).astype(np.float32)
    idx = orc.sample_indices_hash(4, len(allp), 512)
    counts, models = orc.f8_score(allp, idx, 1.0, 1)
    F, res, mask = orc.f8_finish(allp, idx, counts, models, 1.0, 1, refit=True)
    slip = _numpy_ls(allp[: res[1]])  # the first k points
    slip_count = sum(orc._lib.or_f_inlier_f32(orc._p(np.ascontiguousarray(slip.astype(np.float32).reshape(9))),
                                              orc._p(np.ascontiguousarray(allp[i])), 1.0, 1) for i in range(len(allp)))
    assert slip_count < res[1] // 4
    assert res[3] >= res[1] > 800
    assert not mask[:nout].any() or mask[:nout].sum() < 20


def test_finish_edge_cases(orc):
    pts, _, _ = orc.gen_two_view(3, 300, 100)
    idx = orc.sample_indices_hash(1, 300, 64)
    counts, models = orc.f8_score(pts, idx, 1.0, 1)
    # fewer than 8 points: nothing
    F, res, mask = orc.f8_finish(pts[:7], idx[:, :], np.zeros(64, np.uint32), None, 1.0, 1)
    assert list(res) == [0xFFFFFFFF, 0, 0, 0] and not F.any() and not mask.any()
    # every count zero: the reference's fallback sample 0..7
    F, res, mask = orc.f8_finish(pts, idx, np.zeros(64, np.uint32), None, 1.0, 1, refit=False)
    ok, F07 = orc.f8_fit_f32(pts[:8])
    assert res[0] == 0xFFFFFFFF and ok and np.array_equal(F, F07)
    # models absent: the best sample is refitted, identically
    F1, res1, m1 = orc.f8_finish(pts, idx, counts, models, 1.0, 1)
    F2, res2, m2 = orc.f8_finish(pts, idx, counts, None, 1.0, 1)
    assert np.array_equal(F1, F2) and np.array_equal(res1, res2) and np.array_equal(m1, m2)
    # the adaptive cut walks a prefix and picks its best
    F3, res3, _ = orc.f8_finish(pts, idx, counts, models, 1.0, 1, conf=0.99, max_iter=5000)
    w, bh, bc = orc.adaptive_cut(counts, 300, 0.99, 5000)
    assert res3[0] == bh and res3[2] == w and res3[1] == bc
