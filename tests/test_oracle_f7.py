"""Pins of the 7-point oracle (no buildable reference: the reference's solver needs
Eigen; its companion-matrix route is reproduced with numpy here)."""
import numpy as np


def _cubic_from_numpy(p7):
    """Independent construction: numpy SVD null space of the normalised 7x9
    system and numpy companion roots of det(F2 + t (F1 - F2))."""
    p = p7.astype(np.float64)
    c = p.mean(0)
    s0 = np.sqrt(2) / np.linalg.norm(p[:, :2] - c[:2], axis=1).mean()
    s1 = np.sqrt(2) / np.linalg.norm(p[:, 2:] - c[2:], axis=1).mean()
    x0, y0 = (p[:, 0] - c[0]) * s0, (p[:, 1] - c[1]) * s0
    x1, y1 = (p[:, 2] - c[2]) * s1, (p[:, 3] - c[3]) * s1
    P = np.stack([x1 * x0, y1 * x0, x0, x1 * y0, y1 * y0, y0, x1, y1, np.ones(7)], 1)
    V = np.linalg.svd(P)[2]
    F1 = V[8].reshape(3, 3).T
    F2 = V[7].reshape(3, 3).T
    D = F1 - F2
    ts = np.array([0.0, 1.0, -1.0, 2.0])
    d = [np.linalg.det(F2 + t * D) for t in ts]
    coef = np.polyfit(ts, d, 3)
    r = np.roots(coef)
    return int(np.sum(np.abs(r.imag) < 1e-9 * np.maximum(1, np.abs(r.real))))


def _err(Fs, F):
    Ft = F / np.linalg.norm(F)
    return min(min(np.abs(f - Ft).max(), np.abs(f + Ft).max()) for f in Fs)


def test_noise_free_geometry_recovers_true_F(orc):
    for seed in range(6):
        pts, _, F = orc.gen_two_view(seed, 300, 0, noise_px=0.0)
        rng = np.random.default_rng(seed)
        for _ in range(10):
            p7 = pts[rng.choice(300, 7, replace=False)]
            F64 = orc.f7_fit_f64(p7)
            F32 = orc.f7_fit_f32(p7)
            assert 1 <= len(F64) <= 3 and len(F32) == len(F64)
            assert _err(F64, F) < 2e-5  # inputs are FP32 pixel coordinates
            assert _err(F32, F) < 5e-3


def test_every_solution_is_an_epipolar_singular_matrix(orc):
    rng = np.random.default_rng(4)
    for _ in range(30):
        p7 = np.c_[rng.uniform(0, 4000, (7, 1)), rng.uniform(0, 3000, (7, 1)), rng.uniform(0, 4000, (7, 1)),
                   rng.uniform(0, 3000, (7, 1))].astype(np.float32)
        X0 = np.c_[p7[:, :2], np.ones(7)]
        X1 = np.c_[p7[:, 2:], np.ones(7)]
        for F in orc.f7_fit_f64(p7):
            assert abs(np.linalg.norm(F) - 1) < 1e-12
            assert abs(np.linalg.det(F)) < 1e-12
            # the seven constraints, relative to the size of the terms
            res = np.abs(np.einsum("ni,ij,nj->n", X1, F, X0))
            scale = np.einsum("ni,ij,nj->n", np.abs(X1), np.abs(F), np.abs(X0))
            assert np.all(res <= 1e-9 * scale)


def test_root_count_matches_companion_roots(orc):
    rng = np.random.default_rng(7)
    agree = 0
    for _ in range(200):
        p7 = np.c_[rng.uniform(0, 4000, (7, 1)), rng.uniform(0, 3000, (7, 1)), rng.uniform(0, 4000, (7, 1)),
                   rng.uniform(0, 3000, (7, 1))].astype(np.float32)
        agree += len(orc.f7_fit_f64(p7)) == _cubic_from_numpy(p7)
    assert agree >= 198  # near-double roots may be split differently


def test_f32_and_f64_solutions_agree(orc):
    pts, _, F = orc.gen_two_view(11, 500, 0, noise_px=0.5)
    rng = np.random.default_rng(1)
    for _ in range(50):
        p7 = pts[rng.choice(500, 7, replace=False)]
        a, b = orc.f7_fit_f32(p7), orc.f7_fit_f64(p7)
        if len(a) != len(b):
            continue
        for f in b:
            assert min(min(np.abs(g - f).max(), np.abs(g + f).max()) for g in a) < 5e-3


def test_score_picks_best_model(orc):
    pts, inl, F = orc.gen_two_view(5, 800, 300, noise_px=0.5)
    rng = np.random.default_rng(2)
    idx = np.stack([rng.choice(800, 7, replace=False) for _ in range(64)]).astype(np.uint32)
    counts, models = orc.f7_score(pts, idx, 2.0, 0)
    for h in range(0, 64, 9):
        Fs = orc.f7_fit_f32(pts[idx[h]])
        cs = [sum(orc.f_inlier_f32(f.ravel(), q, 2.0, 0) for q in pts) for f in Fs]
        assert counts[h] == max(cs)
        assert np.array_equal(models[h], Fs[int(np.argmax(cs))].ravel())
    # out-of-range indices reject the hypothesis
    bad = idx[:2].copy()
    bad[0, 3] = 10000
    c2, m2 = orc.f7_score(pts, bad, 2.0, 0)
    assert c2[0] == 0 and not m2[0].any()
