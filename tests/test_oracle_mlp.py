"""Pins of the CPU oracle for the learned proposal network and the weighted sampler.

The network has no reference counterpart; the FP64 path is pinned by central
finite differences of its own loss (a dropped term, a wrong sign or a
transposed weight fails them) and by closed forms; the bf16 path is the same
code with roundings inserted, pinned against known round-to-nearest-even
values and by staying within bf16 error of the FP64 path.
"""
import numpy as np
import pytest


def _net(orc, seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.normal(0, scale, orc.MLP_NPARAM)


def _data(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, 8)).astype(np.float32)
    y = (rng.random(n) < 0.4).astype(np.float32)
    return x, y


def test_finite_differences(orc):
    p = _net(orc, 1)
    x, y = _data(2, 40)
    _, g = orc.mlp_loss_grad(p, x, y)
    rng = np.random.default_rng(3)
    # every parameter group: W1 b1 W2 b2 W3 b3 W4 b4
    groups = [(0, 512), (512, 576), (576, 4672), (4672, 4736), (4736, 8832), (8832, 8896), (8896, 8960), (8960, 8961)]
    eps = 1e-6
    for lo, hi in groups:
        for k in rng.choice(np.arange(lo, hi), size=min(6, hi - lo), replace=False):
            pp, pm = p.copy(), p.copy()
            pp[k] += eps
            pm[k] -= eps
            fd = (orc.mlp_loss_grad(pp, x, y, with_grad=False)[0] - orc.mlp_loss_grad(pm, x, y, with_grad=False)[0]) / (
                2 * eps)
            assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])


def test_closed_forms(orc):
    x, y = _data(4, 25)
    p = np.zeros(orc.MLP_NPARAM)
    p[-1] = 0.7  # b4: every output is sigmoid(b4), hidden units all zero
    prob = orc.mlp_forward(p, x)
    assert np.allclose(prob, 1 / (1 + np.exp(-0.7)), rtol=0, atol=1e-15)
    loss, g = orc.mlp_loss_grad(p, x, y)
    s = 1 / (1 + np.exp(-0.7))
    assert abs(loss - np.mean(-(y * np.log(s) + (1 - y) * np.log(1 - s)))) < 1e-12
    assert abs(g[-1] - np.mean(s - y)) < 1e-12
    assert np.all(g[:-1] == 0)  # relu'(0) = 0 and zero activations


def test_single_unit_chain(orc):
    # one path x0 -> h1[0] -> h2[0] -> h3[0] -> z with positive weights: z = w4 w3 w2 (w1 x0 + b1) + ...
    p = np.zeros(orc.MLP_NPARAM)
    p[0] = 0.5          # W1[0][0]
    p[512] = 0.25       # b1[0]
    p[576] = 2.0        # W2[0][0]
    p[4736] = 1.5       # W3[0][0]
    p[8896] = -0.8      # W4[0]
    p[8960] = 0.1       # b4
    x = np.zeros((1, 8), np.float32)
    x[0, 0] = 1.5
    z = -0.8 * 1.5 * 2.0 * (0.5 * 1.5 + 0.25) + 0.1
    assert abs(orc.mlp_forward(p, x)[0] - 1 / (1 + np.exp(-z))) < 1e-14


def test_bf16_rounding(orc):
    # round to nearest even on the 8-bit mantissa
    assert orc.bf16_round(1.0 + 2 ** -8) == 1.0
    assert orc.bf16_round(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6
    assert orc.bf16_round(1.0 + 2 ** -7) == 1.0 + 2 ** -7
    assert orc.bf16_round(-2.0 - 2 ** -7) == -2.0  # halfway between -2 and -2 - 2^-6: even
    assert orc.bf16_round(-2.0 - 3 * 2 ** -7) == -2.0 - 2 ** -5
    assert np.isnan(orc.bf16_round(float("nan")))


def test_bf16_mode_tracks_fp64(orc):
    p = _net(orc, 5, 0.2)
    x, y = _data(6, 200)
    l0, g0 = orc.mlp_loss_grad(p, x, y)
    l1, g1 = orc.mlp_loss_grad(p, x, y, bf16=True)
    assert abs(l0 - l1) < 5e-3
    assert np.abs(g0 - g1).max() < 0.05 * np.abs(g0).max()


def test_weighted_sampler(orc):
    # all-zero weights: identical to the uniform hash sampler
    z = np.zeros(50, np.float32)
    assert np.array_equal(orc.sample_weighted(z, 64, 9), orc.sample_indices_hash(9, 50, 64))
    # zero-weight points are never drawn
    w = np.zeros(300, np.float32)
    w[100:140] = 1.0
    s = orc.sample_weighted(w, 500, 3)
    assert np.all((s >= 100) & (s < 140))
    assert all(len(set(r)) == 8 for r in s)
    # fewer than 8 points with weight: no 8 distinct draws possible -> ~0u
    w2 = np.zeros(100, np.float32)
    w2[:5] = 1
    assert np.all(orc.sample_weighted(w2, 4, 1) == 0xFFFFFFFF)
    # first draws follow the weights (chi-square against q_i / total)
    w3 = np.array([0.1, 0.2, 0.3, 0.4] + [0.0] * 4 + [1.0] * 8, np.float32)
    s3 = orc.sample_weighted(w3, 20000, 5)
    counts = np.bincount(s3[:, 0], minlength=16)
    q = np.floor(np.clip(w3, 0, 1) * 65536 + 0.5)
    exp = q / q.sum() * 20000
    m = exp > 0
    chi2 = ((counts[m] - exp[m]) ** 2 / exp[m]).sum()
    assert chi2 < 40  # 11 dof, p ~ 1e-4
    assert np.all(counts[~m] == 0)
