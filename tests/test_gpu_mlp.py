"""GPU parity of the learned-proposal kernels: bf16 MLP forward/gradient against the
CPU oracle (bf16-emulating and exact FP64 modes), Adam step, determinism, and the
weighted sampler bit for bit."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _problem(seed, n, scale=0.25):
    rng = np.random.default_rng(seed)
    from rapidsfm_amd import learned

    p = learned.init_params(seed, device="cpu").numpy().astype(np.float64) * (scale / 0.5)
    p[learned.OFF_B1:learned.OFF_W2] = rng.normal(0, 0.1, 64)
    p[learned.OFF_B2:learned.OFF_W3] = rng.normal(0, 0.1, 64)
    p[learned.OFF_B3:learned.OFF_W4] = rng.normal(0, 0.1, 64)
    p[learned.OFF_B4] = 0.2
    x = rng.normal(0, 1, (n, 8)).astype(np.float32)
    y = (rng.random(n) < 0.35).astype(np.float32)
    return p.astype(np.float32), x, y


@pytest.mark.parametrize("n", [1, 31, 1000, 4099])
def test_forward_matches_bf16_oracle(orc, dev, n):
    from rapidsfm_amd import learned

    _say("mlp forward n", n)
    p, x, _ = _problem(11 + n, n)
    prob = learned.forward(torch.from_numpy(p).to(dev), torch.from_numpy(x).to(dev)).cpu().numpy()
    ref_q = orc.mlp_forward(p.astype(np.float64), x, bf16=True)
    ref = orc.mlp_forward(p.astype(np.float64), x)
    assert np.abs(prob - ref_q).max() < 2e-3
    assert np.abs(prob - ref).max() < 3e-2


@pytest.mark.parametrize("n", [7, 128, 1000, 9000])
def test_gradient_matches_oracle(orc, dev, n):
    from rapidsfm_amd import learned

    _say("mlp grad n", n)
    p, x, y = _problem(3 + n, n)
    pt, xt, yt = (torch.from_numpy(a).to(dev) for a in (p, x, y))
    g, loss = learned.grad(pt, xt, yt)
    g = g.cpu().numpy().astype(np.float64)
    lq, gq = orc.mlp_loss_grad(p.astype(np.float64), x, y, bf16=True)
    l0, g0 = orc.mlp_loss_grad(p.astype(np.float64), x, y)
    assert abs(loss.item() - lq) < 1e-3 * max(1.0, abs(lq))
    # per parameter group, against the bf16-emulating oracle (FP32 accumulation
    # order and rare one-ulp bf16 rounding flips are the only differences)
    from rapidsfm_amd.learned import OFF_B1, OFF_B2, OFF_B3, OFF_B4, OFF_W2, OFF_W3, OFF_W4
    groups = [(0, OFF_B1), (OFF_B1, OFF_W2), (OFF_W2, OFF_B2), (OFF_B2, OFF_W3), (OFF_W3, OFF_B3), (OFF_B3, OFF_W4),
              (OFF_W4, OFF_B4), (OFF_B4, OFF_B4 + 1)]
    for lo, hi in groups:
        scale = np.abs(gq[lo:hi]).max() + 1e-12
        assert np.abs(g[lo:hi] - gq[lo:hi]).max() <= 1e-2 * scale + 1e-7, (lo, hi)
    # and within bf16 error of the exact FP64 gradient
    assert np.abs(g - g0).max() <= 6e-2 * np.abs(g0).max()
    cos = float(g @ g0 / np.linalg.norm(g) / np.linalg.norm(g0))
    assert cos > 0.995


def test_gradient_deterministic(dev):
    from rapidsfm_amd import learned

    _say("mlp grad determinism")
    p, x, y = _problem(5, 20000)
    pt, xt, yt = (torch.from_numpy(a).to(dev) for a in (p, x, y))
    g1, l1 = learned.grad(pt, xt, yt)
    g2, l2 = learned.grad(pt, xt, yt)
    assert torch.equal(g1, g2) and torch.equal(l1, l2)


def test_adam_step(dev):
    from rapidsfm_amd import learned

    _say("mlp adam")
    rng = np.random.default_rng(2)
    p = rng.normal(0, 1, learned.N_PARAM).astype(np.float32)
    pt = torch.from_numpy(p.copy()).to(dev)
    opt = learned.Adam(pt, lr=0.01)
    m = np.zeros_like(p, dtype=np.float64)
    v = np.zeros_like(p, dtype=np.float64)
    pr = p.astype(np.float64)
    for t in range(1, 4):
        g = rng.normal(0, 1, learned.N_PARAM).astype(np.float32)
        opt.step(torch.from_numpy(g).to(dev))
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g.astype(np.float64) ** 2
        pr = pr - 0.01 * (m / (1 - 0.9 ** t)) / (np.sqrt(v / (1 - 0.999 ** t)) + 1e-8)
    assert np.abs(pt.cpu().numpy() - pr).max() < 1e-5


def test_training_separates(dev):
    """A few hundred Adam steps on a separable problem drive the loss down."""
    from rapidsfm_amd import learned

    _say("mlp training")
    rng = np.random.default_rng(8)
    n = 8192
    x = rng.normal(0, 1, (n, 8)).astype(np.float32)
    y = ((x[:, 0] + 0.5 * x[:, 1] - x[:, 2] * x[:, 3] * 0.3) > 0.2).astype(np.float32)
    xt, yt = torch.from_numpy(x).to(dev), torch.from_numpy(y).to(dev)
    pt = learned.init_params(1, device=dev)
    opt = learned.Adam(pt, lr=3e-3)
    first = None
    for it in range(300):
        g, loss = learned.grad(pt, xt, yt)
        opt.step(g)
        if first is None:
            first = loss.item()
    last = learned.grad(pt, xt, yt)[1].item()
    acc = ((learned.forward(pt, xt) > 0.5).float() == yt).float().mean().item()
    assert last < 0.3 * first and acc > 0.93, (first, last, acc)


@pytest.mark.parametrize("n,H,kind", [(50, 64, "zero"), (4096, 512, "rand"), (300, 256, "sparse"), (5, 4, "rand"),
                                      (16000, 1024, "peaked")])
def test_weighted_sampler_bit_exact(orc, dev, n, H, kind):
    from rapidsfm_amd import learned

    _say("weighted sampler", n, H, kind)
    rng = np.random.default_rng(n + H)
    if kind == "zero":
        w = np.zeros(n, np.float32)
    elif kind == "sparse":
        w = np.zeros(n, np.float32)
        w[rng.choice(n, 20, replace=False)] = rng.random(20).astype(np.float32)
    elif kind == "peaked":
        w = (rng.random(n) ** 8).astype(np.float32)
    else:
        w = rng.random(n).astype(np.float32)
    w[: min(3, n)] = [np.nan, -1.0, 2.0][: min(3, n)]  # clamped like the oracle
    seed = 1234 + n
    (idx,) = learned.sample_weighted([(torch.from_numpy(w).to(dev), H, seed)])
    ref = orc.sample_weighted(w, H, seed)
    assert np.array_equal(idx.cpu().numpy().view(np.uint32), ref)


def test_weighted_sampler_device_count(orc, dev):
    from rapidsfm_amd import learned

    _say("weighted sampler device count")
    rng = np.random.default_rng(4)
    w = rng.random(1000).astype(np.float32)
    nd = torch.tensor([600], dtype=torch.int32, device=dev)
    (idx,) = learned.sample_weighted([(torch.from_numpy(w).to(dev), 128, 77, nd)])
    assert np.array_equal(idx.cpu().numpy().view(np.uint32), orc.sample_weighted(w[:600], 128, 77))
