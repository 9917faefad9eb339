"""Learned proposals end to end on the GPU: feature parity with the oracle, plain-RANSAC
labels, training with the project's own kernels, and the gain of guided over uniform
sampling at 70% outliers (plain uniform RANSAC is the judge of every count)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def test_features_match_oracle(orc, dev):
    from rapidsfm_amd import proposal, synth

    _say("proposal features")
    pairs = [synth.make_pair(s, 1500, 1700, 0.5, 0.8, device=dev) for s in (3, 4)]
    res = proposal.match_pairs(pairs, dev)
    for pr, r in zip(pairs, res):
        m = r["matches"].cpu().numpy().view(np.uint32)
        ref = orc.proposal_features(m, r["secA"].cpu().numpy().view(np.uint32), pr["kpA"], pr["kpB"], synth.WIDTH,
                                    synth.HEIGHT)
        x = r["x"].cpu().numpy()
        assert x.shape == ref.shape and len(x) > 100
        assert np.abs(x - ref).max() <= 1e-6 * np.abs(ref).max()
        # the cross-checked matches and gathered points agree with the oracle's filter
        ba = orc.bf_match_top2_u8(pr["descA"].cpu().numpy(), pr["descB"].cpu().numpy())
        ref_m = orc.cross_check(ba[0], ba[2])
        assert np.array_equal(m[:, :2], ref_m[:, :2].astype(np.uint32))


def test_guided_sampling_gain(dev):
    from rapidsfm_amd import proposal

    _say("proposal end to end (small)")
    rng = np.random.default_rng(0)
    n_train = 192
    train = proposal.make_split(11, n_train, rng.integers(1024, 2049, n_train), rng.uniform(0.5, 0.9, n_train), dev)
    proposal.label_with_plain_ransac(train, H=4096)
    acc = np.mean(np.concatenate([(s["y"].cpu().numpy() > 0.5) == s["gt"] for s in train]))
    assert acc > 0.9, acc  # plain-RANSAC labels agree with the synthetic truth
    params, losses = proposal.train(train, epochs=6, batch=4096)
    assert losses[-1] < losses[0]
    held = proposal.make_split(99, 24, 4096, 0.7, dev)
    rep = proposal.evaluate(params, held, H_plain=32768, H_guided=2048)
    print(rep, flush=True)
    assert rep["guided_reaches_99pct"] >= 0.95
    assert rep["gain"] >= 3.0
