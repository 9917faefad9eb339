"""Learned RANSAC proposals end to end (no reference counterpart; the reference
samples minimal sets uniformly, ransac.hpp:23-82).

  synthetic pairs -> top-2 match + cross-check (rsfm_bf_match_top2,
  rsfm_filter_matches_batch) -> per-match features (rsfm_proposal_features_batch)
  -> labels from plain RANSAC (uniform 8-point draws among the ratio-test
  survivors, scored over all matches; the inlier set of its best model)
  -> training of the 8-64-64-64-1 network on the GPU (rsfm_mlp_grad, Adam)
  -> guided sampling (rsfm_ransac_sample_weighted_batch with the network's
  probabilities) against uniform sampling on held-out pairs, both judged by
  the same plain scorer (rsfm_ransac_f8, Sampson 1 px).

The comparison: for each held-out pair, C* is the best inlier count either
sampler found and a hypothesis "succeeds" if its count is >= 0.99 C*. With the
pooled success rate q of a sampler, ceil(ln(0.01) / ln(1 - q)) hypotheses reach
99% of the best count with 99% confidence; the gain is the ratio of those
counts (plain over guided).
"""
import ctypes
import math

import numpy as np
import torch

from . import learned, synth
from ._abi import BfTask, BfTop2Task, FilterTask, RansacTask, SampleTask, check, lib
from ._abi import bind_extra

RATIO = (4, 5)  # ratio test 0.8 on distances (25 d1 < 16 d2 on squared distances)


class FeatureTask(ctypes.Structure):
    _fields_ = [
        ("matches", ctypes.c_void_p),
        ("count", ctypes.c_void_p),
        ("maxMatches", ctypes.c_uint32),
        ("second", ctypes.c_void_p),
        ("kpQuery", ctypes.c_void_p),
        ("kpTrain", ctypes.c_void_p),
        ("width", ctypes.c_float),
        ("height", ctypes.c_float),
        ("x", ctypes.c_void_p),
    ]


bind_extra("rsfm_proposal_features_batch", ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p])


def _sp():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _arr(cls, items):
    a = (cls * max(1, len(items)))()
    for i, it in enumerate(items):
        a[i] = it
    return a


def match_pairs(pairs, device, hard_outliers=None):
    """Top-2 match + cross-check (no ratio test) + features for a list of synth pairs.

    Returns a list of dicts (device tensors): matches (m, 3) int32 [query, train,
    distance bits], x (m, 8) features, pts (m, 4); m = kept matches.

    hard_outliers: None, or a numpy Generator: then every kept outlier match takes
    the scale ratio and orientation difference of a random kept inlier of its pair
    (its query keypoint's scale and orientation are rewritten before the features
    are computed), so outliers and inliers share the scale / orientation cue.
    """
    P = len(pairs)
    bufs = []
    for pr in pairs:
        nA, nB = pr["descA"].shape[0], pr["descB"].shape[0]
        b = {
            "nA": nA, "nB": nB,
            "bestA": torch.empty((nA, 2), dtype=torch.int32, device=device),
            "secA": torch.empty((nA, 2), dtype=torch.int32, device=device),
            "bestB": torch.empty((nB, 2), dtype=torch.int32, device=device),
            "out": torch.empty((max(1, nA), 3), dtype=torch.int32, device=device),
            "count": torch.zeros(1, dtype=torch.int32, device=device),
            "kA": torch.from_numpy(pr["kpA"]).to(device),
            "kB": torch.from_numpy(pr["kpB"]).to(device),
            "x": torch.empty((max(1, nA), 8), dtype=torch.float32, device=device),
        }
        b["kA2"] = b["kA"][:, :2].contiguous()
        b["kB2"] = b["kB"][:, :2].contiguous()
        b["pts"] = torch.empty((max(1, nA), 4), dtype=torch.float32, device=device)
        bufs.append(b)
    p = lambda t: t.data_ptr()  # noqa: E731
    top2 = _arr(BfTop2Task, [BfTop2Task(BfTask(b["nA"], b["nB"], p(pr["descA"]), p(pr["descB"]), p(b["bestA"]),
                                               p(b["bestB"]), None, None), p(b["secA"]))
                             for pr, b in zip(pairs, bufs)])
    need = ctypes.c_size_t(0)
    check(lib.rsfm_bf_match_top2_workspace_size(ctypes.cast(top2, ctypes.c_void_p), P, ctypes.byref(need)), "ws")
    ws = torch.empty(need.value + 256, dtype=torch.uint8, device=device)
    wsp = (ws.data_ptr() + 255) // 256 * 256 if need.value else 0
    check(lib.rsfm_bf_match_top2(ctypes.cast(top2, ctypes.c_void_p), P, ctypes.c_void_p(wsp), need.value, _sp()),
          "rsfm_bf_match_top2")
    filt = _arr(FilterTask, [FilterTask(p(b["bestA"]), None, b["nA"], p(b["bestB"]), b["nB"], p(b["out"]),
                                        p(b["count"]), p(b["kA2"]), p(b["kB2"]), p(b["pts"])) for b in bufs])
    check(lib.rsfm_filter_matches_batch(ctypes.cast(filt, ctypes.c_void_p), P, RATIO[0], RATIO[1], _sp()), "filter")
    if hard_outliers is not None:
        for pr, b in zip(pairs, bufs):
            c = int(b["count"].item())
            m = b["out"][:c].cpu().numpy()
            q, t = m[:, 0], m[:, 1]
            gt = (pr["src"][q] == t) & pr["geo"][q]
            if gt.sum() == 0 or (~gt).sum() == 0:
                continue
            kA, kB = pr["kpA"].copy(), pr["kpB"]
            qi, ti = q[gt], t[gt]
            pick = hard_outliers.integers(0, len(qi), int((~gt).sum()))
            ratio = kA[qi[pick], 2] / kB[ti[pick], 2]
            delta = kA[qi[pick], 3] - kB[ti[pick], 3]
            qo, to = q[~gt], t[~gt]
            kA[qo, 2] = kB[to, 2] * ratio
            kA[qo, 3] = np.angle(np.exp(1j * (kB[to, 3] + delta)))
            b["kA"].copy_(torch.from_numpy(kA.astype(np.float32)))
    feats = _arr(FeatureTask, [FeatureTask(p(b["out"]), p(b["count"]), b["nA"], p(b["secA"]), p(b["kA"]), p(b["kB"]),
                                           synth.WIDTH, synth.HEIGHT, p(b["x"])) for b in bufs])
    check(lib.rsfm_proposal_features_batch(ctypes.cast(feats, ctypes.c_void_p), P, _sp()), "features")
    counts = torch.stack([b["count"] for b in bufs]).cpu().numpy().ravel()
    res = []
    for b, c in zip(bufs, counts):
        res.append({"matches": b["out"][:c], "x": b["x"][:c], "pts": b["pts"][:c], "secA": b["secA"]})
    return res


def truth(pair, matches):
    """Ground truth per kept match: the query copies its train and sits on the geometry."""
    m = matches.cpu().numpy()
    q, t = m[:, 0], m[:, 1]
    return (pair["src"][q] == t) & pair["geo"][q]


def subsample(rng, gt, outlier_rate):
    """Indices of a subset with the given outlier rate (keeps as many matches as possible)."""
    inl = np.flatnonzero(gt)
    out = np.flatnonzero(~gt)
    n_in, n_out = len(inl), len(out)
    if n_in == 0 or n_out == 0:
        return np.arange(len(gt))
    want_out = int(round(n_in * outlier_rate / (1 - outlier_rate)))
    if want_out <= n_out:
        keep_out = rng.choice(out, want_out, replace=False)
        keep_in = inl
    else:
        keep_out = out
        keep_in = rng.choice(inl, max(8, int(round(n_out * (1 - outlier_rate) / outlier_rate))), replace=False)
    idx = np.concatenate([keep_in, keep_out])
    rng.shuffle(idx)
    return idx


def score(problems, threshold=1.0):
    """problems: list of (pts (n,4), idx (H,8)) -> per-hypothesis counts and best masks."""
    tasks, outs = [], []
    for pts, idx in problems:
        n, H = pts.shape[0], idx.shape[0]
        o = {"counts": torch.empty(H, dtype=torch.int32, device=pts.device),
             "best": torch.empty(2, dtype=torch.int32, device=pts.device),
             "mask": torch.empty(max(1, n), dtype=torch.uint8, device=pts.device)}
        tasks.append(RansacTask(pts.data_ptr(), n, None, idx.data_ptr(), H, o["counts"].data_ptr(), None,
                                o["best"].data_ptr(), o["mask"].data_ptr()))
        outs.append(o)
    arr = _arr(RansacTask, tasks)
    check(lib.rsfm_ransac_f8(ctypes.cast(arr, ctypes.c_void_p), len(tasks), float(threshold), 1, 0.0, 0.0, _sp()),
          "rsfm_ransac_f8")
    return outs


def uniform_samples(ns, H, seeds, device):
    idx = [torch.empty((H, 8), dtype=torch.int32, device=device) for _ in ns]
    arr = _arr(SampleTask, [SampleTask(i.data_ptr(), H, n, None, s) for i, n, s in zip(idx, ns, seeds)])
    check(lib.rsfm_ransac_sample_indices_batch(ctypes.cast(arr, ctypes.c_void_p), len(ns), _sp()), "sample")
    return idx


def make_split(seed, n_pairs, kpts, outlier_rates, device, copy_frac=0.5, geo_frac=0.8, batch=32,
               hard_outliers=False):
    """Pairs -> matched, featurised, subsampled to each pair's outlier rate.

    Returns a list of dicts: x (m,8), pts (m,4) device tensors, gt (m,) bool numpy.
    hard_outliers: outliers copy an inlier's scale ratio and orientation
    difference (match_pairs).
    """
    rng = np.random.default_rng(seed)
    hard = np.random.default_rng(seed + 1) if hard_outliers else None
    out = []
    for b0 in range(0, n_pairs, batch):
        pairs = []
        for i in range(b0, min(n_pairs, b0 + batch)):
            k = int(kpts[i]) if hasattr(kpts, "__len__") else int(kpts)
            pairs.append(synth.make_pair(seed * 1000003 + i, k, k, copy_frac, geo_frac, device=device))
        res = match_pairs(pairs, device, hard)
        for i, (pr, r) in enumerate(zip(pairs, res)):
            gt = truth(pr, r["matches"])
            rate = outlier_rates[b0 + i] if hasattr(outlier_rates, "__len__") else outlier_rates
            sel = subsample(rng, gt, rate)
            selt = torch.from_numpy(sel).to(device)
            out.append({"x": r["x"].index_select(0, selt).contiguous(),
                        "pts": r["pts"].index_select(0, selt).contiguous(), "gt": gt[sel]})
    return out


def label_with_plain_ransac(split, H=4096, seed=5, batch=32):
    """Plain RANSAC labels: uniform draws among the ratio-test survivors, every
    match scored; the inlier mask of the best hypothesis is the label."""
    for b0 in range(0, len(split), batch):
        part = split[b0:b0 + batch]
        tasks = []
        for j, s in enumerate(part):
            ratio = s["x"][:, 0]
            w = (ratio * RATIO[1] < RATIO[0]).float()  # sqrt(d1/d2) < 0.8
            tasks.append((w.contiguous(), H, seed + b0 + j))
        idx = learned.sample_weighted(tasks)
        outs = score([(s["pts"], i) for s, i in zip(part, idx)])
        for s, o in zip(part, outs):
            s["y"] = o["mask"][:s["pts"].shape[0]].float()
    return split


def zero_features(split, cols):
    """A copy of the split with the feature columns cols set to zero (ablation)."""
    out = []
    for s in split:
        x = s["x"].clone()
        x[:, list(cols)] = 0.0
        out.append(dict(s, x=x))
    return out


def train(split, epochs=4, batch=8192, lr=3e-3, seed=1, log=None, curve=None):
    """Adam on the whole split; returns (params, per-epoch loss on the first 65536
    samples). curve: optional list that receives (step, loss) every 64 steps and
    the loss of the initial network at step 0."""
    dev = split[0]["x"].device
    X = torch.cat([s["x"] for s in split]).contiguous()
    Y = torch.cat([s["y"] for s in split]).contiguous()
    params = learned.init_params(seed, device=dev)
    opt = learned.Adam(params, lr=lr)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    n = X.shape[0]
    losses = []
    step = 0
    if curve is not None:
        curve.append((0, learned.grad(params, X[:65536].contiguous(), Y[:65536].contiguous())[1].item()))
    for ep in range(epochs):
        perm = torch.randperm(n, generator=gen, device=dev)
        for b0 in range(0, n - batch + 1, batch):
            sel = perm[b0:b0 + batch]
            g, loss = learned.grad(params, X[sel].contiguous(), Y[sel].contiguous())
            opt.step(g)
            step += 1
            if curve is not None and step % 64 == 0:
                curve.append((step, loss.item()))
        losses.append(learned.grad(params, X[:65536].contiguous(), Y[:65536].contiguous())[1].item())
        if log:
            log("epoch %d loss %.4f" % (ep, losses[-1]))
    return params, losses


def h99(q):
    if q <= 0:
        return math.inf
    if q >= 1:
        return 1
    return math.ceil(math.log(0.01) / math.log(1 - q))


def evaluate(params, split, H_plain=65536, H_guided=4096, seed=9, batch=16, floor=0.02):
    """Guided vs uniform sampling on held-out pairs; returns a summary dict.

    Per pair, an independent plain run of H_plain uniform hypotheses fixes the
    plain method's inlier count C; a hypothesis succeeds if its count is >=
    ceil(0.99 C). A second plain run (other seed, H_plain) and the guided run
    (H_guided) give the success rates. Guided weights: network probability +
    floor (every match keeps a small chance).
    """
    succ_p = succ_g = 0
    hyp_p = hyp_g = 0
    per_pair = []
    reach = 0
    dev = split[0]["pts"].device
    for b0 in range(0, len(split), batch):
        part = split[b0:b0 + batch]
        ns = [s["pts"].shape[0] for s in part]
        idx_ref = uniform_samples(ns, H_plain, [seed + b0 + j for j in range(len(part))], dev)
        idx_p = uniform_samples(ns, H_plain, [seed + 104729 + b0 + j for j in range(len(part))], dev)
        probs = [learned.forward(params, s["x"]) for s in part]
        idx_g = learned.sample_weighted([((p + floor).contiguous(), H_guided, seed + 7919 + b0 + j)
                                         for j, p in enumerate(probs)])
        out_r = score([(s["pts"], i) for s, i in zip(part, idx_ref)])
        out_p = score([(s["pts"], i) for s, i in zip(part, idx_p)])
        out_g = score([(s["pts"], i) for s, i in zip(part, idx_g)])
        for orf, op, og in zip(out_r, out_p, out_g):
            cref = int(orf["best"][1].item())
            cp = op["counts"].cpu().numpy()
            cg = og["counts"].cpu().numpy()
            thr = math.ceil(0.99 * cref)
            sp, sg = int((cp >= thr).sum()), int((cg >= thr).sum())
            succ_p += sp
            succ_g += sg
            hyp_p += len(cp)
            hyp_g += len(cg)
            reach += int(cg.max() >= thr)
            per_pair.append((sp / len(cp), sg / len(cg)))
    qp, qg = succ_p / hyp_p, succ_g / hyp_g
    ratios = [h99(a) / h99(b) for a, b in per_pair if a > 0 and b > 0]
    return {
        "pairs": len(split),
        "q_plain": qp,
        "q_guided": qg,
        "h99_plain": h99(qp),
        "h99_guided": h99(qg),
        "gain": h99(qp) / h99(qg) if qg > 0 else 0.0,
        "gain_median_pair": float(np.median(ratios)) if ratios else None,
        "pairs_with_both": len(ratios),
        "guided_reaches_99pct": reach / len(split),
    }
