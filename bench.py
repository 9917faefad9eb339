"""Two-view matching throughput: pairs/s at 8K keypoints (match + RANSAC) on one MI355X.

One step = one pass of the hot path over a batch of image pairs already
resident in HBM, in groups of 32 pairs (one matcher launch each):
  1. rsfm_bf_match_top2           int8-MFMA brute force, 8192 x 8192 per pair
                                  (top-2 query side, top-1 train side)
  2. rsfm_filter_matches_batch    ratio test 0.8 + cross-check + keypoint gather
  3. rsfm_ransac_sample_indices_batch   H seeded 8-point samples per pair
  4. rsfm_ransac_f8               FP32 8-point fits + Sampson scoring (1 px),
                                  every count and model stored
  5. rsfm_ransac_f8_finish        best hypothesis, FP64 least-squares refit of
                                  F on its inliers (kept when it has more),
                                  final F, result and inlier mask per pair
The matcher calls run back to back on the main stream; each group's steps
2-5 run on a second stream (default priority) once its matches are written,
under the next group's matcher (--streams 1: everything on one stream).
Synthetic data (rapidsfm_amd.synth, drawn on the device): 64 distinct pairs of
SIFT-like descriptors (30% noisy true matches) with keypoints from one
two-view geometry per pair (30% of true matches geometric outliers).

Prints one JSON line (metric, value, ..., roofline, cpu_baseline).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_KPTS = 8192
INT8_PEAK_TOPS = 5000.0  # MI355X dense int8 MFMA, 5.0 POPS
MATCH_KERNEL = "bf::match_top2_kernel"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--pairs", type=int, default=64, help="image pairs per step")
    ap.add_argument("--streams", type=int, default=2, choices=(1, 2),
                    help="2: each 32-pair group's filter and RANSAC on a second stream under the next matcher call")
    ap.add_argument("--no-other", action="store_true",
                    help="skip the timed pass in the other stream mode (for profiling one mode)")
    ap.add_argument("--hyp", type=int, default=4096, help="RANSAC hypotheses per pair")
    ap.add_argument("--n", type=int, default=N_KPTS, help="keypoints per image")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="bound of the all-core CPU baseline sample")
    ap.add_argument("--cpu-seconds-1core", type=float, default=8.0, help="bound of the one-thread CPU sample")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one 256 MiB device copy (HBM-counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-pipeline leg")
    ap.add_argument("--batched-pairs", type=int, default=256)
    ap.add_argument("--no-degensac", action="store_true", help="skip the batched DEGENSAC + pose run")
    ap.add_argument("--no-cashash", action="store_true", help="skip the cascade-hashing leg")
    ap.add_argument("--no-aux", action="store_true", help="skip the top-1 / 4-bit / IDP matcher and affine-filter leg")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal leg")
    ap.add_argument("--learned-train", type=int, default=10000, help="training pairs (1K-2K keypoints)")
    ap.add_argument("--learned-eval", type=int, default=1024, help="held-out 8K-keypoint pairs at 70%% outliers")
    return ap.parse_args()


def batched_leg(args, dev, log):
    """256 pairs, keypoints U[1K,16K] per image, geometric outlier rate U[30%,80%]
    among the true descriptor matches, one HIP-graph launch per batch."""
    import numpy as np
    import torch

    from rapidsfm_amd import synth
    from rapidsfm_amd.pipeline import Pipeline

    rng = np.random.default_rng(2024)
    P = args.batched_pairs
    pairs, truth = [], []
    for i in range(P):
        nA, nB = (int(v) for v in rng.integers(1024, 16385, 2))
        rate = float(rng.uniform(0.3, 0.8))
        pr = synth.make_pair(5000 + i, nA, nB, copy_frac=0.5, geo_frac=1.0 - rate, device=dev)
        kA = torch.from_numpy(np.ascontiguousarray(pr["kpA"][:, :2])).to(dev)
        kB = torch.from_numpy(np.ascontiguousarray(pr["kpB"][:, :2])).to(dev)
        pairs.append((pr["descA"], pr["descB"], kA, kB))
        truth.append((pr["src"], pr["geo"], pr["R"]))
    t0 = time.perf_counter()
    pl = Pipeline(pairs, n_hyp=args.hyp, threshold=1.0, mode=1, seed=77, refit=True)
    t_create = time.perf_counter() - t0
    for _ in range(2):
        pl.run()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = max(3, args.steps // 2)
    e0.record()
    for _ in range(reps):
        pl.run()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    # realised outlier rate of the kept matches (ground truth of the generator)
    kept = out = 0
    for (src, geo, _), o in zip(truth, pl.out):
        c = int(o["count"].item())
        m = o["matches"][:c].cpu().numpy()
        ok = (src[m[:, 0]] == m[:, 1]) & geo[m[:, 0]]
        kept += c
        out += int((~ok).sum())
    info = pl.info()
    ops = sum(2.0 * a.shape[0] * b.shape[0] * 128 for a, b, _, _ in pairs)
    pl.close()
    log("batched: %.3f ms per batch of %d" % (ms, P))
    dg = None if args.no_degensac else degensac_batch(args, pairs, truth, info, log)
    return {
        "workload": "%d pairs, keypoints U[1024,16384] per image, top2+ratio0.8+crosscheck, %d-hypothesis "
                    "F-RANSAC (Sampson 1px) + refit on the best sample's inliers per pair, one HIP graph launch"
                    % (P, args.hyp),
        "value": round(P / (ms / 1e3), 1),
        "unit": "pairs/s",
        "ms_per_batch": round(ms, 4),
        "match_tops_equiv": round(ops / (ms / 1e3) / 1e12, 1),
        "kept_matches": kept,
        "outlier_rate_kept": round(out / max(1, kept), 4),
        "tiling": info,
        "create_s": round(t_create, 2),
        "degensac": dg,
    }


def degensac_batch(args, pairs, truth, info, log):
    """The same batch through the reference's default pair solver in one graph:
    7-point samples, DEGENSAC finisher (F and H), relative pose from F and H.
    Thresholds as the reference derives them for 4000 x 3000 images (two-sided
    line test 0.008 x 3000 px, homography 0.005 x 3000 px, Config.h:172-173,
    degensac.cpp:119-120); the matcher tiling is the one tuned above."""
    import numpy as np
    import torch

    from rapidsfm_amd import synth
    from rapidsfm_amd.pipeline import Pipeline

    cam = (synth.FOCAL, synth.FOCAL, synth.WIDTH / 2, synth.HEIGHT / 2)
    P = len(pairs)
    dim = min(synth.WIDTH, synth.HEIGHT)
    t0 = time.perf_counter()
    pl = Pipeline(pairs, n_hyp=args.hyp, threshold=0.008 * dim, mode=0, seed=77, solver="degensac",
                  threshold_h=0.005 * dim, n_candidates=8, n_plane_samples=256, front_ratio=0.95,
                  cams=[(cam, cam)] * P, seg_cols=info["segCols"], wg_per_xcd=info["wgPerXcd"])
    t_create = time.perf_counter() - t0
    for _ in range(2):
        pl.run()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = max(3, args.steps // 2)
    e0.record()
    for _ in range(reps):
        pl.run()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    # pose quality: rotation error of the leading kept F pose against the generator's pose
    errs, kept_one = [], 0
    for (_, _, R), o in zip(truth, pl.out):
        res = o["poseF"].cpu().numpy()
        if res[8] < 1:
            errs.append(180.0)
            continue
        kept_one += int(res[8] == 1)
        Re = o["posesF"][int(res[4])].cpu().numpy()[:9].reshape(3, 3).astype(np.float64)
        c = np.clip((np.trace(Re @ R.T) - 1.0) / 2.0, -1.0, 1.0)
        errs.append(float(np.degrees(np.arccos(c))))
    errs = np.array(errs)
    pl.close()
    log("batched DEGENSAC + pose: %.3f ms per batch of %d" % (ms, P))
    return {
        "workload": "same batch: 7-point samples (%d per pair), DEGENSAC finisher (F and H, 8 candidates, "
                    "256 plane-and-parallax samples), relative pose from F and H; one HIP graph launch" % args.hyp,
        "value": round(P / (ms / 1e3), 1),
        "unit": "pairs/s",
        "ms_per_batch": round(ms, 4),
        "pose_rot_err_deg_median": round(float(np.median(errs)), 4),
        "pose_within_1deg": int((errs < 1.0).sum()),
        "pose_unique": kept_one,
        "pairs": P,
        "create_s": round(t_create, 2),
    }


def aux_leg(args, dev, log):
    """The other matchers and the affine match filter at the benchmark's size:
    top-1 bidirectional int8 MFMA (rsfm_bf_match_u8), 4-bit (rsfm_bf_match_u4,
    descriptors halved first), packed-dot VALU (rsfm_bf_match_idp_u8), each on
    distinct 8192 x 8192 pairs, and the batched affine filter
    (rsfm_rmf_filter_batch) on 64 pairs of 2500 correspondences (60% on the
    pair's geometry, the rest uniform over the image)."""
    import numpy as np
    import torch

    import rapidsfm_amd as rs
    from rapidsfm_amd import synth
    from rapidsfm_amd.rmf import filter_batch

    n = args.n

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    out = {}
    base = [synth.make_pair(9000 + i, n, n, 0.3, 0.7, device=dev) for i in range(32)]
    pairs = [(p["descA"], p["descB"]) for p in base]
    ops = 2.0 * n * n * 128
    for impl, P, reps in (("mfma", 32, 5), ("u4", 32, 5), ("idp", 4, 2)):
        pp = pairs[:P]
        if impl == "u4":
            pp = [(rs.desc8b_to_4b(a), rs.desc8b_to_4b(b)) for a, b in pp]
        ms = timed(lambda: rs.bf_match(pp, impl=impl), reps)
        out["bf_" + impl] = {"pairs_per_call": P, "ms_per_call": round(ms, 4),
                             "ms_per_pair": round(ms / P, 4),
                             "tops": round(ops * P / (ms / 1e3) / 1e12, 1)}
        log("aux %s: %.3f ms per %d pairs" % (impl, ms, P))
    rng = np.random.default_rng(31)
    probs = []
    for i in range(64):
        p = base[i % 32]
        idx = np.nonzero(p["geo"])[0][:1500]
        a = p["kpA"][idx, :2]
        b = p["kpB"][p["src"][idx], :2]
        k = 2500 - len(idx)
        ra = rng.uniform((0, 0), (synth.WIDTH, synth.HEIGHT), (k, 2)).astype(np.float32)
        rb = rng.uniform((0, 0), (synth.WIDTH, synth.HEIGHT), (k, 2)).astype(np.float32)
        pts = np.concatenate([np.concatenate([a, b], 1), np.concatenate([ra, rb], 1)]).astype(np.float32)
        probs.append(torch.from_numpy(np.ascontiguousarray(pts)).to(dev))
    cnt = [torch.zeros(1, dtype=torch.int32, device=dev) for _ in probs]
    work = [x.clone() for x in probs]

    def rmf_once():
        for w, x, c in zip(work, probs, cnt):
            w.copy_(x)
            c.fill_(x.shape[0])
        filter_batch([(w, c, (int(synth.WIDTH), int(synth.HEIGHT)), 5 + j) for j, (w, c) in enumerate(zip(work, cnt))])

    ms = timed(rmf_once, 5)
    kept = sum(int(c.item()) for c in cnt)
    out["rmf_batch"] = {"pairs_per_call": len(probs), "correspondences_per_pair": 2500, "ms_per_call": round(ms, 4),
                        "ms_per_pair": round(ms / len(probs), 4), "kept_per_pair": round(kept / len(probs), 1)}
    log("aux rmf: %.3f ms per %d pairs" % (ms, len(probs)))
    return out


def cashash_leg(args, dev, log):
    """Cascade-hash matching of 8192 x 8192 pairs (30% noisy true matches), 192
    projection channels (6 bucket tables of 8-bit codes, 128 remap bits),
    everything on the device per pair: descriptor accumulation and bias, hash +
    histogram, bucket bounds, buckets for both images, then the bidirectional
    best match. Two match modes over the same registrations: config 1's re-rank
    (the 8 Hamming-nearest of the bucket union by 128-bit remap, then exact L2^2;
    the leg's `value`) and the reference's own exact L2^2 over the whole bucket
    union (`union`). Each mode is one HIP graph per 8 pairs, its replays checked
    against eager results."""
    import torch
    from rapidsfm_amd import cashash as ch
    from rapidsfm_amd.synth import cashash_weights, make_pair

    n, pairs, steps, keep = 8192, 8, 10, 8
    w, base = cashash_weights(3, 192, dev)
    sets = [make_pair(40 + k, n, n, 0.3, 1.0, dev) for k in range(pairs)]
    stream = torch.cuda.current_stream()

    def one(p, rerank):
        q, t = p["descA"], p["descB"]
        acc = ch.accumulate(q)
        bias_q = ch.compute_bias(w, base, acc, n)
        acc = ch.accumulate(t, acc=acc, init=False)
        bias_t = ch.compute_bias(w, base, acc, 2 * n)
        qi = ch.register_image(q, w, bias_q)
        ti = ch.register_image(t, w, bias_t)
        if rerank:
            return ch.find_best_match_rerank([qi], ti, 6, keep)[0]
        return ch.find_best_match([qi], ti, 6)[0]

    def measure(rerank):
        # eager reference results, then one HIP graph of the whole batch of pairs
        side = torch.cuda.Stream()
        side.wait_stream(stream)
        with torch.cuda.stream(side):
            eager = [one(p, rerank) for p in sets]
        stream.wait_stream(side)
        torch.cuda.synchronize()
        ref = [(q.cpu(), t.cpu()) for q, t in eager]
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            outs = [one(p, rerank) for p in sets]
        graphed = True
        for rep in range(3):  # every replay must reproduce the eager results bit for bit
            graph.replay()
            torch.cuda.synchronize()
            for (q, t), (rq, rt) in zip(outs, ref):
                if not (torch.equal(q.cpu(), rq) and torch.equal(t.cpu(), rt)):
                    graphed = False
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            if graphed:
                graph.replay()
            else:
                for p in sets:
                    one(p, rerank)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / (steps * pairs)
        last = ref[-1]
        src = torch.from_numpy(sets[-1]["src"])
        hits = int(((last[0][:, 0] == src) & (src >= 0)).sum())
        log("cashash (%s): %.3f ms per 8K x 8K pair (%s)" % ("re-rank" if rerank else "union", ms,
                                                              "HIP graph" if graphed else "eager"))
        return {
            "value": round(1000.0 / ms, 1),
            "unit": "pairs/s",
            "ms_per_pair": round(ms, 4),
            "launch": "one HIP graph per %d pairs, replays checked against eager results" % pairs if graphed
                      else "eager (graph replay differed from eager results)",
            "true_matches_found_last_pair": hits,
            "true_matches_last_pair": int((sets[-1]["src"] >= 0).sum()),
        }

    out = measure(True)
    out["workload"] = ("cascade-hash match 8192 x 8192 (30%% noisy true matches), 6 bucket tables + 128 remap bits, "
                       "re-rank of the %d Hamming-nearest by exact L2^2; registration of both images + "
                       "bidirectional best match per pair" % keep)
    uni = measure(False)
    uni["workload"] = "same pairs and registrations, exact L2^2 over the whole union of the 6 tables' buckets"
    out["union"] = uni
    return out


def learned_leg(args, dev, log):
    """Learned proposals: train on plain-RANSAC labels, compare guided vs uniform
    sampling on held-out 8K pairs at 70% outliers (plain scoring judges both).
    Ablations: the same data with the scale-ratio and orientation-difference
    features (f6, f7) zeroed, and a generator variant whose outliers copy an
    inlier's scale ratio and orientation difference (no such cue at all)."""
    import numpy as np

    from rapidsfm_amd import proposal

    def run(hard, zero, tag):
        t0 = time.perf_counter()
        rng = np.random.default_rng(31)
        n = args.learned_train
        train = proposal.make_split(17, n, rng.integers(1024, 2049, n), rng.uniform(0.5, 0.9, n), dev,
                                    hard_outliers=hard)
        proposal.label_with_plain_ransac(train, H=4096)
        t_data = time.perf_counter() - t0
        label_acc = float(np.mean(np.concatenate([(s["y"].cpu().numpy() > 0.5) == s["gt"] for s in train])))
        n_samples = sum(int(s["x"].shape[0]) for s in train)
        held = proposal.make_split(7777, args.learned_eval, 8192, 0.7, dev, hard_outliers=hard)
        reps = {}
        for z in ([False, True] if zero else [False]):
            tr = proposal.zero_features(train, (6, 7)) if z else train
            he = proposal.zero_features(held, (6, 7)) if z else held
            t1 = time.perf_counter()
            curve = []
            params, losses = proposal.train(tr, epochs=4, batch=8192, curve=curve)
            t_train = time.perf_counter() - t1
            rep = proposal.evaluate(params, he, H_plain=65536, H_guided=4096)
            rep.update({
                "train_pairs": n,
                "train_samples": n_samples,
                "label_agreement_with_truth": round(label_acc, 4),
                "train_loss_epochs": [round(v, 4) for v in losses],
                "train_loss_curve": [(k, round(v, 4)) for k, v in curve[:: max(1, len(curve) // 12)]],
                "train_s": round(t_train, 2),
                "data_s": round(t_data, 2),
            })
            key = tag + ("_no_scale_orientation" if z else "")
            reps[key] = rep
            log("learned %s: gain %.2f (h99 plain %s, guided %s)" % (key, rep["gain"], rep["h99_plain"],
                                                                    rep["h99_guided"]))
        return reps

    reps = run(False, True, "default")
    reps.update(run(True, False, "outliers_share_scale_orientation"))
    out = dict(reps["default"])
    out.update({
        "held_out": "%d pairs, 8192 keypoints per image, 70%% outliers among the kept matches" % args.learned_eval,
        "h_plain": 65536,
        "h_guided": 4096,
        "ablations": {k: {q: v[q] for q in ("gain", "gain_median_pair", "h99_plain", "h99_guided",
                                            "guided_reaches_99pct", "q_plain", "q_guided", "train_loss_epochs")}
                      for k, v in reps.items() if k != "default"},
    })
    return out


def measured_traffic(n, per_launch):
    """HBM bytes per matcher launch from the committed counter profile
    (profiles/match_8k_hbm_counters.md, two rocprofv3 --pmc passes, calibrated),
    when it was taken for this kernel at this shape; else None."""
    path = os.path.join(ROOT, "profiles", "match_8k_hbm_counters.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                f = dict(kv.split("=") for kv in line.split(":", 1)[1].split())
                if (f.get("kernel") == MATCH_KERNEL and int(f["keypoints"]) == n and
                        int(f["pairs_per_launch"]) == per_launch):
                    return int(f["bytes"])
    except (OSError, ValueError, KeyError):
        pass
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (image pairs are independent and "
              "would shard across GPUs without communication). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import rapidsfm_amd as rs  # noqa: F401  (fails loudly without the HIP library)
    from rapidsfm_amd._abi import BfTask, BfTop2Task, FilterTask, FinishTask, RansacTask, SampleTask, check, lib

    def log(msg):
        print("[bench] " + msg, file=sys.stderr, flush=True)

    dev = torch.device("cuda:0")
    n, P, H = args.n, args.pairs, args.hyp
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)

    # ---- synthetic pairs: P distinct pairs, generated on the device, resident in HBM
    from rapidsfm_amd import synth

    t_gen = time.perf_counter()
    src = [synth.make_pair(1000 + p, n, n, copy_frac=0.3, geo_frac=0.7, device=dev) for p in range(P)]
    bufs = []
    for u in src:
        b = {
            "dA": u["descA"], "dB": u["descB"],
            "kA": torch.from_numpy(np.ascontiguousarray(u["kpA"][:, :2])).to(dev),
            "kB": torch.from_numpy(np.ascontiguousarray(u["kpB"][:, :2])).to(dev),
            "bestA": torch.empty((n, 2), dtype=torch.int32, device=dev),
            "secA": torch.empty((n, 2), dtype=torch.int32, device=dev),
            "bestB": torch.empty((n, 2), dtype=torch.int32, device=dev),
            "pts": torch.empty((n, 4), dtype=torch.float32, device=dev),
            "count": torch.zeros(1, dtype=torch.int32, device=dev),
            "idx": torch.empty((H, 8), dtype=torch.int32, device=dev),
            "counts": torch.empty(H, dtype=torch.int32, device=dev),
            "models": torch.empty((H, 9), dtype=torch.float32, device=dev),
            "F": torch.empty(9, dtype=torch.float32, device=dev),
            "result": torch.empty(4, dtype=torch.int32, device=dev),
            "mask": torch.empty(n, dtype=torch.uint8, device=dev),
        }
        bufs.append(b)
    torch.cuda.synchronize()
    log("generated %d pairs in %.1f s" % (P, time.perf_counter() - t_gen))
    ptr = lambda t: t.data_ptr()  # noqa: E731
    top2 = (BfTop2Task * P)()
    filt = (FilterTask * P)()
    samp = (SampleTask * P)()
    rtask = (RansacTask * P)()
    fin = (FinishTask * P)()
    for p, b in enumerate(bufs):
        top2[p] = BfTop2Task(BfTask(n, n, ptr(b["dA"]), ptr(b["dB"]), ptr(b["bestA"]), ptr(b["bestB"]), None, None),
                             ptr(b["secA"]))
        filt[p] = FilterTask(ptr(b["bestA"]), ptr(b["secA"]), n, ptr(b["bestB"]), n, None, ptr(b["count"]),
                             ptr(b["kA"]), ptr(b["kB"]), ptr(b["pts"]))
        samp[p] = SampleTask(ptr(b["idx"]), H, n, ptr(b["count"]), 77 + p)
        # the scorer stores every hypothesis' count and model; the finisher picks
        # the best, refits F on its inliers and writes F, result and the mask
        rtask[p] = RansacTask(ptr(b["pts"]), n, ptr(b["count"]), ptr(b["idx"]), H, ptr(b["counts"]),
                              ptr(b["models"]), None, None)
        fin[p] = FinishTask(ptr(b["pts"]), n, ptr(b["count"]), ptr(b["idx"]), ptr(b["counts"]), ptr(b["models"]), H,
                            ptr(b["F"]), ptr(b["result"]), ptr(b["mask"]))
    ws_need = ctypes.c_size_t(0)
    check(lib.rsfm_bf_match_top2_workspace_size(ctypes.cast(top2, ctypes.c_void_p), P, ctypes.byref(ws_need)), "ws")
    ws = torch.empty(ws_need.value + 256, dtype=torch.uint8, device=dev)
    ws_ptr = (ws.data_ptr() + 255) // 256 * 256 if ws_need.value else 0

    side = torch.cuda.Stream(device=dev)  # default priority
    ssp = ctypes.c_void_p(side.cuda_stream)
    groups = [(g, min(32, P - g)) for g in range(0, P, 32)]
    at = lambda arr, cls, g: ctypes.c_void_p(ctypes.addressof(arr) + g * ctypes.sizeof(cls))  # noqa: E731
    ev_m0 = [[torch.cuda.Event(enable_timing=True) for _ in groups] for _ in range(args.steps)]
    ev_m1 = [[torch.cuda.Event(enable_timing=True) for _ in groups] for _ in range(args.steps)]
    ev_matched = [torch.cuda.Event() for _ in groups]
    ev_filtered = [torch.cuda.Event() for _ in groups]

    def step(i=None, streams=args.streams):
        for gi, (g, gn) in enumerate(groups):
            # the group's match results are rewritten only after its previous filter read them
            if streams == 2:
                stream.wait_event(ev_filtered[gi])
            if i is not None:
                ev_m0[i][gi].record(stream)
            check(lib.rsfm_bf_match_top2(at(top2, BfTop2Task, g), gn, ctypes.c_void_p(ws_ptr), ws_need.value, sp),
                  "match")
            if i is not None:
                ev_m1[i][gi].record(stream)
            rs = sp
            if streams == 2:
                ev_matched[gi].record(stream)
                side.wait_event(ev_matched[gi])
                rs = ssp
            check(lib.rsfm_filter_matches_batch(at(filt, FilterTask, g), gn, 4, 5, rs), "filter")
            if streams == 2:
                ev_filtered[gi].record(side)
            check(lib.rsfm_ransac_sample_indices_batch(at(samp, SampleTask, g), gn, rs), "sample")
            check(lib.rsfm_ransac_f8(at(rtask, RansacTask, g), gn, 1.0, 1, 0.0, 0.0, rs), "ransac")
            check(lib.rsfm_ransac_f8_finish(at(fin, FinishTask, g), gn, 1.0, 1, 0.0, 5000, 1, rs), "finish")

    def timed(streams):
        for _ in range(args.warmup):
            step(None, streams)
        torch.cuda.synchronize()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record(stream)
        for i in range(args.steps):
            step(i, streams)
        stream.wait_stream(side)  # the last group's RANSAC is inside the timed region
        t1.record(stream)
        torch.cuda.synchronize()
        ms_total = t0.elapsed_time(t1)
        match_ms = sum(a.elapsed_time(b) for ea, eb in zip(ev_m0, ev_m1) for a, b in zip(ea, eb))
        return ms_total, match_ms / (args.steps * len(groups))

    # the other mode first (same steps, same buffers), then the reported one
    other = 3 - args.streams
    ms_total_o, match_ms_o = timed(other) if not args.no_other else (None, None)
    ms_total, match_ms = timed(args.streams)
    # the roofline is the matcher's own: its calls in the one-stream pass, where
    # no other kernel shares the GPU with it (in the two-stream pass the scorer
    # of the previous group runs beside it and each call lasts longer)
    rl_other = args.streams == 2 and match_ms_o is not None
    rl_ms = match_ms_o if rl_other else match_ms
    ms_step = ms_total / args.steps
    value = P * args.steps / (ms_total / 1e3)
    # the matcher takes at most 32 pairs per kernel launch: report per launch
    launches = len(groups)
    per_launch = min(P, 32)
    ops = 2.0 * n * n * 128 * P / launches
    achieved = ops / (rl_ms / 1e3) / 1e12

    cnts = np.array([int(b["count"].item()) for b in bufs])
    res = np.stack([b["result"].cpu().numpy().view(np.uint32) for b in bufs])
    result = {
        "metric": "pairs/s at 8K kpts (match+RANSAC)",
        "value": round(value, 1),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8 (match, int32 accumulate) + fp32 (RANSAC) + fp64 (refit)",
        "data": "synthetic",
        "config": {
            "workload": "bf-match top2+ratio0.8+crosscheck, then F 8-point RANSAC (Sampson 1px, fixed table of "
                        "%d seeded hypotheses) + refit on the best sample's inliers, per pair" % H,
            "keypoints_per_image": n,
            "pairs_per_step": P,
            "distinct_pairs": P,
            "ransac_hypotheses": H,
            "matches_per_pair_median": int(np.median(cnts)),
            "refit_kept_pairs": int((res[:, 3] > res[:, 1]).sum()),
            "parallelism": "single GPU",
        },
        "parity_pair0": None,
        "roofline": {
            "kernel": MATCH_KERNEL + " (rsfm_bf_match_top2)",
            "bound": "mfma",
            "achieved": round(achieved, 1),
            "peak": INT8_PEAK_TOPS,
            "unit": "TOP/s",
            "frac": round(achieved / INT8_PEAK_TOPS, 4),
            "traffic": measured_traffic(n, per_launch),
            "traffic_unit": "HBM bytes per launch (rocprofv3 FETCH_SIZE / WRITE_SIZE passes, calibrated; "
                            "profiles/match_8k_hbm_counters.md)",
            "algorithmic_bytes": per_launch * (2 * n * 128 + 3 * n * 8),
            "avg_launch_ms": round(rl_ms, 4),
            "timed": "HIP events around each rsfm_bf_match_top2 call on the launch stream (preparation kernel "
                     "included), per launch of %d pairs, over the %d timed steps of the %s" % (
                         per_launch, args.steps, "one-stream pass (the matcher alone on the GPU)"
                         if rl_other else "reported pass"),
            "avg_launch_ms_in_reported_steps": round(match_ms, 4),
        },
        "streams": args.streams,
        "other_mode": None if ms_total_o is None else {
            "streams": other,
            "value": round(P * args.steps / (ms_total_o / 1e3), 1),
            "ms_per_step": round(ms_total_o / args.steps, 4),
            "note": "the same steps timed in the other stream mode, before the reported pass",
        },
        "cpu_baseline": None,
    }

    if args.calib_copy:
        # one elementwise kernel that reads 256 MiB and writes 256 MiB: the
        # FETCH_SIZE / WRITE_SIZE readings of this dispatch calibrate the counters
        srcb = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(srcb)
        torch.add(srcb, 1.0, out=dst)
        torch.cuda.synchronize()

    if not args.no_cpu:
        # ---- parity of pair 0 of the measured pass, and the CPU baseline: the
        # oracle's restatement of the same pass on the host cores
        from oracle import cpu_baseline as cb
        from oracle import pyoracle as orc

        host = [(u["descA"].cpu().numpy(), u["descB"].cpu().numpy(), u["kpA"][:, :2], u["kpB"][:, :2])
                for u in src[:8]]
        b0 = bufs[0]
        F, r0, m0 = cb.one_pair(orc, *host[0], H, 77, 0)
        ra, sa, rb = orc.bf_match_top2_fast(host[0][0], host[0][1])
        u32 = lambda t: t.cpu().numpy().view(np.uint32)  # noqa: E731
        cnt = int(cnts[0])
        result["parity_pair0"] = bool(
            np.array_equal(u32(b0["bestA"]), ra) and np.array_equal(u32(b0["secA"]), sa)
            and np.array_equal(u32(b0["bestB"]), rb) and cnt == len(m0)
            and np.array_equal(b0["F"].cpu().numpy().view(np.uint32), F.reshape(9).view(np.uint32))
            and np.array_equal(u32(b0["result"]), r0) and np.array_equal(b0["mask"][:cnt].cpu().numpy(), m0))
        native, how = cb.load_native()
        cores = len(os.sched_getaffinity(0))
        v_all, k_all, t_all = cb.measure(native, host, H, args.cpu_seconds, cores)
        v_one, k_one, t_one = cb.measure(native, host, H, args.cpu_seconds_1core, 1)
        result["cpu_baseline"] = {
            "value": round(v_all, 3),
            "unit": "pairs/s",
            "cores": cores,
            "kind": "port",
            "sample": "%d pairs (8 distinct) of the same workload (%dx%d top-2 match + filter + %d-hypothesis "
                      "F-RANSAC + refit), %.1f s, OpenMP over %d threads" % (k_all, n, n, H, t_all, cores),
            "build": how,
            "cpu_model": cb.cpu_model(),
            "value_1core": round(v_one, 3),
            "sample_1core": "%d pairs, %.1f s, one thread" % (k_one, t_one),
            "gpu_over_cpu": round(value / v_all, 1),
        }
        log("cpu baseline: %.2f pairs/s on %d threads, %.3f on one (%s)" % (v_all, cores, v_one, how))

    if not args.no_batched:
        try:
            result["batched"] = batched_leg(args, dev, log)
        except Exception as e:  # report, keep the main line
            result["batched"] = {"error": repr(e)}
    if not args.no_cashash:
        try:
            result["cashash"] = cashash_leg(args, dev, log)
        except Exception as e:
            result["cashash"] = {"error": repr(e)}
    if not args.no_aux:
        try:
            result["aux"] = aux_leg(args, dev, log)
        except Exception as e:
            result["aux"] = {"error": repr(e)}
    if not args.no_learned:
        try:
            result["learned"] = learned_leg(args, dev, log)
        except Exception as e:
            result["learned"] = {"error": repr(e)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
