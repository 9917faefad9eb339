"""CPU baseline of bench.py: the oracle's restatement of the measured pass
(top-2 match of 8192 x 8192 u8 descriptors, ratio test + cross-check + keypoint
gather, seeded 8-point samples, FP32 scoring, refit on the best sample's
inliers) timed on the host cores.

TEST INFRASTRUCTURE ONLY: used by bench.py's cpu_baseline leg. The library is
compiled from oracle/*.c on the measuring host itself with -O3 -march=native
(make target `native`), into a temporary directory, so the build matches that
host's CPU; when the build fails the oracle's portable build (x86-64-v3) is used
and the report says so.
"""
import importlib.util
import os
import subprocess
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def load_native():
    """(oracle module bound to a -march=native build, build description)"""
    out = os.path.join(tempfile.mkdtemp(prefix="rsfm_native_"), "liborsfm_native.so")
    r = subprocess.run(["make", "-s", "-B", "-C", HERE, "native", "NATIVE_OUT=" + out], capture_output=True, text=True)
    if r.returncode != 0 or not os.path.exists(out):
        from oracle import pyoracle

        return pyoracle, "portable oracle build (-O3 -march=x86-64-v3; the native build failed)"
    spec = importlib.util.spec_from_file_location("pyoracle_native", os.path.join(HERE, "pyoracle.py"))
    mod = importlib.util.module_from_spec(spec)
    old = os.environ.get("RSFM_ORACLE_LIB_OVERRIDE")
    os.environ["RSFM_ORACLE_LIB_OVERRIDE"] = out
    try:
        spec.loader.exec_module(mod)
    finally:
        if old is None:
            del os.environ["RSFM_ORACLE_LIB_OVERRIDE"]
        else:
            os.environ["RSFM_ORACLE_LIB_OVERRIDE"] = old
    return mod, "gcc -O3 -march=native, built on this host"


def one_pair(orc, dA, dB, kA, kB, H, seed, threads):
    ra, sa, rb = orc.bf_match_top2_fast(dA, dB, threads)
    m = orc.filter_matches(ra, sa, rb, (4, 5))
    pts = np.concatenate([kA[m[:, 0]], kB[m[:, 1]]], 1).astype(np.float32)
    idx = orc.sample_indices_hash(seed, len(m), H)
    counts, models = orc.f8_score(pts, idx, 1.0, 1, nthreads=threads)
    return orc.f8_finish(pts, idx, counts, models, 1.0, 1, 0.0, 5000, True)


def measure(orc, pairs, H, seconds, threads, seed0=77):
    """pairs/s over a bounded sample: pairs in order, at least one, until `seconds`."""
    t_total, k = 0.0, 0
    while t_total < seconds and k < 100000:
        dA, dB, kA, kB = pairs[k % len(pairs)]
        t = time.perf_counter()
        one_pair(orc, dA, dB, kA, kB, H, seed0 + k % len(pairs), threads)
        t_total += time.perf_counter() - t
        k += 1
    return k / t_total, k, t_total
