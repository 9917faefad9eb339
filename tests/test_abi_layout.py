"""The ctypes mirrors of the C ABI structs have the C layout: a program built
from include/rsfm_amd.h with the host C compiler prints every struct's size and
each field's offset, and the ctypes classes must agree field by field."""
import ctypes
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PAIRS = [
    ("rsfm_best_match", "_abi", "BestMatch"),
    ("rsfm_valid_match", "_abi", "ValidMatch"),
    ("rsfm_bf_task", "_abi", "BfTask"),
    ("rsfm_bf_top2_task", "_abi", "BfTop2Task"),
    ("rsfm_ransac_task", "_abi", "RansacTask"),
    ("rsfm_degensac_task", "_abi", "DegensacTask"),
    ("rsfm_camera", "_abi", "Camera"),
    ("rsfm_pose_task", "_abi", "PoseTask"),
    ("rsfm_pose_h_task", "_abi", "PoseHTask"),
    ("rsfm_real_camera", "_abi", "RealCamera"),
    ("rsfm_undistort_task", "_abi", "UndistortTask"),
    ("rsfm_filter_task", "_abi", "FilterTask"),
    ("rsfm_sample_task", "_abi", "SampleTask"),
    ("rsfm_weighted_sample_task", "_abi", "WSampleTask"),
    ("rsfm_cashash_image", "cashash", "CasHashImage"),
    ("rsfm_cashash_result", "cashash", "CasHashResult"),
    ("rsfm_pipeline_pair", "pipeline", "PipelinePair"),
    ("rsfm_pipeline_config", "pipeline", "PipelineConfig"),
    ("rsfm_pipeline_tuning_entry", "pipeline", "TuningEntry"),
    ("rsfm_pipeline_info_t", "pipeline", "PipelineInfo"),
    ("rsfm_feature_task", "proposal", "FeatureTask"),
    ("rsfm_ransac_finish_task", "_abi", "FinishTask"),
    ("rsfm_cashash_rerank_image", "cashash", "CasHashRerankImage"),
    ("rsfm_rmf_task", "rmf", "RmfTask"),
]


def _classes():
    import importlib

    out = []
    for cname, mod, cls in PAIRS:
        m = importlib.import_module("rapidsfm_amd." + mod)
        out.append((cname, getattr(m, cls)))
    return out


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "rsfm_amd.h"', "int main(void) {"]
    for cname, cls in _classes():
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in cls._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, f[0], cname, f[0]))
    lines += ["return 0;", "}"]
    d = tmp_path_factory.mktemp("abi")
    src, exe = d / "abi.c", d / "abi"
    src.write_text("\n".join(lines))
    subprocess.run([cc, "-std=c11", "-I" + os.path.join(ROOT, "include"), "-I/opt/rocm/include",
                    "-D__HIP_PLATFORM_AMD__", str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        a, b, v = line.split()
        got[(a, b)] = int(v)
    return got


@pytest.mark.parametrize("cname,mod,cls", PAIRS)
def test_ctypes_mirror_matches_c(c_layout, cname, mod, cls):
    import importlib

    k = getattr(importlib.import_module("rapidsfm_amd." + mod), cls)
    assert ctypes.sizeof(k) == c_layout[(cname, "size")], cname
    for f in k._fields_:
        assert getattr(k, f[0]).offset == c_layout[(cname, f[0])], (cname, f[0])
