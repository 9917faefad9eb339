"""rsfm_bf_match_top2_noalloc (fixed device scratch, no allocation) against
rsfm_bf_match_top2 with a caller workspace: identical results on two calls in
a row, the second with a smaller input; an oversized batch is rejected."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _tasks(pairs, dev):
    from rapidsfm_amd._abi import BfTask, BfTop2Task

    tasks = (BfTop2Task * len(pairs))()
    outs = []
    for i, (a, b) in enumerate(pairs):
        o = [torch.full((max(1, n), 2), -5, dtype=torch.int32, device=dev) for n in (a.shape[0], a.shape[0], b.shape[0])]
        tasks[i] = BfTop2Task(BfTask(a.shape[0], b.shape[0], a.data_ptr(), b.data_ptr(), o[0].data_ptr(),
                                     o[2].data_ptr(), None, None), o[1].data_ptr())
        outs.append(o)
    return tasks, outs


def _with_workspace(pairs, dev, stream):
    from rapidsfm_amd._abi import check, lib

    tasks, outs = _tasks(pairs, dev)
    tp = ctypes.cast(tasks, ctypes.c_void_p)
    need = ctypes.c_size_t(0)
    check(lib.rsfm_bf_match_top2_workspace_size(tp, len(pairs), ctypes.byref(need)), "ws")
    ws = torch.empty(need.value + 256, dtype=torch.uint8, device=dev)
    check(lib.rsfm_bf_match_top2(tp, len(pairs), ctypes.c_void_p((ws.data_ptr() + 255) // 256 * 256), need.value,
                                 stream), "top2")
    return outs, need.value


def test_noalloc_matches_workspace_entry(dev):
    from rapidsfm_amd import synth
    from rapidsfm_amd._abi import check, lib

    lib.rsfm_bf_match_top2_noalloc.restype = ctypes.c_int
    lib.rsfm_bf_match_top2_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    big = [synth.make_pair(60 + k, n, m, 0.4, 0.6, device=dev) for k, (n, m) in
           enumerate([(8192, 8192), (5000, 7000), (300, 9000), (8192, 100)])]
    small = [synth.make_pair(80 + k, n, m, 0.4, 0.6, device=dev) for k, (n, m) in enumerate([(2048, 2048), (700, 900)])]
    for name, prs in (("first call, 4 pairs up to 8K", big), ("second call, 2 smaller pairs", small)):
        _say("noalloc:", name)
        pairs = [(p["descA"], p["descB"]) for p in prs]
        ref, need = _with_workspace(pairs, dev, s)
        tasks, outs = _tasks(pairs, dev)
        check(lib.rsfm_bf_match_top2_noalloc(ctypes.cast(tasks, ctypes.c_void_p), len(pairs), s), "noalloc")
        torch.cuda.synchronize()
        assert need > 0
        for (a, b), r, o in zip(pairs, ref, outs):
            for k, n in enumerate((a.shape[0], a.shape[0], b.shape[0])):
                assert torch.equal(o[k][:n], r[k][:n]), (name, k)
    # the oracle on one pair of the second call
    from oracle import pyoracle as orc

    a, b = small[1]["descA"].cpu().numpy(), small[1]["descB"].cpu().numpy()
    ra, sa, rb = orc.bf_match_top2_u8(a, b)
    u32 = lambda t: t.cpu().numpy().view(np.uint32)  # noqa: E731
    assert np.array_equal(u32(outs[1][0][:700]), ra) and np.array_equal(u32(outs[1][1][:700]), sa)
    assert np.array_equal(u32(outs[1][2][:900]), rb)


def test_noalloc_rejects_oversized(dev):
    from rapidsfm_amd._abi import RsfmError, check, lib

    lib.rsfm_bf_match_top2_noalloc.restype = ctypes.c_int
    lib.rsfm_bf_match_top2_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    _say("noalloc: oversized batch rejected before any launch")
    # 70 train sets of 8192 descriptors need ~75 MiB of staging: more than the 64 MiB array
    a = torch.zeros((16, 128), dtype=torch.uint8, device=dev)
    b = torch.zeros((8192, 128), dtype=torch.uint8, device=dev)
    tasks, _ = _tasks([(a, b)] * 70, dev)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    with pytest.raises(RsfmError, match="64 MiB"):
        check(lib.rsfm_bf_match_top2_noalloc(ctypes.cast(tasks, ctypes.c_void_p), 70, s), "noalloc")
