"""Learned RANSAC proposals: the 8-64-64-64-1 network on bf16 matrix cores and
the weighted minimal-sample sampler (torch tensors for device memory only).

No reference counterpart: the reference draws minimal samples uniformly
(NonDuplicateSampler, ransac.hpp:23-82). Plain uniform sampling stays the
fallback (all-zero weights) and the judge of inlier counts.
"""
import ctypes

import numpy as np
import torch

from ._abi import WSampleTask, check, lib
from .ops import _ptr, _side, _stream

N_PARAM = lib.rsfm_mlp_param_count()
H = 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * 8
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H * H
OFF_W4 = OFF_B3 + H
OFF_B4 = OFF_W4 + H
assert OFF_B4 + 1 == N_PARAM


def init_params(seed=0, device=None):
    """He-normal weights, zero biases (FP32 master vector)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAM, np.float32)
    p[OFF_W1:OFF_B1] = rng.normal(0, np.sqrt(2.0 / 8), H * 8)
    p[OFF_W2:OFF_B2] = rng.normal(0, np.sqrt(2.0 / H), H * H)
    p[OFF_W3:OFF_B3] = rng.normal(0, np.sqrt(2.0 / H), H * H)
    p[OFF_W4:OFF_B4] = rng.normal(0, np.sqrt(1.0 / H), H)
    return torch.from_numpy(p).to(device or "cuda")


def _check_x(x):
    if x.dtype != torch.float32 or x.dim() != 2 or x.shape[1] != 8 or not x.is_cuda or not x.is_contiguous():
        raise ValueError("features must be a contiguous CUDA float32 tensor of shape (n, 8)")


def forward(params, x, stream=None):
    """Sigmoid outputs (n,) FP32."""
    _check_x(x)
    out = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
    check(lib.rsfm_mlp_forward(_ptr(params), _ptr(x), x.shape[0], _ptr(out), _stream(stream)), "rsfm_mlp_forward")
    return out


class GradWorkspace:
    """Gradient workspace, one buffer per (device, stream): calls on different
    streams never share one."""

    def __init__(self):
        self.bufs = {}

    def get(self, n, device, stream=None):
        need = ctypes.c_size_t(0)
        check(lib.rsfm_mlp_grad_workspace_size(n, ctypes.byref(need)), "rsfm_mlp_grad_workspace_size")
        s = stream or torch.cuda.current_stream(device)
        key = (str(device), s.cuda_stream)
        buf = self.bufs.get(key)
        if buf is None or buf.numel() < need.value:
            buf = torch.empty(need.value, dtype=torch.uint8, device=device)
            if stream is not None:
                buf.record_stream(stream)
            self.bufs[key] = buf
        return buf, need.value


_ws = GradWorkspace()


def grad(params, x, y, stream=None, workspace=None):
    """(gradient of the mean BCE (N_PARAM,), mean loss (1,)) as device tensors."""
    _check_x(x)
    n = x.shape[0]
    g = torch.empty(N_PARAM, dtype=torch.float32, device=x.device)
    loss = torch.empty(1, dtype=torch.float32, device=x.device)
    buf, nbytes = (workspace or _ws).get(n, x.device, stream)
    _side(stream, g, loss)
    check(lib.rsfm_mlp_grad(_ptr(params), _ptr(x), _ptr(y), n, _ptr(g), _ptr(loss), _ptr(buf), nbytes,
                            _stream(stream)), "rsfm_mlp_grad")
    return g, loss


class Adam:
    """Adam on the FP32 master parameters (device kernel)."""

    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999), eps=1e-8):
        self.params = params
        self.m = torch.zeros_like(params)
        self.v = torch.zeros_like(params)
        self.lr, self.b1, self.b2, self.eps = lr, betas[0], betas[1], eps
        self.t = 0

    def step(self, g, stream=None):
        self.t += 1
        check(lib.rsfm_mlp_adam_step(_ptr(self.params), _ptr(g), _ptr(self.m), _ptr(self.v), self.t, self.lr,
                                     self.b1, self.b2, self.eps, _stream(stream)), "rsfm_mlp_adam_step")


def sample_weighted(tasks, stream=None):
    """tasks: list of (weights float32 (n,), H, seed[, n_dev int32 (1,)]) -> list of (H, 8) int32 index sets."""
    outs, keep = [], []
    arr = (WSampleTask * max(1, len(tasks)))()
    for i, t in enumerate(tasks):
        w, h, seed = t[0], t[1], t[2]
        n_dev = t[3] if len(t) > 3 else None
        idx = torch.empty((h, 8), dtype=torch.int32, device=w.device)
        cdf = torch.empty(max(1, w.shape[0]), dtype=torch.int64, device=w.device)
        arr[i] = WSampleTask(idx.data_ptr() if h else None, h, w.shape[0],
                             n_dev.data_ptr() if n_dev is not None else None,
                             w.data_ptr() if w.numel() else None, cdf.data_ptr(), seed)
        outs.append(idx)
        keep.append(cdf)
    _side(stream, *outs, *keep)
    check(lib.rsfm_ransac_sample_weighted_batch(ctypes.cast(arr, ctypes.c_void_p), len(tasks), _stream(stream)),
          "rsfm_ransac_sample_weighted_batch")
    return outs
