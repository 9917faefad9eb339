"""Synthetic two-view inputs for benchmarks and training (seeded).

Descriptors follow the SIFT-like recipe: 128 half-normal floats, L2-normalise,
clip at 0.2, renormalise, q = min(127, round(512 v)); a "copy" of a train
descriptor adds N(0, sigma^2) noise to its float vector before that chain.
They are drawn on the device with torch (data generation, not the matching
path).

Geometry: camera 0 at the origin, camera 1 rotated by <= 15 degrees about a
random axis with a baseline of 1-3 units; 3D points uniform in the depth 10-50
frustum of camera 0 and visible in camera 1; focal 2000 px, 4000 x 3000
images, pixel noise sigma 0.5. Keypoints carry (x, y, scale, orientation):
a consistent copy sees its train point's scale times a per-pair zoom and its
orientation plus a per-pair rotation (with small noise); every other keypoint
has independent position, scale and orientation.
"""
import numpy as np
import torch

FOCAL, WIDTH, HEIGHT, NOISE_PX = 2000.0, 4000.0, 3000.0, 0.5


def sift_like(gen, n, device, base=None, sigma=0.05):
    """(n, 128) uint8; with base (n, 128 float), noisy copies of those vectors."""
    if base is None:
        v = torch.randn((n, 128), generator=gen, device=device).abs()
    else:
        v = base + sigma * torch.randn((n, 128), generator=gen, device=device)
        v = v.clamp_min(0)
    v = v / v.norm(dim=1, keepdim=True).clamp_min(1e-12)
    v = v.clamp_max(0.2)
    v = v / v.norm(dim=1, keepdim=True).clamp_min(1e-12)
    q = torch.round(v * 512).clamp(0, 127)
    return q.to(torch.uint8), v


def _rodrigues(w):
    th = np.linalg.norm(w)
    if th < 1e-12:
        return np.eye(3)
    k = w / th
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K


class TwoView:
    """Random relative pose; project() maps 3D points of camera 0 to both images."""

    def __init__(self, rng):
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        self.R = _rodrigues(axis * np.deg2rad(15.0) * rng.random())
        t = rng.normal(size=3)
        self.t = t / np.linalg.norm(t) * (1.0 + 2.0 * rng.random())
        Kinv = np.array([[1 / FOCAL, 0, -WIDTH / 2 / FOCAL], [0, 1 / FOCAL, -HEIGHT / 2 / FOCAL], [0, 0, 1]])
        tx = np.array([[0, -self.t[2], self.t[1]], [self.t[2], 0, -self.t[0]], [-self.t[1], self.t[0], 0]])
        self.F = Kinv.T @ tx @ self.R @ Kinv  # x1^T F x0 = 0

    def points(self, rng, n):
        """n points visible in both images: (x0, y0, x1, y1) without noise."""
        out = np.zeros((0, 4))
        while len(out) < n:
            m = 2 * (n - len(out)) + 16
            u = WIDTH * rng.random(m)
            v = HEIGHT * rng.random(m)
            d = 10.0 + 40.0 * rng.random(m)
            X = np.stack([(u - WIDTH / 2) / FOCAL * d, (v - HEIGHT / 2) / FOCAL * d, d], 1)
            Y = X @ self.R.T + self.t
            ok = Y[:, 2] > 1e-3
            u1 = np.where(ok, FOCAL * Y[:, 0] / np.where(ok, Y[:, 2], 1) + WIDTH / 2, -1)
            v1 = np.where(ok, FOCAL * Y[:, 1] / np.where(ok, Y[:, 2], 1) + HEIGHT / 2, -1)
            ok &= (u1 >= 0) & (u1 < WIDTH) & (v1 >= 0) & (v1 < HEIGHT)
            out = np.concatenate([out, np.stack([u, v, u1, v1], 1)[ok]])
        return out[:n]


def _random_kp(rng, n):
    kp = np.zeros((n, 4), np.float32)
    kp[:, 0] = WIDTH * rng.random(n)
    kp[:, 1] = HEIGHT * rng.random(n)
    kp[:, 2] = np.exp(rng.uniform(np.log(1.0), np.log(12.0), n))
    kp[:, 3] = rng.uniform(-np.pi, np.pi, n)
    return kp


def make_pair(seed, n_a, n_b, copy_frac, geo_frac, device="cuda", sigma=0.05):
    """One image pair.

    Returns dict: descA (n_a,128) / descB (n_b,128) uint8 device tensors; kpA /
    kpB (n,4) float32 numpy (x, y, scale, orientation); src (n_a,) int32 =
    train index a query copies or -1; geo (n_a,) bool = copy placed
    consistently with the two-view geometry; F (3,3) true fundamental matrix.
    copy_frac of the queries copy a distinct random train descriptor; geo_frac
    of those copies are geometrically consistent.
    """
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed) & 0x7FFFFFFFFFFFFFFF)
    view = TwoView(rng)

    desc_b, vb = sift_like(gen, n_b, device)
    n_copy = min(int(round(copy_frac * n_a)), n_b)
    src_rows = rng.choice(n_b, n_copy, replace=False)
    slots = rng.choice(n_a, n_copy, replace=False)
    copies, _ = sift_like(gen, n_copy, device, base=vb[torch.from_numpy(src_rows).to(device)], sigma=sigma)
    fresh, _ = sift_like(gen, n_a, device)
    desc_a = fresh
    if n_copy:
        desc_a[torch.from_numpy(slots).to(device)] = copies

    kp_b = _random_kp(rng, n_b)
    kp_a = _random_kp(rng, n_a)
    src = np.full(n_a, -1, np.int32)
    src[slots] = src_rows
    geo = np.zeros(n_a, bool)
    is_geo = rng.random(n_copy) < geo_frac
    g_slots, g_src = slots[is_geo], src_rows[is_geo]
    if len(g_slots):
        pts = view.points(rng, len(g_slots))
        kp_a[g_slots, 0:2] = pts[:, 0:2] + NOISE_PX * rng.normal(size=(len(g_slots), 2))
        kp_b[g_src, 0:2] = pts[:, 2:4] + NOISE_PX * rng.normal(size=(len(g_slots), 2))
        zoom = np.exp(rng.uniform(-0.25, 0.25))
        rot = rng.normal(0, 0.2)
        kp_a[g_slots, 2] = kp_b[g_src, 2] / zoom * np.exp(rng.normal(0, 0.1, len(g_slots)))
        kp_a[g_slots, 3] = np.angle(np.exp(1j * (kp_b[g_src, 3] - rot + rng.normal(0, 0.1, len(g_slots)))))
        geo[g_slots] = True
    return dict(descA=desc_a, descB=desc_b, kpA=kp_a.astype(np.float32), kpB=kp_b.astype(np.float32), src=src,
                geo=geo, F=view.F, R=view.R, t=view.t)


def cashash_weights(seed, n_channels=192, device="cuda"):
    """Gaussian random projections for cascade hashing (seeded): int8 weights
    [n_channels][128] = round(clamp(32 N(0,1), -128, 127)) and int32 bias_base[c]
    = round(sum_k w[c][k] r_k), r_k = clamp(16 N(0,1), -64, 64)."""
    gen = torch.Generator()
    gen.manual_seed(int(seed) & 0x7FFFFFFFFFFFFFFF)
    w = torch.round((32.0 * torch.randn((n_channels, 128), generator=gen, dtype=torch.float64)).clamp(-128, 127))
    r = (16.0 * torch.randn((n_channels, 128), generator=gen, dtype=torch.float64)).clamp(-64, 64)
    base = torch.round((w * r).sum(dim=1))
    return w.to(torch.int8).to(device), base.to(torch.int32).to(device)
