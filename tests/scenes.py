"""Synthetic two-view scenes with a dominant plane (test data only).

Camera 0 is K[I|0], camera 1 is K[R|t]. Plane points lie on n.X = d; off-plane
points at random depths; outliers are uniform over the image. Returns float32
(x0, y0, x1, y1) rows, a label per row (0 plane, 1 off-plane, 2 outlier), the
true F (x1^T F x0 = 0) and the plane's H (x1 ~ H x0), both unit Frobenius.
"""
import numpy as np


def _rot(rng, max_deg):
    ax = rng.normal(size=3)
    ax /= np.linalg.norm(ax)
    a = np.deg2rad(rng.uniform(0.3, 1.0) * max_deg)
    Kx = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    return np.eye(3) + np.sin(a) * Kx + (1 - np.cos(a)) * Kx @ Kx


def _skew(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])


def plane_scene(seed, n_plane, n_off, n_out, noise=0.0, width=1600.0, height=1200.0, focal=1200.0):
    rng = np.random.default_rng(seed)
    K = np.array([[focal, 0, width / 2], [0, focal, height / 2], [0, 0, 1.0]])
    Ki = np.linalg.inv(K)
    R = _rot(rng, 8.0)
    t = rng.normal(size=3)
    t[2] *= 0.3
    t = t / np.linalg.norm(t) * 0.8
    nrm = np.array([rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), 1.0])
    nrm /= np.linalg.norm(nrm)
    d = 6.0

    def project(X):
        x0 = (K @ X.T).T
        X1 = (R @ X.T).T + t
        x1 = (K @ X1.T).T
        return np.hstack([x0[:, :2] / x0[:, 2:], x1[:, :2] / x1[:, 2:]]), X1[:, 2]

    rows, labels = [], []
    for kind, count in ((0, n_plane), (1, n_off)):
        got = 0
        while got < count:
            u = np.column_stack([rng.uniform(0, width, 4 * count), rng.uniform(0, height, 4 * count),
                                 np.ones(4 * count)])
            ray = (Ki @ u.T).T
            depth = d / (ray @ nrm) if kind == 0 else rng.uniform(3.0, 14.0, ray.shape[0])
            X = ray * depth[:, None]
            p, z1 = project(X)
            ok = (depth > 0.5) & (z1 > 0.5) & (p[:, 2] > 0) & (p[:, 2] < width) & (p[:, 3] > 0) & (p[:, 3] < height)
            p = p[ok][: count - got]
            rows.append(p)
            labels.append(np.full(p.shape[0], kind))
            got += p.shape[0]
    pts = np.vstack(rows)
    pts = pts + rng.normal(scale=noise, size=pts.shape) if noise > 0 else pts
    out = np.column_stack([rng.uniform(0, width, n_out), rng.uniform(0, height, n_out), rng.uniform(0, width, n_out),
                           rng.uniform(0, height, n_out)])
    pts = np.vstack([pts, out])
    lab = np.concatenate(labels + [np.full(n_out, 2)])
    perm = rng.permutation(pts.shape[0])
    F = Ki.T @ _skew(t) @ R @ Ki
    H = K @ (R + np.outer(t, nrm) / d) @ Ki
    return (pts[perm].astype(np.float32), lab[perm], F / np.linalg.norm(F), H / np.linalg.norm(H))


def same_up_to_scale(A, B):
    """max |A/|A| - s B/|B|| over the sign s"""
    a = np.asarray(A, np.float64).ravel()
    b = np.asarray(B, np.float64).ravel()
    a = a / np.linalg.norm(a)
    b = b / np.linalg.norm(b)
    return min(np.abs(a - b).max(), np.abs(a + b).max())


def posed_scene(seed, n, n_out=0, noise=0.0, width=1600.0, height=1200.0, focal=1200.0, focal1=None):
    """General two-view scene with its pose: returns float32 rows (x0, y0, x1, y1)
    (inliers first, then n_out outliers), cameras as (fx, fy, cx, cy), the true
    R, t (x_cam1 = R x_cam0 + t, |t| = 1), the 3D points in camera 0 and F."""
    rng = np.random.default_rng(seed)
    f1 = focal if focal1 is None else focal1
    cam0 = np.array([focal, focal, width / 2, height / 2])
    cam1 = np.array([f1, f1, width / 2 + 20.0, height / 2 - 15.0])
    K0 = np.array([[cam0[0], 0, cam0[2]], [0, cam0[1], cam0[3]], [0, 0, 1.0]])
    K1 = np.array([[cam1[0], 0, cam1[2]], [0, cam1[1], cam1[3]], [0, 0, 1.0]])
    R = _rot(rng, 10.0)
    t = rng.normal(size=3)
    t[2] *= 0.3
    t /= np.linalg.norm(t)
    Xs, rows = [], []
    while len(rows) < n:
        u = np.array([rng.uniform(0, width), rng.uniform(0, height), 1.0])
        X = np.linalg.inv(K0) @ u * rng.uniform(4.0, 20.0)
        X1 = R @ X + t
        if X1[2] < 0.5:
            continue
        p1 = K1 @ X1
        p1 = p1[:2] / p1[2]
        if not (0 < p1[0] < width and 0 < p1[1] < height):
            continue
        Xs.append(X)
        rows.append([u[0], u[1], p1[0], p1[1]])
    pts = np.array(rows)
    if noise > 0:
        pts = pts + rng.normal(scale=noise, size=pts.shape)
    out = np.column_stack([rng.uniform(0, width, n_out), rng.uniform(0, height, n_out), rng.uniform(0, width, n_out),
                           rng.uniform(0, height, n_out)])
    F = np.linalg.inv(K1).T @ _skew(t) @ R @ np.linalg.inv(K0)
    return (np.vstack([pts, out]).astype(np.float32), cam0.astype(np.float32), cam1.astype(np.float32), R, t,
            np.array(Xs), F / np.linalg.norm(F))


def posed_plane_scene(seed, n, n_out=0, noise=0.0, width=1600.0, height=1200.0, focal=1200.0, focal1=None):
    """Plane-only two-view scene with its pose: returns float32 rows (x0, y0, x1, y1)
    (plane points first, then n_out outliers), cameras as (fx, fy, cx, cy), the true
    R, t (x_cam1 = R x_cam0 + t, |t| = 1), the plane (normal nrm, |nrm| = 1, and
    distance d: nrm . X = d in camera 0) and the pixel homography (unit Frobenius)."""
    rng = np.random.default_rng(seed)
    f1 = focal if focal1 is None else focal1
    cam0 = np.array([focal, focal, width / 2, height / 2])
    cam1 = np.array([f1, f1, width / 2 - 12.0, height / 2 + 9.0])
    K0 = np.array([[cam0[0], 0, cam0[2]], [0, cam0[1], cam0[3]], [0, 0, 1.0]])
    K1 = np.array([[cam1[0], 0, cam1[2]], [0, cam1[1], cam1[3]], [0, 0, 1.0]])
    R = _rot(rng, 10.0)
    t = rng.normal(size=3)
    t[2] *= 0.3
    t /= np.linalg.norm(t)
    nrm = np.array([rng.uniform(-0.4, 0.4), rng.uniform(-0.4, 0.4), 1.0])
    nrm /= np.linalg.norm(nrm)
    d = rng.uniform(5.0, 9.0)
    rows = []
    while len(rows) < n:
        u = np.array([rng.uniform(0, width), rng.uniform(0, height), 1.0])
        ray = np.linalg.inv(K0) @ u
        X = ray * (d / (ray @ nrm))
        X1 = R @ X + t
        if X[2] < 0.5 or X1[2] < 0.5:
            continue
        p1 = K1 @ X1
        p1 = p1[:2] / p1[2]
        if not (0 < p1[0] < width and 0 < p1[1] < height):
            continue
        rows.append([u[0], u[1], p1[0], p1[1]])
    pts = np.array(rows)
    if noise > 0:
        pts = pts + rng.normal(scale=noise, size=pts.shape)
    out = np.column_stack([rng.uniform(0, width, n_out), rng.uniform(0, height, n_out), rng.uniform(0, width, n_out),
                           rng.uniform(0, height, n_out)])
    H = K1 @ (R + np.outer(t, nrm) / d) @ np.linalg.inv(K0)
    return (np.vstack([pts, out]).astype(np.float32), cam0.astype(np.float32), cam1.astype(np.float32), R, t, nrm, d,
            H / np.linalg.norm(H))
