"""Oracle (test infrastructure only): the reference's dispersion-search pieces restated in numpy:
the sharpness metrics (octrecon/estimation/dispersion.py:46-63), the B-scan signal score
(:145-150) and the B-scan selection (:111-142).  Pinned by closed forms and special cases
(tests/test_search_oracle.py).  Not used by the product."""
from __future__ import annotations

import numpy as np

Z_MIN_PIX = 20


def sharpness_metric(A, name, z_min=Z_MIN_PIX):
    """octrecon/estimation/dispersion.py:46-63 with numpy: A (B, nX, nZ) float32."""
    A = np.asarray(A, np.float32)
    B = A.shape[0]
    if name == "legacy_peak":
        return np.log(A[:, :, z_min:] + 1e-12).mean(axis=1).max(axis=1).astype(np.float64)
    I = (A[:, :, z_min:] ** 2).reshape(B, -1).astype(np.float64)
    if name == "top_percentile":
        M = I.shape[1]
        k = max(1, int(M * 0.01))
        return np.log(np.partition(I, M - k, axis=1)[:, M - k:].mean(axis=1))
    if name == "kurtosis":
        return (I ** 2).sum(axis=1) / I.sum(axis=1) ** 2
    if name == "entropy":
        p = I / I.sum(axis=1, keepdims=True)
        return (p * np.log(p + 1e-30)).sum(axis=1)
    raise ValueError(name)


def legacy_peak_fp64(A, z_min=Z_MIN_PIX):
    """The legacy metric with the log and the mean in float64 of the float32 A + 1e-12."""
    A = np.asarray(A, np.float32)
    return np.log((A[:, :, z_min:] + np.float32(1e-12)).astype(np.float64)).mean(axis=1).max(axis=1)


def signal_score(A, z_min=Z_MIN_PIX):
    """octrecon/estimation/dispersion.py:145-150 for the magnitudes of one B-scan (all files)."""
    I = np.asarray(A, np.float32)[..., z_min:].ravel().astype(float) ** 2
    return 10 * np.log10(np.percentile(I, 99.5) / (np.median(I) + 1e-30) + 1e-30)


def select_bscans(x_centers, y_centers, z_depths, n_frames, n, rng_seed=0, exists=None):
    """octrecon/estimation/dispersion.py:111-142: (xi, yi, zi, frame) candidates; exists(xi, yi, zi)
    stands for the reference's tile lookup."""
    z = np.asarray(z_depths, float)
    depths = [i for i in range(len(z)) if z[i] > 1e-9]
    if not depths:
        depths = [i for i in range(len(z)) if z[i] >= -1e-9] or list(range(len(z)))

    def central(c):
        c = np.asarray(c)
        if len(c) <= 2:
            return list(range(len(c)))
        span = np.ptp(c)
        ok = np.nonzero(np.abs(c - c.mean()) <= 0.3 * span + 1e-9)[0]
        return list(ok) if len(ok) else [int(np.argmin(np.abs(c - c.mean())))]

    xs, ys = central(x_centers), central(y_centers)
    rng = np.random.default_rng(rng_seed)
    nf = n_frames
    frames = np.unique(np.round(np.linspace(0.1, 0.9, max(n, 2)) * (nf - 1)).astype(int))
    out, seen = [], set()
    for i in range(4 * n):
        if len(out) >= n:
            break
        zi = depths[i % len(depths)]
        xi, yi = int(rng.choice(xs)), int(rng.choice(ys))
        fr = int(frames[i % len(frames)]) if nf > 1 else 0
        if (exists is not None and not exists(xi, yi, zi)) or (xi, yi, zi, fr) in seen:
            continue
        seen.add((xi, yi, zi, fr))
        out.append((xi, yi, zi, fr))
    return out
