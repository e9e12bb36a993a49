"""Dispersion (beta) sharpness search and focus detection on the hip device.

Reference: octrecon/estimation/dispersion.py (DispersionScorer :66-107, _metric :46-63,
estimate_dispersion :154-267) and octrecon/estimation/focus.py (imgaussfilt_1d :44-52,
findpeaks_matlab :55-72, detect_peak_near :191-206, interface_pix :209-218).

A search evaluation is the spectral path with the candidate window hann/rms * exp(-i beta dk^2)
(the reference's DispersionScorer window equals the reconstruction window of that beta:
tests/test_estimation.py:98-106), so the device scorer runs the fused spectral kernel with a
per-candidate window and the exact top-1 % metric kernel.  The grid of candidates is queued
without host synchronisation; the bounded refinement (scipy minimize_scalar, as the reference)
evaluates one candidate per call.
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib
from .geometry import SpectralGeometry, build_spectral_geometry

Z_MIN_PIX = 20
TOP_FRACTION = 0.01


class HipDispersionScorer:
    """Scores beta candidates on a fixed set of raw B-scans (B, lines, N) on the device."""

    def __init__(self, lambda_nm, raw, apod_size: int, xp, n_medium: float = 1.33, apod_mode: str = "lines",
                 metric: str = "top_percentile"):
        if metric not in ("top_percentile", "auto"):
            raise NotImplementedError("the device scorer implements the reference's default metric "
                                      "'top_percentile' only")
        from .spectral import HipSpectralProcessor
        self.xp = xp
        self.torch = xp.torch
        self.sg0: SpectralGeometry = build_spectral_geometry(lambda_nm, 0.0, n_medium)
        self.sp = HipSpectralProcessor(self.sg0, apod_size, xp, apod_mode=apod_mode)
        N = self.sg0.n
        hann = 0.5 * (1 - np.cos(2 * np.pi * np.arange(N) / (N - 1)))
        self.hann = hann / np.sqrt(np.mean(hann ** 2))
        k_eq = 2 * np.pi / self.sg0.lambda_eq_nm
        self.dk2 = (k_eq - k_eq.mean()) ** 2
        torch = self.torch
        self._hann_n = torch.tensor(self.hann / N, dtype=torch.float64, device=xp.device)
        self._dk2 = torch.tensor(self.dk2, dtype=torch.float64, device=xp.device)
        self.nz = N // 2
        self._mag = None
        self._ws = None
        self.n_evals = 0
        self.set_raw(raw)

    def set_raw(self, raw):
        """Score another set of B-scans (same spectrometer)."""
        self.raw = self.sp._device_raw(raw)
        self.B = self.raw.shape[0]
        self.nX = self.sp.n_scan_lines(self.raw.shape[1])
        shape = (self.B, self.nX, self.nz)
        if self._mag is None or tuple(self._mag.shape) != shape:
            self._mag = self.torch.empty(shape, dtype=self.torch.float32, device=self.xp.device)

    def windows(self, betas=None, head=None, head_scale=1.0, offset=0.0, to_ref=1.0, count=None, out=None):
        """(count, N, 2) float32 device windows hann/rms/N exp(-i beta dk2), float64 phase
        (oct_dispersion_windows): beta = betas (reference units), or per B-scan from a device
        network head: (head[b, 0] * head_scale + offset) * to_ref."""
        torch = self.torch
        if head is None:
            b = torch.tensor(np.ascontiguousarray(np.atleast_1d(np.asarray(betas, float))), dtype=torch.float64,
                             device=self.xp.device)
            count, hp, hs = b.numel(), None, 1
        else:
            b, hp, hs = None, head.data_ptr(), head.stride(0)
            count = head.shape[0] if count is None else count
        if out is None:
            out = torch.empty((count, self.sg0.n, 2), dtype=torch.float32, device=self.xp.device)
        elif tuple(out.shape) != (count, self.sg0.n, 2) or not out.is_contiguous():
            raise ValueError(f"window buffer must be contiguous ({count}, {self.sg0.n}, 2)")
        _lib.check(_lib.lib().oct_dispersion_windows(None if b is None else b.data_ptr(), hp, hs, head_scale, offset,
                                                     to_ref, count, self._hann_n.data_ptr(), self._dk2.data_ptr(),
                                                     self.sg0.n, out.data_ptr(), self.xp.stream()),
                   "oct_dispersion_windows")
        self._keep_b = b                  # the launch is asynchronous
        return out

    def _metric(self, out_row):
        L = _lib.lib()
        nb = L.oct_topfrac_metric_workspace_bytes(self.B, self.nX, self.nz, Z_MIN_PIX)
        if self._ws is None or self._ws.numel() < nb:
            self._ws = self.torch.empty((nb,), dtype=self.torch.uint8, device=self.xp.device)
        _lib.check(L.oct_topfrac_metric(self._mag.data_ptr(), self.B, self.nX, self.nz, Z_MIN_PIX, TOP_FRACTION,
                                        out_row.data_ptr(), self._ws.data_ptr(), self._ws.numel(), self.xp.stream()),
                   "oct_topfrac_metric")

    def magnitude_at(self, betas, out=None):
        """|A| (B, nX, N/2) with one candidate for all B-scans (scalar) or one per B-scan."""
        betas = np.atleast_1d(np.asarray(betas, float))
        if betas.size not in (1, self.B):
            raise ValueError(f"{betas.size} candidates for {self.B} B-scans")
        win = self.windows(betas)
        return self.magnitude_with(win[0] if betas.size == 1 else win, out)

    def magnitude_with(self, win, out=None):
        """|A| with device windows win: (N, 2) for every B-scan or (B, N, 2) one per B-scan."""
        per_file = win.dim() == 3
        if per_file and win.shape[0] != self.B:
            raise ValueError(f"{win.shape[0]} windows for {self.B} B-scans")
        out = self._mag if out is None else out
        L = _lib.lib()
        try:
            _lib.check(L.oct_plan_use_windows(self.sp.plan.handle, win.data_ptr(), 2 * self.sg0.n if per_file else 0,
                                              0), "oct_plan_use_windows")
            self.sp.run(self.raw, _lib.OCT_OUT_MAG, 0, self.nz, out=out)
        finally:
            L.oct_plan_use_window(self.sp.plan.handle, None)
        self._keep = win                  # alive until the next call (the launch is asynchronous)
        return out

    def magnitude_repeat(self, win, n_repeat: int, bg, out, x0: int = 0, nx: int | None = None):
        """|A| of the batch at n_repeat candidate sets in one launch: win (n_repeat, N, 2) one window
        per repeat, or (n_repeat B, N, 2) one per repeat and B-scan; scan lines [x0, x0 + nx);
        out (n_repeat B, nx, N/2)."""
        N = self.sg0.n
        per_file = win.shape[0] == n_repeat * self.B and not (self.B == 1)
        if win.shape[0] not in (n_repeat, n_repeat * self.B):
            raise ValueError(f"{win.shape[0]} windows for {n_repeat} repeats of {self.B} B-scans")
        fstride, rstride = (2 * N, 2 * N * self.B) if per_file else (0, 2 * N)
        L = _lib.lib()
        try:
            _lib.check(L.oct_plan_use_windows(self.sp.plan.handle, win.data_ptr(), fstride, rstride),
                       "oct_plan_use_windows")
            self.sp.run_repeat(self.raw, n_repeat, bg, out, x0=x0, nx=nx)
        finally:
            L.oct_plan_use_window(self.sp.plan.handle, None)
        self._keep = win
        return out

    def scores_many(self, betas):
        """(G, B) float64 scores for a batch of candidates, one host synchronisation."""
        torch = self.torch
        betas = np.atleast_1d(np.asarray(betas, float))
        win = self.windows(betas)
        out = torch.empty((len(betas), self.B), dtype=torch.float64, device=self.xp.device)
        L = _lib.lib()
        try:
            for g in range(len(betas)):
                _lib.check(L.oct_plan_use_window(self.sp.plan.handle, win[g].data_ptr()), "oct_plan_use_window")
                self.sp.run(self.raw, _lib.OCT_OUT_MAG, 0, self.nz, out=self._mag)
                self._metric(out[g])
        finally:
            L.oct_plan_use_window(self.sp.plan.handle, None)
        self.n_evals += len(betas)
        res = out.cpu().numpy()          # synchronises; keeps `win` alive until the launches ran
        del win
        return res

    def scores_per_file(self, betas):
        """(B,) float64: B-scan b scored at its own candidate betas[b] (one launch)."""
        out = self.torch.empty((self.B,), dtype=self.torch.float64, device=self.xp.device)
        self.magnitude_at(np.asarray(betas, float).reshape(self.B))
        self._metric(out)
        self.n_evals += 1
        return out.cpu().numpy()

    def scores(self, beta, idx=None):
        s = self.scores_many([beta])[0]
        return s if idx is None else s[idx:idx + 1]


def depth_median(mag, xp, out=None):
    """(B, nX, nZ) float32 device magnitudes -> (B, nZ) float64 device median over x (exact, as
    numpy.median of the float32 values in float64; octrecon/estimation/focus.py:213-216)."""
    torch = xp.torch
    B, nX, nZ = mag.shape
    if mag.stride(2) != 1 or mag.stride(1) < nZ or mag.stride(0) != nX * mag.stride(1):
        raise ValueError("magnitudes must be (B, nX, rows) with contiguous rows")
    out = torch.empty((B, nZ), dtype=torch.float64, device=xp.device) if out is None else out
    _lib.check(_lib.lib().oct_depth_median(mag.data_ptr(), B, nX, mag.stride(1), nZ, out.data_ptr(), xp.stream()),
               "oct_depth_median")
    return out


def search_per_bscan(scorer: HipDispersionScorer, lo: float, hi: float, n_grid: int = 41, iters: int = 24):
    """Each B-scan's own sharpness maximum over [lo, hi]: the reference's coarse grid
    (octrecon/estimation/dispersion.py:206-214), then the bracket around each B-scan's best grid
    point refined by golden-section search, all B-scans advanced together (one launch per step).
    Returns (beta (B,), score (B,), grid scores (G, B))."""
    grid = np.linspace(lo, hi, int(n_grid))
    S = scorer.scores_many(grid)
    gi = np.argmax(S, 0)
    a = grid[np.maximum(gi - 1, 0)].astype(float)
    b = grid[np.minimum(gi + 1, len(grid) - 1)].astype(float)
    r = (np.sqrt(5.0) - 1) / 2
    c = b - r * (b - a)
    d = a + r * (b - a)
    fc = scorer.scores_per_file(c)
    fd = scorer.scores_per_file(d)
    for _ in range(iters):
        left = fc >= fd                           # maximum in [a, d]
        b = np.where(left, d, b)
        a = np.where(left, a, c)
        nc = np.where(left, b - r * (b - a), d)
        nd = np.where(left, c, a + r * (b - a))
        fnew = scorer.scores_per_file(np.where(left, nc, nd))
        fc, fd = np.where(left, fnew, fd), np.where(left, fc, fnew)
        c, d = nc, nd
    x = 0.5 * (c + d)
    fx = np.maximum(fc, fd)
    best_grid = S[gi, np.arange(S.shape[1])]
    use_grid = best_grid > fx
    return np.where(use_grid, grid[gi], x), np.where(use_grid, best_grid, fx), S


def estimate_dispersion(scorer: HipDispersionScorer, initial: float, bounds=None, n_grid: int = 41, log=None):
    """Coarse grid over [0.5, 1.5] * initial (or `bounds`), per-B-scan normalisation of the
    score curves, bounded refinement of the mean normalised score and of each B-scan's own
    curve (octrecon/estimation/dispersion.py:206-239)."""
    from scipy.optimize import minimize_scalar
    log = log or (lambda *a: None)
    t0 = time.perf_counter()
    if bounds is None:
        lo, hi = sorted((0.5 * initial, 1.5 * initial)) if initial != 0 else (-3e8, 3e8)
    else:
        lo, hi = sorted(float(b) for b in bounds)
    grid = np.linspace(lo, hi, int(n_grid))
    S = scorer.scores_many(grid)                        # (G, B)
    smin, smax = S.min(0), S.max(0)
    rngs = np.where(smax > smin, smax - smin, 1.0)
    agg = ((S - smin) / rngs).mean(1)
    gi = int(np.argmax(agg))

    def bracket(i):
        return grid[max(i - 1, 0)], grid[min(i + 1, len(grid) - 1)]

    tol = max(abs(initial), abs(lo), abs(hi)) * 2e-5
    a, b = bracket(gi)
    r = minimize_scalar(lambda x: -float(np.mean((scorer.scores(x) - smin) / rngs)), bounds=(a, b),
                        method="bounded", options={"xatol": tol})
    value = float(r.x) if -r.fun >= agg[gi] else float(grid[gi])
    per = []
    for j in range(S.shape[1]):
        i = int(np.argmax(S[:, j]))
        a, b = bracket(i)
        rj = minimize_scalar(lambda x: -float(scorer.scores(x, j)[0]), bounds=(a, b), method="bounded",
                             options={"xatol": tol})
        per.append(float(rj.x) if -rj.fun >= S[i, j] else float(grid[i]))
    per = np.array(per)
    el = time.perf_counter() - t0
    at_bound = bool(value <= lo + (hi - lo) / (n_grid - 1) or value >= hi - (hi - lo) / (n_grid - 1))
    log(f"dispersion estimate {value:.5g} ({scorer.n_evals} evaluations, {el * 1e3:.1f} ms)")
    return {"value": value, "bounds": [lo, hi], "candidates": grid.tolist(), "score_curve": agg.tolist(),
            "per_bscan_values": per.tolist(), "median_per_bscan": float(np.median(per)),
            "spread_mad": float(1.4826 * np.median(np.abs(per - np.median(per)))), "at_bound": at_bound,
            "n_evaluations": scorer.n_evals, "elapsed_s": el}


# --------------------------------------------------------------------------- focus (host, 1-D)
def matlab_round(x):
    return np.floor(np.abs(x) + 0.5) * np.sign(x)


def imgaussfilt_1d(v, sigma=4.0):
    """imgaussfilt for a vector: kernel 2*ceil(2 sigma)+1 taps, replicate padding."""
    v = np.asarray(v, float).ravel()
    h = int(np.ceil(2 * sigma))
    x = np.arange(-h, h + 1)
    k = np.exp(-x ** 2 / (2 * sigma ** 2))
    k /= k.sum()
    p = np.concatenate([np.full(h, v[0]), v, np.full(h, v[-1])])
    return np.convolve(p, k, mode="valid")


def findpeaks_matlab(v):
    """0-based indices of strict local maxima (flat peaks at their first sample, no endpoints)."""
    v = np.asarray(v, float)
    n = len(v)
    out = []
    i = 1
    while i < n - 1:
        if v[i] > v[i - 1]:
            j = i
            while j < n - 1 and v[j + 1] == v[i]:
                j += 1
            if j < n - 1 and v[j + 1] < v[i]:
                out.append(i)
            i = j + 1
        else:
            i += 1
    return np.asarray(out, int)


def detect_peak_near(profile, guess_pix, search_px=60, baseline_sigma=30.0, smooth_sigma=4.0):
    """Confocal peak of a depth profile nearest a 1-based guess: dB, detrend (smooth minus
    baseline), strict local maxima within +-search_px, the most prominent.  (pix or nan, dB)."""
    prof = np.asarray(profile, float)
    with np.errstate(divide="ignore", invalid="ignore"):
        lp = 10 * np.log10(np.where(np.isfinite(prof) & (prof > 0), prof, np.nan))
    ok = np.isfinite(lp)
    if ok.sum() < 10:
        return np.nan, 0.0
    lp = np.interp(np.arange(len(lp)), np.nonzero(ok)[0], lp[ok])
    detr = imgaussfilt_1d(lp, smooth_sigma) - imgaussfilt_1d(lp, baseline_sigma)
    p = findpeaks_matlab(detr) + 1
    p = p[np.abs(p - guess_pix) <= search_px]
    if p.size == 0:
        return np.nan, 0.0
    best = p[np.argmax(detr[p - 1])]
    return int(best), float(detr[best - 1])


def depth_profile(mag):
    """mean over frames of the median over x of |A|: (F, nX, nZ) -> (nZ,)."""
    m = np.asarray(mag, float)
    return np.nanmean(np.nanmedian(m, axis=1), axis=0)


def interface_pix(mag):
    """Brightest pixel of the x-median profile below the top third, median over frames (1-based)."""
    m = np.asarray(mag, float)
    nz = m.shape[-1]
    vals = []
    for f in m:
        med = np.nanmedian(f, axis=0)
        med[: int(matlab_round(nz / 3))] = np.nan
        vals.append(int(np.nanargmax(med)) + 1)
    return int(matlab_round(np.median(vals))), vals
