"""Dispersion (beta) sharpness search on the hip device.

Reference: octrecon/estimation/dispersion.py (_metric :46-63, DispersionScorer :66-107,
select_bscans :111-142, _signal_score :145-150, estimate_dispersion :154-267).

A search evaluation is the spectral path with the candidate window hann/rms * exp(-i beta dk^2)
(the reference's DispersionScorer window equals the reconstruction window of that beta:
tests/test_estimation.py:98-106), so the device scorer runs the fused spectral kernel with a
per-candidate window and a sharpness-metric kernel (the exact top-1 % mean by radix selection,
or the moment / entropy / legacy-peak reductions).  The grid of candidates is queued without
host synchronisation; the bounded refinement (scipy minimize_scalar, as the reference) evaluates
one candidate per call.  The B-scans of a volume are chosen as the reference chooses them: spread
over central tiles at tissue depths, twice as many candidates as needed, the ones with the
largest bright-pixel dynamic range (exact percentile and median by device order statistics) kept.
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib
from .geometry import SpectralGeometry, build_spectral_geometry

Z_MIN_PIX = 20
TOP_FRACTION = 0.01
METRICS = ("top_percentile", "kurtosis", "entropy", "legacy_peak")


def sharpness_metric(mag, metric, xp, ws=None, out=None):
    """(B,) float64 device scores of magnitudes (B, nX, nZ) (contiguous float32) by one of the
    reference's metrics over rows z >= 20 (octrecon/estimation/dispersion.py:46-63); `ws(nbytes)`
    may supply the workspace, `out` the (B,) contiguous float64 device result."""
    torch = xp.torch
    B, nX, nZ = mag.shape
    if not mag.is_contiguous() or mag.dtype != torch.float32:
        raise ValueError("magnitudes must be contiguous float32 (B, nX, nZ)")
    code = _lib.OCT_METRICS[metric]
    L = _lib.lib()
    nb = L.oct_sharpness_metric_workspace_bytes(code, B, nX, nZ, Z_MIN_PIX)
    w = ws(nb) if ws is not None else torch.empty((max(nb, 16),), dtype=torch.uint8, device=xp.device)
    if out is None:
        out = torch.empty((B,), dtype=torch.float64, device=xp.device)
    elif out.numel() != B or not out.is_contiguous() or out.dtype != torch.float64:
        raise ValueError(f"metric output must be {B} contiguous float64 values")
    _lib.check(L.oct_sharpness_metric(code, mag.data_ptr(), B, nX, nZ, Z_MIN_PIX, out.data_ptr(), w.data_ptr(),
                                      w.numel(), xp.stream()), "oct_sharpness_metric")
    return out


def order_stats(mag, ranks_desc, xp, square: bool = False):
    """Exact order statistics of each B-scan's values over rows z >= 20: mag (B, nX, nZ) device
    float32, ranks_desc (B, R) host int (1-based, from the largest) -> (B, R) host float64 of
    |A| (square=False) or of the float32 |A|^2 (oct_order_stats)."""
    torch = xp.torch
    B, nX, nZ = mag.shape
    r = np.ascontiguousarray(np.asarray(ranks_desc, np.int64).reshape(B, -1))
    M = nX * (nZ - Z_MIN_PIX)
    if r.size and (r.min() < 1 or r.max() > M):
        raise ValueError(f"ranks must lie in [1, {M}]")
    if not mag.is_contiguous():
        raise ValueError("magnitudes must be contiguous")
    L = _lib.lib()
    nb = L.oct_order_stats_workspace_bytes(B, r.shape[1], nX, nZ, Z_MIN_PIX)
    ws = torch.empty((max(nb, 16),), dtype=torch.uint8, device=xp.device)
    rd = torch.from_numpy(r).to(xp.device)
    out = torch.empty(r.shape, dtype=torch.float64, device=xp.device)
    _lib.check(L.oct_order_stats(mag.data_ptr(), B, nX, nZ, Z_MIN_PIX, int(square), rd.data_ptr(), r.shape[1],
                                 out.data_ptr(), ws.data_ptr(), ws.numel(), xp.stream()), "oct_order_stats")
    return out.cpu().numpy()


def _linear_percentile(value_at, n: int, q: float) -> float:
    """numpy.percentile's default ('linear') rule for the q-th percentile of n sorted values, given
    value_at(i) = the i-th smallest (0-based): the virtual index (n - 1) q / 100, and the
    interpolation numpy uses between the two neighbouring values (from the upper one when the
    fraction is at least 1/2)."""
    qq = q / 100.0
    h = (n - 1) * qq
    i0 = int(np.floor(h))
    i0 = min(max(i0, 0), n - 1)
    i1 = min(i0 + 1, n - 1)
    t = h - np.floor(h)
    a, b = value_at(i0), value_at(i1)
    d = b - a
    return b - d * (1.0 - t) if t >= 0.5 else a + d * t


def signal_scores(mag, xp, files_per_bscan: int = 1):
    """Bright-pixel dynamic range (dB) of each B-scan, the reference's B-scan screen
    (octrecon/estimation/dispersion.py:145-150): 10 log10(P99.5(I) / (median(I) + 1e-30) + 1e-30)
    over I = |A|^2 (float64) of rows z >= 20 of every file of the B-scan.  mag: device magnitudes
    (n files_per_bscan, nX, nZ); the files of one B-scan are consecutive.  Exact: the four order
    statistics the percentile and the median need come from the device selection."""
    F, nX, nZ = mag.shape
    if F % files_per_bscan:
        raise ValueError(f"{F} files are not whole B-scans of {files_per_bscan}")
    n_b = F // files_per_bscan
    m = mag.contiguous().view(n_b, files_per_bscan * nX, nZ)
    n = files_per_bscan * nX * (nZ - Z_MIN_PIX)
    asc = signal_order_indices(n)
    vals = order_stats(m, np.tile(np.array([n - i for i in asc]), (n_b, 1)), xp)
    return np.array([signal_from_order_stats(dict(zip(asc, v.astype(float) ** 2)), n) for v in vals])


def signal_order_indices(n: int):
    """Ascending 0-based positions among n values that the signal score reads: the two around the
    99.5th percentile's virtual index and the one or two middle ones."""
    h = (n - 1) * (99.5 / 100.0)
    i0 = min(max(int(np.floor(h)), 0), n - 1)
    return sorted({i0, min(i0 + 1, n - 1), (n - 1) // 2, n // 2})


def signal_from_order_stats(at: dict, n: int) -> float:
    """10 log10(P99.5 / (median + 1e-30) + 1e-30) from the values `at` {ascending position: I} of
    signal_order_indices(n) (numpy's linear percentile and its median of the middle pair)."""
    p995 = _linear_percentile(lambda i: at[i], n, 99.5)
    med = at[(n - 1) // 2] if n % 2 else np.mean([at[(n - 1) // 2], at[n // 2]])
    return float(10 * np.log10(p995 / (med + 1e-30) + 1e-30))


class HipDispersionScorer:
    """Scores beta candidates on a fixed set of raw B-scans (B, lines, N) on the device, with one
    of the reference's sharpness metrics (default 'top_percentile')."""

    def __init__(self, lambda_nm, raw, apod_size: int, xp, n_medium: float = 1.33, apod_mode: str = "lines",
                 metric: str = "top_percentile"):
        metric = "top_percentile" if metric in (None, "", "auto") else metric
        if metric not in METRICS:
            raise ValueError(f"unknown metric {metric!r}; choose from {METRICS} or 'auto'")
        self.metric = metric
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

    def _metric(self, out_row, mag=None):
        """The scorer's metric of each B-scan of `mag` (default: the scorer's magnitudes) into
        the device float64 row out_row (oct_sharpness_metric)."""
        mag = self._mag if mag is None else mag
        sharpness_metric(mag, self.metric, self.xp, ws=self._workspace, out=out_row)

    def _workspace(self, nbytes):
        if self._ws is None or self._ws.numel() < nbytes:
            self._ws = self.torch.empty((max(nbytes, 16),), dtype=self.torch.uint8, device=self.xp.device)
        return self._ws

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




# --------------------------------------------------------------------------- the search
def _search_bounds(initial: float, bounds):
    """[lo, hi]: the given bounds, else initial x [0.5, 1.5] (either order), else +-3e8 for an
    initial value of 0 (octrecon/estimation/dispersion.py:178-181)."""
    if bounds is not None:
        lo, hi = sorted(float(b) for b in bounds)
    elif initial != 0:
        lo, hi = sorted((0.5 * initial, 1.5 * initial))
    else:
        lo, hi = -3e8, 3e8
    return lo, hi


def _refine_peak(score, grid, curve, xatol):
    """Maximum of score(beta) near the best node of a coarse curve: bounded Brent
    (scipy minimize_scalar) over the node's two neighbouring grid intervals (clipped at the grid
    ends); the grid node stands unless the refinement reaches at least its score."""
    from scipy.optimize import minimize_scalar
    i = int(np.argmax(curve))
    a, b = grid[max(i - 1, 0)], grid[min(i + 1, len(grid) - 1)]
    res = minimize_scalar(lambda x: -score(x), bounds=(a, b), method="bounded", options={"xatol": xatol})
    return float(res.x) if -res.fun >= curve[i] else float(grid[i])


def estimate_dispersion(scorer: HipDispersionScorer, initial: float, bounds=None, n_grid: int = 41, log=None):
    """Sharpness search over the scorer's B-scans, as the reference runs it
    (octrecon/estimation/dispersion.py:213-244):

    1. every B-scan scored on a uniform grid of n_grid candidates over the bounds;
    2. each B-scan's curve rescaled to [0, 1] (a flat curve keeps unit range) and the curves
       averaged; the average's best node refined (xatol = 2e-5 of the largest of |initial|,
       |lo|, |hi|) -> the estimate;
    3. each B-scan's own curve refined the same way -> per-B-scan optima (spread, stability).

    Returns the reference's result keys (value, bounds, candidates, score_curve,
    score_curves_per_bscan, per_bscan_values, median_per_bscan, spread_mad, at_bound,
    n_evaluations, elapsed_s)."""
    log = log or (lambda *a: None)
    t0 = time.perf_counter()
    lo, hi = _search_bounds(initial, bounds)
    grid = np.linspace(lo, hi, int(n_grid))
    curves = scorer.scores_many(grid)                                  # (G, B)
    floor = curves.min(axis=0)
    span = curves.max(axis=0) - floor
    span[~(span > 0)] = 1.0
    unit = (curves - floor) / span
    mean_curve = unit.mean(axis=1)
    xatol = 2e-5 * max(abs(initial), abs(lo), abs(hi))
    value = _refine_peak(lambda x: float(np.mean((scorer.scores(x) - floor) / span)), grid, mean_curve, xatol)
    per = np.array([_refine_peak(lambda x, j=j: float(scorer.scores(x, j)[0]), grid, curves[:, j], xatol)
                    for j in range(curves.shape[1])])
    step = (hi - lo) / (n_grid - 1)
    centre = float(np.median(per))
    elapsed = time.perf_counter() - t0
    log(f"dispersion estimate {value:.5g} ({scorer.n_evals} evaluations, {elapsed * 1e3:.1f} ms)")
    return {"value": value, "metric": getattr(scorer, "metric", "top_percentile"), "bounds": [lo, hi],
            "candidates": grid.tolist(), "score_curve": mean_curve.tolist(), "score_curves_per_bscan": unit.T.tolist(),
            "per_bscan_values": per.tolist(), "median_per_bscan": centre,
            "spread_mad": float(1.4826 * np.median(np.abs(per - centre))),
            "at_bound": bool(value <= lo + step or value >= hi - step),
            "n_evaluations": scorer.n_evals, "elapsed_s": elapsed}


# --------------------------------------------------------------------------- a volume's B-scans
def _central(centres):
    """Indices of the tiles near the middle of one grid axis: within 0.3 of the axis span of
    the mean centre (all of them for at most two tiles; the one nearest the mean if none)."""
    c = np.asarray(centres, float)
    if c.size <= 2:
        return list(range(c.size))
    near = np.flatnonzero(np.abs(c - c.mean()) <= 0.3 * np.ptp(c) + 1e-9)
    return near.tolist() if near.size else [int(np.argmin(np.abs(c - c.mean())))]


def select_bscans(scan, n: int, rng_seed: int = 0, has_tile=None):
    """Candidate B-scans (xi, yi, zi, frame) of a tiled scan for the dispersion search, as the
    reference picks them (octrecon/estimation/dispersion.py:111-142): focus depths inside the
    tissue (z > 0, else z >= 0, else all), tiles drawn at random among the central ones of each
    axis, frames spread over 10-90 % of the tile, no (tile, frame) twice; at most 4 n draws.
    `scan` is a ScanGrid; has_tile(xi, yi, zi) tells whether a tile exists (default: all do).
    The random draws follow the reference's order, so the same seed picks the same B-scans."""
    z = np.asarray(scan.z_depths_mm, float)
    tissue = np.flatnonzero(z > 1e-9)
    if tissue.size == 0:
        tissue = np.flatnonzero(z >= -1e-9)
    depths = tissue.tolist() if tissue.size else list(range(z.size))
    xs, ys = _central(scan.x_centers_mm), _central(scan.y_centers_mm)
    nf = int(scan.n_y_px)
    spread = np.unique(np.round(np.linspace(0.1, 0.9, max(n, 2)) * (nf - 1)).astype(int))
    rng = np.random.default_rng(rng_seed)
    picked = []
    used = set()
    for draw in range(4 * n):
        if len(picked) >= n:
            break
        zi = depths[draw % len(depths)]
        xi, yi = int(rng.choice(xs)), int(rng.choice(ys))
        frame = int(spread[draw % len(spread)]) if nf > 1 else 0
        if (has_tile is not None and not has_tile(xi, yi, zi)) or (xi, yi, zi, frame) in used:
            continue
        used.add((xi, yi, zi, frame))
        picked.append((xi, yi, zi, frame))
    return picked


def estimate_dispersion_volume(rec, n_bscans: int = 8, initial=None, bounds=None, metric: str = "auto", tiles=None,
                               n_grid: int = 41, rng_seed: int = 0, log=None):
    """Dispersion of a tiled scan held by a Reconstructor `rec`, as the reference's
    estimate_dispersion runs it on a volume (octrecon/estimation/dispersion.py:154-267): 2 n
    candidate B-scans (select_bscans, or `tiles` = [(xi, yi, zi, frame)] used as given), each
    screened by its signal score at the initial dispersion (one launch for all), the n strongest
    kept in candidate order, then the search (estimate_dispersion) on them.  initial defaults to
    the Reconstructor's dispersion.  Returns the search's keys plus method, initial and
    tiles_used (tile, frame, depth, signal)."""
    log = log or (lambda *a: None)
    t0 = time.perf_counter()
    torch = rec.torch
    if rec.nb > 1 or rec.na > 1 or rec.sp.ascan_binning > 1:
        raise NotImplementedError("the device search scores B-scans without repeats or A-scan binning")
    metric_used = "top_percentile" if metric in (None, "", "auto") else metric
    if metric_used not in METRICS:
        raise ValueError(f"unknown metric {metric!r}")
    init = float(rec.cfg.dispersion_quadratic_term if initial in (None, "", "auto") else initial)
    lo, hi = _search_bounds(init, bounds)
    if tiles:
        cands = [tuple(int(v) for v in t) for t in tiles]
    else:
        cands = select_bscans(rec.scan, 2 * n_bscans, rng_seed,
                              has_tile=None if callable(rec.tiles) else (lambda xi, yi, zi: (xi, yi, zi) in rec.tiles))
    if not cands:
        raise ValueError("no B-scan to score")
    parts = [rec.tile_frames(xi, yi, zi, [fr]) for xi, yi, zi, fr in cands]
    if parts[0].dtype == torch.uint16:                 # no uint16 cat / gather on the device: move the bits as int16
        raw = torch.cat([p.view(torch.int16) for p in parts]).view(torch.uint16)
    else:
        raw = torch.cat(parts)
    scorer = HipDispersionScorer(rec.sg.lambda_nm, raw, rec.sp.raw_apod_size, rec.xp, n_medium=rec.cfg.n_medium,
                                 apod_mode=rec.sp.apod_mode, metric=metric_used)
    sig = signal_scores(scorer.magnitude_at(init), rec.xp)
    keep = np.arange(len(cands)) if tiles else np.argsort(sig)[::-1][:n_bscans]
    keep = sorted(int(i) for i in keep)
    if len(keep) < len(cands):
        idx = torch.as_tensor(np.asarray(keep, np.int64), device=raw.device)
        raw = (torch.index_select(raw.view(torch.int16), 0, idx).view(torch.uint16) if raw.dtype == torch.uint16
               else torch.index_select(raw, 0, idx))
    scorer.set_raw(raw)
    scorer.n_evals = 0
    used = [{"tile": list(cands[i][:3]), "frame": cands[i][3],
             "z_depth_mm": float(rec.scan.z_depths_mm[cands[i][2]]), "signal_dB": float(sig[i])} for i in keep]
    for u in used:
        log(f"  B-scan of tile {u['tile']} frame {u['frame']} (signal {u['signal_dB']:.1f} dB)")
    res = estimate_dispersion(scorer, init, bounds=(lo, hi), n_grid=n_grid, log=log)
    res.update(method=f"sharpness search ({metric_used}) on {len(keep)} B-scans; coarse grid {n_grid} points + "
                      "bounded refinement of the mean normalised score",
               initial=init, tiles_used=used, elapsed_s=time.perf_counter() - t0)
    return res
