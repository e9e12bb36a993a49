"""Focus-depth detection of a tiled scan: the legacy two-step detector, the per-depth drift
measurement and the stationary confocal peak, over depth profiles taken on the device.

Reference: octrecon/estimation/focus.py (imgaussfilt_1d :44-52, findpeaks_matlab :55-72,
legacy_frames :84-86, legacy_focus :132-172, detect_peak_near :191-206, interface_pix :209-218,
stationary_confocal_peak :282-314, detect_focus :318-388).

The heavy part of every rule is the depth profile: the median over x of |A| of a few B-scans,
optionally after the optical-path correction.  Those come from the device (the spectral kernel,
repeat averaging, then oct_depth_median / oct_depth_median_opc, exact medians in float64); the
1-D steps that follow (dB, Gaussian detrending, MATLAB-style peak picking, medians over frames
or depths) act on a few thousand values and stay on the host, as in the reference.
"""
from __future__ import annotations

import time
import warnings

import numpy as np

DRIFT_MIN_DEPTH_UM = 50.0


def _mround(x):
    """MATLAB round: halves away from zero."""
    x = np.asarray(x, float)
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


# --------------------------------------------------------------------------- 1-D MATLAB helpers
def imgaussfilt_1d(v, sigma=4.0):
    """MATLAB imgaussfilt of a vector: a normalised Gaussian of 2 ceil(2 sigma) + 1 taps applied
    with replicated edges (the output has the input's length)."""
    from scipy.ndimage import correlate1d
    half = int(np.ceil(2 * sigma))
    taps = np.exp(-np.arange(-half, half + 1) ** 2 / (2.0 * sigma * sigma))
    return correlate1d(np.asarray(v, float).ravel(), taps / taps.sum(), mode="nearest")


def findpeaks_matlab(v):
    """0-based positions of MATLAB findpeaks maxima: runs of equal values higher than the runs
    on both sides, reported at the run's first sample; runs touching either end never count.
    NaN compares false, so it is never a peak and never lets a neighbour be one."""
    v = np.asarray(v, float).ravel()
    if v.size < 3:
        return np.zeros(0, int)
    starts = np.concatenate([[0], np.flatnonzero(v[1:] != v[:-1]) + 1])
    level = v[starts]
    inner = np.arange(1, starts.size - 1)                # runs that touch neither end
    rise = level[inner] > level[inner - 1]
    fall = level[inner] > level[inner + 1]
    return starts[inner[rise & fall]].astype(int)


def detect_peak_near(profile, guess_pix, search_px=60, baseline_sigma=30.0, smooth_sigma=4.0):
    """Confocal peak of a depth profile near a 1-based guess: the profile in dB (non-positive and
    non-finite samples bridged linearly; fewer than 10 usable samples give no peak), detrended as
    a sigma-4 minus a sigma-30 Gaussian smoothing, then among the detrended maxima within
    +-search_px of the guess the most prominent (the first of equals).
    Returns (1-based pixel or nan, detrended height in dB)."""
    prof = np.asarray(profile, float).ravel()
    ok = np.isfinite(prof) & (prof > 0)
    if np.count_nonzero(ok) < 10:
        return np.nan, 0.0
    where = np.flatnonzero(ok)
    db = np.interp(np.arange(prof.size), where, 10.0 * np.log10(prof[where]))
    detr = imgaussfilt_1d(db, smooth_sigma) - imgaussfilt_1d(db, baseline_sigma)
    cand = findpeaks_matlab(detr)                         # 0-based
    cand = cand[np.abs(cand + 1 - guess_pix) <= search_px]
    if cand.size == 0:
        return np.nan, 0.0
    top = int(cand[np.argmax(detr[cand])])
    return top + 1, float(detr[top])


def depth_profile(mag):
    """Mean over frames of the median over x of |A|: (F, nX, nZ) -> (nZ,)."""
    m = np.asarray(mag, float)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        return np.nanmean(np.nanmedian(m, axis=1), axis=0)


def interface_from_profiles(profiles):
    """Step-1 rule on x-median depth profiles (F, nZ): per frame the brightest row below the top
    third (the first of equal maxima, 1-based; row 1 if nothing below is finite), then the
    MATLAB-rounded median over frames.  Returns (pixel, per-frame pixels)."""
    P = np.array(profiles, float, ndmin=2)
    skip = int(_mround(P.shape[-1] / 3))
    below = P[:, skip:]
    fin = np.isfinite(below)
    masked = np.where(fin, below, -np.inf)
    per = [int(np.argmax(m)) + skip + 1 if f.any() else 1 for m, f in zip(masked, fin)]
    return int(_mround(np.median(per))), per


def interface_pix(mag):
    """interface_from_profiles of host magnitudes (F, nX, nZ) (x-nanmedian per frame)."""
    m = np.asarray(mag, float)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        return interface_from_profiles(np.nanmedian(m, axis=1))


def legacy_frames(n_y: int):
    """The legacy detector's five frames unique(round(linspace(1, nY, 5))), 0-based."""
    return sorted({int(f) - 1 for f in _mround(np.linspace(1, n_y, 5))})


# --------------------------------------------------------------------------- device profiles
def _profiles(rec, key, frames, opc: bool):
    """(F, nZ) float64 host x-median depth profiles of tile `key`'s frames at the volume's
    dispersion (repeats averaged), optical-path corrected (NaN-aware) when opc and the volume
    has a correction."""
    from .drift import depth_profiles
    mag = rec.bscan_magnitudes(*key, frames)
    tables = rec.opc_tables() if opc else None
    return depth_profiles(rec.xp, mag, frames if tables is not None else None, tables)


def _central_xy(scan):
    """Tile indices nearest the scan origin (the reference's central tile)."""
    return int(np.argmin(np.abs(np.asarray(scan.x_centers_mm)))), int(np.argmin(np.abs(np.asarray(scan.y_centers_mm))))


def _tile_of_depth(rec, zi: int, tile_selection: str):
    """The tile the legacy steps read at depth zi: the central one ('center'), or the legacy
    quirk 'legacy' (the zi-th folder of a z-innermost scan order: the first x/y tile), falling
    back to any tile of that depth."""
    nz = len(rec.scan.z_depths_mm)
    if tile_selection == "legacy":
        k = (0, 0, min(zi, nz - 1))
    elif tile_selection == "center":
        k = (*_central_xy(rec.scan), zi)
    else:
        raise ValueError("tile_selection must be center|legacy")
    if callable(rec.tiles) or k in rec.tiles:
        return k
    return next(t for t in sorted(rec.tiles) if t[2] == k[2])


def legacy_focus(rec, tile_selection: str = "center", log=None):
    """Steps 1 and 2 of the legacy detector (octrecon/estimation/focus.py:132-172), on the
    device-reconstructed B-scans of a Reconstructor `rec`:

    1. at the focus depth nearest z = 0, five frames spread over the tile: per frame the
       brightest x-median row below the top third; the rounded median -> focus1;
    2. one depth shallower (the gel): the frames' x-median profiles averaged, smoothed
       (imgaussfilt sigma 4), and the MATLAB peak nearest focus1 (focus1 if there is none).
    Neither step applies the optical-path correction (the legacy tool does not).  Pixels are
    1-based.  Returns {step1, step2, focus} with the reference's fields."""
    log = log or (lambda *a: None)
    zd = np.asarray(rec.scan.z_depths_mm, float)
    i1 = int(np.argmin(np.abs(zd)))
    i2 = min(max(1, i1 - 1), zd.size - 1)
    frames = legacy_frames(rec.scan.n_y_px)
    k1 = _tile_of_depth(rec, i1, tile_selection)
    k2 = _tile_of_depth(rec, i2, tile_selection)
    res = legacy_from_profiles(_profiles(rec, k1, frames, opc=False), _profiles(rec, k2, frames, opc=False))
    res["step1"].update(tile=list(k1), z_depth_mm=float(zd[i1]), frames=frames)
    res["step2"].update(tile=list(k2), z_depth_mm=float(zd[i2]), frames=frames)
    log(f"focus step 1: tile {k1} brightest pixel per frame {res['step1']['per_frame_pix']} -> "
        f"{res['step1']['value']}; step 2: tile {k2} nearest smoothed peak -> {res['focus']}")
    return res


def legacy_from_profiles(prof1, prof2):
    """The two legacy steps given the x-median depth profiles (F, nZ) of the step-1 tile's frames
    (prof1) and of the step-2 tile's frames (prof2); see legacy_focus."""
    focus1, per_frame = interface_from_profiles(prof1)
    mean2 = np.asarray(prof2, float).mean(axis=0)
    smooth = imgaussfilt_1d(mean2, 4)
    peaks = findpeaks_matlab(smooth) + 1
    focus2 = int(peaks[np.argmin(np.abs(peaks - focus1))]) if peaks.size else focus1
    return {"step1": {"per_frame_pix": per_frame, "value": focus1,
                      "profile": np.asarray(prof1, float).mean(axis=0).tolist()},
            "step2": {"profile": mean2.tolist(), "profile_smoothed": smooth.tolist(), "peaks": peaks.tolist(),
                      "value": focus2},
            "focus": focus2}


def stationary_confocal_peak(rec, tiles_xy=None, frames=None, log=None):
    """Diagnostic (octrecon/estimation/focus.py:282-314): the confocal peak is the one depth
    feature that does not move with the stage.  Per focus depth the dB of the mean over the
    given x/y tiles (default: the central one) of their frames' mean optical-path corrected
    x-median profiles; the pointwise median over depths keeps what stays put; that median
    (gaps bridged), detrended as sigma 4 minus sigma 50, has its most prominent MATLAB peak below
    the top third taken (else its largest value there).  Pixels are 1-based."""
    log = log or (lambda *a: None)
    xy = [tuple(t) for t in (tiles_xy or [_central_xy(rec.scan)])]
    nF = rec.scan.n_y_px
    frames = list(frames) if frames else [int(round(f * (nF - 1))) for f in (0.25, 0.5, 0.75)]
    per_depth = []
    for zi in range(len(rec.scan.z_depths_mm)):
        keys = [(x, y, zi) for x, y in xy if callable(rec.tiles) or (x, y, zi) in rec.tiles]
        per_depth.append([_profiles(rec, k, frames, opc=True) for k in keys])
    res = stationary_from_profiles(per_depth)
    log(f"stationary confocal peak (median over {res['depths']} depths): {res['value']}")
    res.update(tiles_xy=[list(map(int, t)) for t in xy], frames=frames)
    return res


def stationary_from_profiles(per_depth):
    """The stationary confocal peak given, per focus depth, a list over tiles of the frames'
    optical-path corrected x-median profiles (F, nZ); see stationary_confocal_peak."""
    db = []
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        for tiles in per_depth:
            if len(tiles):
                db.append(20 * np.log10(np.nanmean([np.nanmean(np.asarray(t, float), axis=0) for t in tiles], axis=0)))
        stay = np.nanmedian(np.array(db), axis=0)
    fin = np.flatnonzero(np.isfinite(stay))
    stay = np.interp(np.arange(stay.size), fin, stay[fin])
    d = imgaussfilt_1d(stay, 4) - imgaussfilt_1d(stay, 50)
    skip = int(_mround(d.size / 3))
    pk = findpeaks_matlab(d) + 1
    pk = pk[pk > skip]
    best = int(pk[np.argmax(d[pk - 1])]) if pk.size else int(np.argmax(d[skip:]) + skip + 1)
    return {"value": best, "prominence_dB": float(d[best - 1]), "profile_detrended": d.tolist(), "depths": len(db)}


def detect_focus(rec, mode: str = "auto", tile_selection: str = "center", confocal_check: bool = True,
                 tissue_ri: float = 1.4, log=None):
    """Focus pixel (1-based) of every focus depth of the scan held by `rec`, as the reference's
    detect_focus (octrecon/estimation/focus.py:318-388) at the volume's dispersion:

    mode 'auto' | 'measure' | 'drift': the drift measurement on the central tiles
    (Reconstructor.measure_focus_drift, seeded with the legacy step-1 pixel); 'legacy': the
    legacy step-2 pixel for every depth.  The legacy steps are always run and returned, and with
    confocal_check (and at least 3 depths) the stationary confocal peak as a diagnostic."""
    log = log or (lambda *a: None)
    t0 = time.perf_counter()
    modes = {"auto": "measure", "measure": "measure", "drift": "measure", "legacy": "legacy"}
    if mode not in modes:
        raise ValueError("mode must be auto|measure|drift|legacy")
    used = modes[mode]
    zd = np.asarray(rec.scan.z_depths_mm, float)
    depth_um = float(max(zd.max(), 0.0) * 1e3)
    leg = legacy_focus(rec, tile_selection, log)
    dz_um = float(np.median(np.diff(rec.sg.z_um)))
    out = {"step1": leg["step1"]["value"], "step2": leg["step2"]["value"],
           "legacy": {"step1": leg["step1"], "step2": leg["step2"]},
           "dispersion_quadratic_term": float(rec.cfg.dispersion_quadratic_term), "z_pixel_size_um": dz_um,
           "n_z": rec.sp.nz, "z_depths_mm": zd.tolist(), "positive_depth_range_um": depth_um,
           "tile_selection": tile_selection, "device": "hip"}
    if used == "legacy":
        focus = [leg["focus"]] * zd.size
        out["method"] = "legacy detector (step 1 interface pixel, step 2 nearest smoothed gel peak), constant"
        out["per_depth"] = [{"zi": i, "z_depth_mm": float(z), "focus_pix": leg["focus"], "source": "constant (legacy)"}
                            for i, z in enumerate(zd)]
    else:
        f, diag, per = rec.measure_focus_drift(leg["step1"]["value"], tissue_ri=tissue_ri, log=log)
        focus = [int(v) for v in f]
        deep = depth_um >= DRIFT_MIN_DEPTH_UM - 1e-6
        out["method"] = ("focus drift measured on the central tiles " +
                         ("every 50 um from 50 um" if deep else "at z = 0 (shallow scan)") + " + drift fit")
        out["drift"] = {"measurements": per, "fit": diag}
        out["per_depth"] = [{"zi": i, "z_depth_mm": float(z), "focus_pix": focus[i],
                             "se_pix": diag["seAtDepth_pix"][i], "source": "drift fit"} for i, z in enumerate(zd)]
    if confocal_check and zd.size >= 3:
        try:
            out["confocal_peak"] = stationary_confocal_peak(rec, log=log)
        except Exception as e:                     # a diagnostic: its failure does not stop the detector
            out["confocal_peak"] = {"error": str(e)}
    out["focus_positions"] = [int(v) for v in focus]
    out["mode"] = used
    out["elapsed_s"] = time.perf_counter() - t0
    log(f"focus positions: {out['focus_positions']} ({out['elapsed_s']:.2f} s)")
    return out
