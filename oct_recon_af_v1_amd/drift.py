"""Focus drift over the focus depths of a tiled scan.

Reference: octrecon/estimation/focus.py:175-189 (which depths to measure), :209-279 (per-depth
measurement on the central tile), octrecon/estimation/drift_fit.py (robust line through the
measurements, the port of yOCTMeasureFocusDrift_fitDrift.m).

Per measured depth the device reconstructs the central tile's central B-scans at the volume's
dispersion and takes the optical-path corrected x-median depth profile of each
(oct_depth_median_opc); the host averages the profiles over the B-scans and finds the confocal
peak nearest the running prediction (or, at the top depth of a shallow scan, the brightest
interface pixel).  The measurements ("clicks") then get a Theil-Sen line with iterative outlier
rejection, clamped to the physically possible slopes, evaluated at every depth with no drift
above the tissue (z <= 0).
"""
from __future__ import annotations

import warnings

import numpy as np

from . import _lib
from .estimation import depth_median
from .focus import detect_peak_near, interface_from_profiles

REJECT_FLOOR_PIX = 15          # smallest residual threshold of the outlier rejection (pixels)
MAX_REJECT_ITERATIONS = 6
THEIL_SEN_EFFICIENCY = 0.91    # relative efficiency of the Theil-Sen slope (standard errors)


def _mround(x):
    """MATLAB round (half away from zero)."""
    x = np.asarray(x, float)
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def theil_sen(z, pix):
    """(slope, intercept, has_slope): the median of the slopes of every pair of points at distinct
    depths, and the median of pix - slope z.  No pair at distinct depths: (nan, nan, False)."""
    z = np.asarray(z, float)
    p = np.asarray(pix, float)
    dz = np.subtract.outer(z, z)
    pair = np.triu(np.abs(dz) > 1e-9, 1)
    if not pair.any():
        return np.nan, np.nan, False
    slope = float(np.median(np.subtract.outer(p, p)[pair] / dz[pair]))
    return slope, float(np.median(p - slope * z)), True


def _robust_line(z, pix, notes):
    """Theil-Sen line through (z, pix) refitted after dropping points whose residual exceeds
    max(3 x 1.4826 x MAD, REJECT_FLOOR_PIX), at most MAX_REJECT_ITERATIONS times and never below
    3 points.  Returns (slope, intercept, kept indices, [(rejected index, residual)])."""
    kept = np.arange(z.size)
    dropped = []
    if z.size == 1:
        notes.append("one measurement: the focus is taken as constant")
        return 0.0, float(pix[0]), kept, dropped
    for attempt in range(MAX_REJECT_ITERATIONS + 1):
        slope, icpt, ok = theil_sen(z[kept], pix[kept])
        if not ok:
            notes.append("every measurement at one stage depth: the focus is taken as constant")
            return 0.0, float(np.median(pix[kept])), kept, dropped
        if kept.size < 3 or attempt == MAX_REJECT_ITERATIONS:
            break
        res = pix[kept] - (slope * z[kept] + icpt)
        spread = 1.4826 * np.median(np.abs(res - np.median(res)))
        far = np.abs(res) > max(3 * spread, REJECT_FLOOR_PIX)
        if not far.any():
            break
        dropped.extend(zip(kept[far].tolist(), res[far].tolist()))
        kept = kept[~far]
    return slope, icpt, kept, dropped


def _line_quality(z, pix, slope, icpt):
    """(standard error of the slope, R^2) of the line on its points; nan below 3 points (and
    for a zero spread in z or pix)."""
    n = z.size
    if n < 3:
        return np.nan, np.nan
    res = pix - (slope * z + icpt)
    sigma = np.sqrt(np.sum(res ** 2) / (n - 2))
    sxx = np.sum((z - z.mean()) ** 2)
    sst = np.sum((pix - pix.mean()) ** 2)
    se = sigma / np.sqrt(THEIL_SEN_EFFICIENCY * sxx) if sxx > 0 else np.nan
    r2 = 1 - np.sum(res ** 2) / sst if sst > 0 else np.nan
    return se, r2


def _regime(slope, used, se, ni, na, dz):
    """Drift regime of a measured slope: 'suspicious-structure-clicks' when it is significantly
    negative (the measurements followed a fixed structure: the tissue index cannot be below the
    immersion medium's), 'no-measurable-drift' within 2 standard errors of 0,
    'clamped-to-max' above the physical maximum, else 'normal'."""
    if np.isnan(se) or se <= 0:
        ratio = np.nan
        structure = slope < -0.5 * (ni / na) / dz
    else:
        ratio = slope / se
        structure = ratio < -3 and slope < -(0.1 * (ni / na) / dz)
    if structure:
        return "suspicious-structure-clicks"
    if abs(ratio) < 2:                      # false for nan
        return "no-measurable-drift"
    if used != slope and slope > 0:
        return "clamped-to-max"
    return "normal"


def fit_drift(z_clicked_mm, focus_clicked_pix, z_all_mm, z_pixel_um, n_z, tissue_ri=1.4, immersion_ri=1.33,
              max_tissue_ri=1.55):
    """Focus pixel (1-based int) at every depth of z_all_mm, and diagnostics, from focus pixels
    measured at stage depths z_clicked_mm: the model of octrecon/estimation/drift_fit.py:46-235.

    Measurements with NaN or at stage z < 0 (above the tissue the focus does not move) are left
    out.  A robust Theil-Sen line (_robust_line) is clamped to the physically possible slopes
    0 .. (n_max^2 - n_i^2) / (n_i n_a) / dz pixels per um, classified (_regime), and evaluated
    with no drift above z = 0; its standard error at each depth uses the pi / 2 variance of a
    median and the Theil-Sen efficiency.  The diagnostics keep the reference's keys
    (slopeMeasured_pixPerUm, slopeUsed_pixPerUm, driftSlope, intercept_pix, tissueRI, keptIdx,
    rejectedIdx, rejectedReason, wasSlopeClamped, driftRegime, seSlope_pixPerUm, seTissueRI,
    r2, seAtDepth_pix, maxMeasuredZ_mm, zPixelSize_um, messages)."""
    na, ni, nmax = float(tissue_ri), float(immersion_ri), float(max_tissue_ri)
    dz = float(z_pixel_um)
    nZ = int(_mround(n_z))
    if dz <= 0 or nZ < 1:
        raise ValueError("z_pixel_um must be > 0 and n_z >= 1")
    notes = []
    z_um = np.atleast_1d(np.asarray(z_clicked_mm, float)).ravel() * 1e3
    pc = np.atleast_1d(np.asarray(focus_clicked_pix, float)).ravel()
    missing = np.isnan(z_um) | np.isnan(pc)
    above = ~missing & (z_um < 0)
    if missing.any():
        notes.append(f"{int(missing.sum())} measurement(s) with NaN left out")
    if above.any():
        notes.append(f"{int(above.sum())} measurement(s) above the tissue (stage z < 0) left out: the focus does "
                     "not drift there")
    src = np.flatnonzero(~missing & ~above)
    if src.size == 0:
        raise ValueError("no usable focus measurement")
    z, pix = z_um[src], pc[src]

    slope, icpt, kept, dropped = _robust_line(z, pix, notes)
    se, r2 = _line_quality(z[kept], pix[kept], slope, icpt)
    m_max = (nmax ** 2 - ni ** 2) / (ni * na) / dz
    used = min(max(slope, 0.0), m_max)
    regime = _regime(slope, used, se, ni, na, dz)
    if used != slope:
        icpt = float(np.median(pix[kept] - used * z[kept]))
    if regime == "suspicious-structure-clicks":
        notes.append("the drift slope is significantly negative: the measurements followed a fixed structure, not "
                     "the focus; slope set to 0")
    elif regime == "clamped-to-max":
        notes.append(f"the drift slope exceeds the maximum for tissue index {nmax}: clamped")
    ri = np.nan if regime == "suspicious-structure-clicks" else float(np.sqrt(max(ni ** 2 + (slope * dz) * ni * na,
                                                                                    ni ** 2)))
    se_ri = (se * dz) * ni * na / (2 * ri) if (np.isfinite(ri) and np.isfinite(se) and ri > 0) else np.nan
    reasons = [f"measurement at z = {z[k] * 1e-3:.3f} mm is {r:+.0f} pixels from the drift line" for k, r in dropped]
    notes.extend("left out: " + r for r in reasons)

    z_eval = np.maximum(np.atleast_1d(np.asarray(z_all_mm, float)).ravel() * 1e3, 0.0)
    focus = _mround(used * z_eval + icpt)
    if ((focus < 1) | (focus > nZ)).any():
        notes.append("the focus leaves the image at some depths: clamped to it")
    focus = np.clip(focus, 1, nZ).astype(int)
    nk = kept.size
    if nk >= 3:
        zk = z[kept]
        res = pix[kept] - (used * zk + icpt)
        sigma = np.sqrt(np.sum(res ** 2) / (nk - 2))
        zbar = zk.mean()
        sxx = np.sum((zk - zbar) ** 2)
        se_depth = (sigma * np.sqrt((np.pi / 2) / nk + (z_eval - zbar) ** 2 / (THEIL_SEN_EFFICIENCY * sxx)) if sxx > 0
                    else sigma * np.sqrt(np.pi / 2) * np.ones_like(z_eval) / np.sqrt(nk))
    else:
        se_depth = np.full(z_eval.shape, np.nan)
    diag = {
        "slopeMeasured_pixPerUm": float(slope), "slopeUsed_pixPerUm": float(used), "driftSlope": float(used * dz),
        "intercept_pix": float(icpt), "tissueRI": ri, "keptIdx": src[kept].tolist(),
        "rejectedIdx": [int(src[k]) for k, _ in dropped], "rejectedReason": reasons,
        "wasSlopeClamped": bool(used != slope), "driftRegime": regime, "seSlope_pixPerUm": float(se),
        "seTissueRI": float(se_ri), "r2": float(r2), "seAtDepth_pix": se_depth.tolist(),
        "maxMeasuredZ_mm": float(np.max(z[kept]) * 1e-3), "zPixelSize_um": dz, "messages": notes,
    }
    return focus, diag


def select_depths_to_measure(z_depths_mm, step_um=50.0, first_um=50.0):
    """0-based indices of the depths to measure: every round(step / depth spacing)-th depth from
    the first at >= first_um, in increasing depth; all depths >= 0 when none is that deep
    (returns (indices, fallback)).  octrecon/estimation/focus.py:175-188."""
    z = np.asarray(z_depths_mm, float)
    d = np.median(np.abs(np.diff(z))) if len(z) > 1 else 0.0
    stride = 1 if d <= 0 else max(1, int(_mround(step_um * 1e-3 / d)))
    cand = np.flatnonzero(z >= first_um * 1e-3 - 1e-9)
    fallback = cand.size == 0
    if fallback:
        cand = np.flatnonzero(z >= -1e-9)
        if cand.size == 0:
            raise ValueError("every focus depth is above the tissue (z < 0): no drift to measure")
    cand = cand[np.argsort(z[cand], kind="stable")]
    return cand[::stride].tolist(), bool(fallback)


def central_frames(n_frames_tile: int, n: int = 5):
    """Tile-local frames around the central B-scan (octrecon/estimation/focus.py:241-244)."""
    c = (n_frames_tile - 1) / 2
    half = max(1, int(round(0.05 * n_frames_tile)))
    f = np.round(np.linspace(c - half, c + half, n))
    return np.unique(np.clip(f, 0, n_frames_tile - 1).astype(int)).tolist()


def depth_profiles(xp, mag, frames=None, opc=None):
    """(F, nZ) float64 x-median depth profiles of magnitudes (F, nX, nZ) on the device; with
    opc = (shift, valid_lo, valid_hi) device int32 (nY, nX) tables, B-scan b is tile frame
    frames[b] and the profile is taken after the optical-path correction (NaN-free median of
    the rows that exist)."""
    if opc is None:
        return depth_median(mag, xp).cpu().numpy()
    torch = xp.torch
    shift, vlo, vhi = opc
    F, nX, nz = mag.shape
    fr = np.asarray(frames, np.int64)
    if fr.shape != (F,) or fr.min() < 0 or fr.max() >= shift.shape[0] or shift.shape[1] != nX:
        raise ValueError("frames do not match the optical-path tables")
    if mag.stride(2) != 1:
        raise ValueError("magnitude rows must be contiguous")
    fdev = torch.as_tensor(fr.astype(np.int32), device=xp.device)
    out = torch.empty((F, nz), dtype=torch.float64, device=xp.device)
    _lib.check(_lib.lib().oct_depth_median_opc(mag.data_ptr(), F, nX, mag.stride(1), nz, fdev.data_ptr(),
                                               shift.data_ptr(), vlo.data_ptr(), vhi.data_ptr(), out.data_ptr(),
                                               xp.stream()), "oct_depth_median_opc")
    return out.cpu().numpy()


def measure_focus_drift(xp, magnitude_of, frames, z_depths_mm, prior_pix, z_pixel_um, n_z, opc=None, tissue_ri=1.4,
                        step_um=50.0, first_um=50.0, search_px=60, min_prominence_db=0.3, log=None):
    """Focus pixel (1-based) per focus depth from the central tile of each measured depth.

    magnitude_of(zi): device magnitudes (len(frames), nX, n_z) of the central tile of depth zi
    at the volume's dispersion, tile-local frames `frames`; opc: optional optical-path tables
    (shift, valid_lo, valid_hi) for those frames; prior_pix: the first guess (the interface
    rule's value).  Returns (focus per depth, fit diagnostics, per-depth measurements), as
    octrecon/estimation/focus.py:221-279."""
    log = log or (lambda *a: None)
    zd = np.asarray(z_depths_mm, float)
    idxs, fallback = select_depths_to_measure(zd, step_um, first_um)
    step_mm = float(np.median(np.abs(np.diff(zd)))) if len(zd) > 1 else 0.0
    nz = int(n_z)
    clicks_z, clicks_pix, per = [], [], []
    guess = float(prior_pix)
    for zi in idxs:
        z = float(zd[zi])
        mag = magnitude_of(zi)
        med = depth_profiles(xp, mag, frames, opc)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", RuntimeWarning)
            prof = np.nanmean(med, axis=0)
        if abs(z) <= 0.5 * step_mm + 1e-9:
            pix, frame_vals = interface_from_profiles(med)
            rule, prom, acc = "interface", None, True
        else:
            pix, prom = detect_peak_near(prof, guess, search_px)
            frame_vals, rule = None, "confocal peak"
            acc = bool(np.isfinite(pix) and prom >= min_prominence_db)
        per.append({"zi": int(zi), "z_depth_mm": z, "rule": rule,
                    "detected_pix": None if not np.isfinite(pix) else int(pix), "per_frame_pix": frame_vals,
                    "prominence_dB": prom, "accepted": acc})
        log(f"focus drift: z = {z * 1e3:.0f} um [{rule}] -> {pix}{'' if acc else ' (not used)'}")
        if acc:
            clicks_z.append(z)
            clicks_pix.append(float(pix))
            if len(clicks_z) >= 2:
                guess = float(fit_drift(clicks_z, clicks_pix, [z + step_um * 1e-3], z_pixel_um, nz,
                                        tissue_ri=tissue_ri)[0][0])
            else:
                guess = float(pix)
    if not clicks_z:
        raise RuntimeError("focus drift: the focus was not found at any measured depth")
    focus, diag = fit_drift(clicks_z, clicks_pix, zd, z_pixel_um, nz, tissue_ri=tissue_ri)
    diag["fallback_shallow"] = fallback
    return focus, diag, per
