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
from .estimation import depth_median, detect_peak_near

REJECT_FLOOR_PIX = 15          # smallest residual threshold of the outlier rejection (pixels)
MAX_REJECT_ITERATIONS = 6


def _mround(x):
    """MATLAB round (half away from zero)."""
    x = np.asarray(x, float)
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def theil_sen(z, pix):
    """(slope, intercept, has_slope): median of the slopes of all pairs at distinct z, then the
    median of pix - slope z."""
    z = np.asarray(z, float)
    pix = np.asarray(pix, float)
    i, j = np.triu_indices(len(z), 1)
    dz = z[j] - z[i]
    use = np.abs(dz) > 1e-9
    if not use.any():
        return np.nan, np.nan, False
    slope = float(np.median((pix[j][use] - pix[i][use]) / dz[use]))
    return slope, float(np.median(pix - slope * z)), True


def fit_drift(z_clicked_mm, focus_clicked_pix, z_all_mm, z_pixel_um, n_z, tissue_ri=1.4, immersion_ri=1.33,
              max_tissue_ri=1.55):
    """Focus pixel (1-based int) at every depth of z_all_mm and diagnostics, from measured focus
    pixels at stage depths z_clicked_mm (octrecon/estimation/drift_fit.py:46-300).

    Slope limits: 0 .. (max_ri^2 - n_i^2) / (n_i n_a) / dz pixels per um; a significantly
    negative measured slope means the clicks followed a fixed structure ("suspicious").
    Diagnostics use the reference's keys (driftRegime, keptIdx, rejectedIdx, slopeUsed_pixPerUm,
    driftSlope, tissueRI, seAtDepth_pix, ...)."""
    na, ni, nmax = float(tissue_ri), float(immersion_ri), float(max_tissue_ri)
    dz = float(z_pixel_um)
    nZ = int(_mround(n_z))
    if dz <= 0 or nZ < 1:
        raise ValueError("z_pixel_um must be > 0 and n_z >= 1")
    notes = []
    zc = np.atleast_1d(np.asarray(z_clicked_mm, float)).ravel() * 1e3
    pc = np.atleast_1d(np.asarray(focus_clicked_pix, float)).ravel()
    nan = np.isnan(zc) | np.isnan(pc)
    if nan.any():
        notes.append(f"ignored {int(nan.sum())} measurement(s) with NaN values")
    above = (zc < 0) & ~nan
    if above.any():
        notes.append(f"ignored {int(above.sum())} measurement(s) at stage z < 0 (above the tissue the focus "
                     "does not drift)")
    use = np.flatnonzero(~nan & ~above)
    z, pix = zc[use], pc[use]
    if len(z) == 0:
        raise ValueError("no usable focus measurement")

    m_max = (nmax ** 2 - ni ** 2) / (ni * na) / dz
    kept = np.arange(len(z))
    out_idx, out_res = [], []
    if len(z) == 1:
        slope, icpt = 0.0, float(pix[0])
        notes.append("one measurement: constant focus")
    else:
        for round_ in range(MAX_REJECT_ITERATIONS + 1):
            slope, icpt, ok = theil_sen(z[kept], pix[kept])
            if not ok:
                slope, icpt = 0.0, float(np.median(pix[kept]))
                notes.append("all measurements at one stage depth: constant focus")
                break
            if len(kept) < 3 or round_ == MAX_REJECT_ITERATIONS:
                break
            r = pix[kept] - (slope * z[kept] + icpt)
            mad = 1.4826 * np.median(np.abs(r - np.median(r)))
            bad = np.abs(r) > max(3.0 * mad, REJECT_FLOOR_PIX)
            if not bad.any():
                break
            out_idx.extend(kept[bad].tolist())
            out_res.extend(r[bad].tolist())
            kept = kept[~bad]

    nk = len(kept)
    se_slope, r2 = np.nan, np.nan
    if nk >= 3:
        res = pix[kept] - (slope * z[kept] + icpt)
        s_res = np.sqrt(np.sum(res ** 2) / (nk - 2))
        sxx = np.sum((z[kept] - z[kept].mean()) ** 2)
        if sxx > 0:
            se_slope = s_res / np.sqrt(0.91 * sxx)          # Theil-Sen efficiency 0.91
        sst = np.sum((pix[kept] - pix[kept].mean()) ** 2)
        if sst > 0:
            r2 = 1.0 - np.sum(res ** 2) / sst

    used = min(max(slope, 0.0), m_max)
    clamped = used != slope
    unit = (ni / na) / dz
    if np.isnan(se_slope) or se_slope <= 0:
        sig = np.nan
        structure = slope < -0.5 * unit
    else:
        sig = slope / se_slope
        structure = sig < -3 and slope < -0.1 * unit
    if structure:
        regime = "suspicious-structure-clicks"
        notes.append("significantly negative drift: the measurements followed a fixed structure; slope set to 0")
    elif abs(sig) < 2:
        regime = "no-measurable-drift"
    elif clamped and slope > 0:
        regime = "clamped-to-max"
        notes.append(f"drift above the physical maximum (tissue index {nmax}): clamped")
    else:
        regime = "normal"
    if clamped:
        icpt = float(np.median(pix[kept] - used * z[kept]))

    ns2 = ni ** 2 + slope * dz * ni * na
    ri = np.nan if structure else float(np.sqrt(max(ns2, ni ** 2)))
    se_ri = (se_slope * dz) * ni * na / (2 * ri) if (np.isfinite(ri) and np.isfinite(se_slope) and ri > 0) else np.nan
    for k, rr in zip(out_idx, out_res):
        notes.append(f"rejected the measurement at z = {z[k] * 1e-3:.3f} mm ({rr:+.0f} pixels from the line)")

    za = np.atleast_1d(np.asarray(z_all_mm, float)).ravel() * 1e3
    zeff = np.maximum(za, 0.0)
    focus = _mround(used * zeff + icpt)
    if ((focus < 1) | (focus > nZ)).any():
        notes.append("focus clamped to the image at some depths")
    focus = np.clip(focus, 1, nZ).astype(int)
    if nk >= 3:
        res = pix[kept] - (used * z[kept] + icpt)
        s_res = np.sqrt(np.sum(res ** 2) / (nk - 2))
        zbar = z[kept].mean()
        sxx = np.sum((z[kept] - zbar) ** 2)
        se_depth = (s_res * np.sqrt((np.pi / 2) / nk + (zeff - zbar) ** 2 / (0.91 * sxx)) if sxx > 0
                    else np.full(zeff.shape, s_res * np.sqrt(np.pi / 2) / np.sqrt(nk)))
    else:
        se_depth = np.full(zeff.shape, np.nan)
    diag = {
        "slopeMeasured_pixPerUm": float(slope), "slopeUsed_pixPerUm": float(used), "driftSlope": float(used * dz),
        "intercept_pix": float(icpt), "tissueRI": ri, "keptIdx": use[kept].tolist(),
        "rejectedIdx": use[np.asarray(out_idx, int)].tolist(), "wasSlopeClamped": bool(clamped),
        "driftRegime": regime, "seSlope_pixPerUm": float(se_slope), "seTissueRI": float(se_ri), "r2": float(r2),
        "seAtDepth_pix": se_depth.tolist(), "maxMeasuredZ_mm": float(np.max(z[kept]) * 1e-3),
        "zPixelSize_um": dz, "messages": notes,
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


def _interface_pix(med):
    """Per frame the brightest pixel of the x-median profile below the top third (first of
    equal maxima), 1-based; the median over frames (octrecon/estimation/focus.py:209-218)."""
    nz = med.shape[-1]
    vals = []
    for m in med:
        m = m.copy()
        m[: int(_mround(nz / 3))] = np.nan
        vals.append(int(np.nanargmax(m)) + 1 if np.isfinite(m).any() else 1)
    return int(_mround(np.median(vals))), vals


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
            pix, frame_vals = _interface_pix(med)
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
