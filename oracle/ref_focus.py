"""Oracle (test infrastructure only): the reference's focus detector restated in numpy: the
optical-path corrected depth profile (octrecon/estimation/focus.py:102-129 apply_opc, :260-262
the x-nanmedian), the 1-D MATLAB helpers, the legacy and stationary-peak rules (focus.py:44-72,
132-172, 191-218, 282-314) and the drift fit (octrecon/estimation/drift_fit.py:30-221).
Pinned by the reference's own known answers and closed forms (tests/test_focus_host.py,
tests/test_drift.py).  Not used by the product."""
import warnings

import numpy as np


def opc_depth_median(mag, frames, shift, pdz):
    """mag (F, nX, nZ) float32, frames (F) tile frame of each B-scan, shift / pdz (nY, nX) ->
    (F, nZ) float64 nanmedian over x of the corrected magnitudes (NaN where r + pdz leaves
    [0, nZ - 1])."""
    m = np.asarray(mag, float)
    nz = m.shape[-1]
    rr = np.arange(nz)
    out = np.empty((m.shape[0], nz))
    for i, f in enumerate(frames):
        src = np.clip(rr[None, :] + shift[f][:, None], 0, nz - 1)
        t = rr[None, :] + pdz[f][:, None]
        g = np.take_along_axis(m[i], src, axis=-1)
        g = np.where((t >= 0) & (t <= nz - 1), g, np.nan)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", RuntimeWarning)
            out[i] = np.nanmedian(g, axis=0)
    return out


# ----------------------------------------------------------------------------- 1-D helpers
# Loop-form restatements of the reference's MATLAB helpers and rules
# (octrecon/estimation/focus.py:44-72, 84-86, 132-172, 191-218, 282-314), the checker of the
# product's vectorised versions (oct_recon_af_v1_amd/focus.py).

def mround(x):
    x = np.asarray(x, float)
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def imgaussfilt_1d(v, sigma=4.0):
    """octrecon/estimation/focus.py:44-52: 2 ceil(2 sigma) + 1 taps, replicate padding, 'valid'."""
    v = np.asarray(v, float).ravel()
    h = int(np.ceil(2 * sigma))
    x = np.arange(-h, h + 1)
    k = np.exp(-x ** 2 / (2 * sigma ** 2))
    k /= k.sum()
    p = np.concatenate([np.full(h, v[0]), v, np.full(h, v[-1])])
    return np.convolve(p, k, mode="valid")


def findpeaks_matlab(v):
    """octrecon/estimation/focus.py:55-72 (scan with plateau handling)."""
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
    """octrecon/estimation/focus.py:191-206."""
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


def interface_from_profiles(profiles):
    """octrecon/estimation/focus.py:209-218 on x-median profiles (F, nZ), 1-based."""
    vals = []
    for med in np.array(profiles, float, ndmin=2):
        med = med.copy()
        med[: int(mround(med.size / 3))] = np.nan
        vals.append(int(np.nanargmax(med)) + 1)
    return int(mround(np.median(vals))), vals


def legacy_frames(n_y):
    return (np.unique(mround(np.linspace(1, n_y, 5))).astype(int) - 1).tolist()


def legacy_focus_from_profiles(prof1, prof2):
    """octrecon/estimation/focus.py:143-163 given the x-median profiles (F, nZ) of the step-1 tile
    (depth nearest 0) and of the step-2 tile (one depth shallower)."""
    focus1, _ = interface_from_profiles(prof1)
    prof = np.asarray(prof2, float).mean(axis=0)
    smooth = imgaussfilt_1d(prof, 4)
    p = findpeaks_matlab(smooth) + 1
    return focus1, (int(mround(p[int(np.argmin(np.abs(p - focus1)))])) if len(p) else focus1)


def stationary_peak_from_profiles(per_depth_tile_profiles):
    """octrecon/estimation/focus.py:291-311 given, per depth, a list over tiles of the
    optical-path corrected x-nanmedian profiles (F, nZ) of the frames."""
    import warnings
    P = []
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        for tiles in per_depth_tile_profiles:
            acc = [np.nanmean(np.asarray(t, float), axis=0) for t in tiles]
            if acc:
                P.append(20 * np.log10(np.nanmean(acc, axis=0)))
        med = np.nanmedian(np.array(P), axis=0)
    med = np.interp(np.arange(len(med)), np.nonzero(np.isfinite(med))[0], med[np.isfinite(med)])
    d = imgaussfilt_1d(med, 4) - imgaussfilt_1d(med, 50)
    lo = int(mround(len(d) / 3))
    pk = findpeaks_matlab(d) + 1
    pk = pk[pk > lo]
    return int(pk[np.argmax(d[pk - 1])]) if pk.size else int(np.argmax(d[lo:]) + lo + 1)


# ----------------------------------------------------------------------------- drift fit
def fit_drift(z_clicked_mm, pix_clicked, z_all_mm, dz_um, n_z, na=1.4, ni=1.33, ns_max=1.55):
    """octrecon/estimation/drift_fit.py:46-221 (numbers only: focus, slopes, intercept, regime,
    tissue index, kept / rejected indices, standard errors, R^2)."""
    nZ = int(mround(n_z))
    zc = np.atleast_1d(np.asarray(z_clicked_mm, float)).ravel() * 1e3
    pc = np.atleast_1d(np.asarray(pix_clicked, float)).ravel()
    bad = np.isnan(zc) | np.isnan(pc)
    use = np.nonzero(~bad & ~((zc < 0) & ~bad))[0]
    z, pix = zc[use], pc[use]
    n = len(z)
    m_max = (ns_max ** 2 - ni ** 2) / (ni * na) / dz_um
    kept = np.arange(n)
    rej = np.zeros(0, int)

    def ts(zz, pp):
        ii, jj = np.triu_indices(len(zz), k=1)
        d = zz[jj] - zz[ii]
        v = np.abs(d) > 1e-9
        if not np.any(v):
            return np.nan, np.nan, False
        s = float(np.median((pp[jj[v]] - pp[ii[v]]) / d[v]))
        return s, float(np.median(pp - s * zz)), True

    if n == 1:
        slope, icpt = 0.0, float(pix[0])
    else:
        for it in range(1, 8):
            slope, icpt, has = ts(z[kept], pix[kept])
            if not has:
                slope, icpt = 0.0, float(np.median(pix[kept]))
                break
            if len(kept) < 3 or it > 6:
                break
            r = pix[kept] - (slope * z[kept] + icpt)
            s_mad = 1.4826 * np.median(np.abs(r - np.median(r)))
            out = np.abs(r) > max(3 * s_mad, 15)
            if not out.any():
                break
            rej = np.concatenate([rej, kept[out]])
            kept = kept[~out]
    nk = len(kept)
    r2 = se = np.nan
    if nk >= 3:
        rm = pix[kept] - (slope * z[kept] + icpt)
        sr = np.sqrt(np.sum(rm ** 2) / (nk - 2))
        sxx = np.sum((z[kept] - z[kept].mean()) ** 2)
        se = sr / np.sqrt(0.91 * sxx) if sxx > 0 else np.nan
        sst = np.sum((pix[kept] - pix[kept].mean()) ** 2)
        if sst > 0:
            r2 = 1 - np.sum(rm ** 2) / sst
    used = min(max(slope, 0.0), m_max)
    clamped = used != slope
    if np.isnan(se) or se <= 0:
        sig = np.nan
        structure = slope < -0.5 * (ni / na) / dz_um
    else:
        sig = slope / se
        structure = sig < -3 and slope < -(0.1 * (ni / na) / dz_um)
    regime = ("suspicious-structure-clicks" if structure else "no-measurable-drift" if abs(sig) < 2
              else "clamped-to-max" if (clamped and slope > 0) else "normal")
    if clamped:
        icpt = float(np.median(pix[kept] - used * z[kept]))
    ri = np.nan if structure else float(np.sqrt(max(ni ** 2 + (slope * dz_um) * ni * na, ni ** 2)))
    za = np.atleast_1d(np.asarray(z_all_mm, float)).ravel() * 1e3
    ze = np.maximum(za, 0)
    focus = np.clip(mround(used * ze + icpt), 1, nZ).astype(int)
    if nk >= 3:
        rk = pix[kept] - (used * z[kept] + icpt)
        sr = np.sqrt(np.sum(rk ** 2) / (nk - 2))
        zb = z[kept].mean()
        sxx = np.sum((z[kept] - zb) ** 2)
        sed = (sr * np.sqrt((np.pi / 2) / nk + (ze - zb) ** 2 / (0.91 * sxx)) if sxx > 0
               else sr * np.sqrt(np.pi / 2) * np.ones_like(ze) / np.sqrt(nk))
    else:
        sed = np.full(ze.shape, np.nan)
    return focus, {"slopeMeasured_pixPerUm": slope, "slopeUsed_pixPerUm": used, "intercept_pix": icpt,
                   "driftRegime": regime, "tissueRI": ri, "keptIdx": use[kept].tolist(), "rejectedIdx": use[rej].tolist(),
                   "seSlope_pixPerUm": se, "r2": r2, "seAtDepth_pix": sed.tolist()}
