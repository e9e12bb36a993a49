"""Oracle (test infrastructure only): the reference's optical-path corrected depth profile,
restated in numpy (octrecon/estimation/focus.py:102-129 apply_opc, :260-262 the x-nanmedian).
Not used by the product."""
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
