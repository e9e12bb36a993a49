"""Host-side (float64) spectral and spatial geometry of a tiled OCT scan.

These builders produce the small operators the device kernels consume: the banded
sinc k-linearisation operator, the complex Hann/rms x dispersion window, the depth
axis, and the tile/output grids used by stitching.  Everything here runs once per
reconstruction on the host in float64.

Reference interfaces mirrored (octrecon/core/geometry.py):
  matlab_linspace            :30-35
  matlab_colon               :38-51
  chirp_to_lambda            :57-60
  _pchip_matlab              :63-68
  SpectralGeometry           :71-81
  build_spectral_geometry    :84-128
  TiledGeometry              :134-143
  build_tiled_geometry       :146-202
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy.interpolate import PchipInterpolator

# Spectrometer wavelength span per OCT system in nm (octrecon/core/geometry.py:18-24).
LAMBDA_RANGE_NM = {
    "ganymede": (796.23, 1010.02),
    "ganymede_srr": (796.23, 1010.02),
    "gan632": (796.0, 1010.0),
    "telesto": (1208.69, 1372.50),
    "telesto_srr": (1208.69, 1372.50),
}


# --------------------------------------------------------------------------------------
# MATLAB grid constructors.  Their arithmetic has to follow MATLAB's own formulas so that
# grid nodes (and therefore interpolation weights and sinc-band membership) agree with the
# legacy reconstruction to the last bit.
# --------------------------------------------------------------------------------------
def matlab_linspace(a: float, b: float, n: int) -> np.ndarray:
    """MATLAB ``linspace(a, b, n)``: ``a + (0:n-1)*(b-a)/(n-1)`` with both ends pinned."""
    n = int(n)
    if n < 2:
        return np.array([float(b)] if n == 1 else [], dtype=float)
    out = a + (np.arange(n, dtype=float) * (b - a)) / (n - 1)
    out[0], out[-1] = a, b
    return out


def matlab_colon(a: float, d: float, b: float) -> np.ndarray:
    """MATLAB ``a:d:b``: element count with a 2-eps tolerance, first half of the
    elements counted up from ``a`` and the second half down from the end point."""
    if d == 0:
        return np.zeros(0)
    tol = 2.0 * np.finfo(float).eps * max(abs(a), abs(b))
    n = int(np.floor((b - a) / d + tol / abs(d)))
    if n < 0:
        return np.zeros(0)
    idx = np.arange(n + 1, dtype=float)
    last = a + n * d
    mid = n // 2
    out = np.empty(n + 1)
    out[: mid + 1] = a + idx[: mid + 1] * d
    out[mid + 1:] = last - (n - idx[mid + 1:]) * d
    return out


# --------------------------------------------------------------------------------------
# Wavelength maps
# --------------------------------------------------------------------------------------
def chirp_to_lambda(chirp: np.ndarray, oct_system: str) -> np.ndarray:
    """Thorlabs chirp table -> per-pixel wavelength in nm (linear in 1/lambda)."""
    lo, hi = LAMBDA_RANGE_NM[oct_system.lower()]
    c = np.asarray(chirp, dtype=float)
    step = (1.0 / hi - 1.0 / lo) / (len(c) - 1)
    return 1.0 / (c * step + 1.0 / lo)


def cubic_lambda_map(n: int = 2048, center_nm: float = 1300.0, a1: float = 0.0625,
                     a2: float = -4e-6, a3: float = 1.5e-9) -> np.ndarray:
    """Nonlinear spectrometer map lambda(p) = c + a1 u + a2 u^2 + a3 u^3 with
    u = p - (n-1)/2 (p = 0..n-1).  Default: 1230-1361 nm over 2048 pixels."""
    u = np.arange(n, dtype=float) - 0.5 * (n - 1)
    return center_nm + u * (a1 + u * (a2 + u * a3))


def pchip(x, v, xq):
    """``interp1(x, v, xq, 'pchip')`` for strictly monotonic ``x`` in either direction
    (shape-preserving Fritsch-Carlson cubic, extrapolating with the end pieces)."""
    x = np.asarray(x, dtype=float)
    v = np.asarray(v, dtype=float)
    if x[0] > x[-1]:
        x, v = x[::-1], v[::-1]
    return PchipInterpolator(x, v, extrapolate=True)(xq)


# --------------------------------------------------------------------------------------
# Spectral geometry
# --------------------------------------------------------------------------------------
@dataclass
class SpectralGeometry:
    """Same fields as the reference's SpectralGeometry, plus the contiguous-band form of
    the sinc operator that the device kernels use."""
    lambda_nm: np.ndarray      # (N,) raw per-pixel wavelength
    k: np.ndarray              # (N,) 2*pi/lambda  [1/nm]
    k_lin: np.ndarray          # (N,) equispaced target k, descending
    lambda_eq_nm: np.ndarray   # (N,) wavelength of the equispaced samples
    sinc_idx: np.ndarray       # (N, T) int64 source indices, ascending, padded with 0
    sinc_w: np.ndarray         # (N, T) float64 sinc weights, 0 on padding
    window: np.ndarray         # (N,) complex128 Hann/rms * exp(-i beta (k_eq - <k_eq>)^2)
    z_um: np.ndarray           # (N/2,) depth in the medium, z=0 at zero delay
    interp_method: str
    band_start: np.ndarray = field(default=None)   # (N,) int32 first source of a T-wide window
    band_w: np.ndarray = field(default=None)       # (N, T) float64 weights over that window
    dispersion_quadratic_term: float = 0.0
    n_medium: float = 1.0

    @property
    def n(self) -> int:
        return len(self.lambda_nm)

    @property
    def taps(self) -> int:
        return self.sinc_idx.shape[1]


def _sinc_band(k: np.ndarray, k_lin: np.ndarray, taps: int):
    """Banded sinc operator.  Output sample i takes every source j whose normalised
    k-distance |nq_i - n_j| is below ``taps``.  For a monotonic k(pixel) map those sources
    form one contiguous run of j, found here with a sorted search and then decided with the
    exact float64 distance test, so membership and weights are those of a dense N x N
    evaluation."""
    N = len(k)
    nq = (k_lin - k_lin.min()) / (k_lin.max() - k_lin.min()) * (N - 1)
    n = (k - k.min()) / (k.max() - k.min()) * (N - 1)
    d = np.diff(n)
    if not (np.all(d > 0) or np.all(d < 0)):
        raise ValueError("the wavelength map must be strictly monotonic in pixel index")
    asc = d[0] > 0
    n_sorted = n if asc else n[::-1]
    # candidate runs in sorted order with one sample of slack on both sides
    lo_s = np.searchsorted(n_sorted, nq - taps, side="left") - 1
    hi_s = np.searchsorted(n_sorted, nq + taps, side="right") + 1
    width = int((hi_s - lo_s).max()) + 1
    js = lo_s[:, None] + np.arange(width)[None, :]
    inb = (js >= 0) & (js < N)
    js_c = np.clip(js, 0, N - 1)
    j = js_c if asc else (N - 1 - js_c)                 # original pixel index
    dist = nq[:, None] - n[j]
    mask = inb & (np.abs(dist) < taps)
    count = mask.sum(1)
    if np.any(count == 0):
        raise ValueError("an output sample has no sinc taps")
    # contiguity: each row's mask must be a single run
    runs = (np.diff(mask.astype(np.int8), axis=1) == 1).sum(1) + mask[:, 0]
    if np.any(runs != 1):
        raise ValueError("sinc band is not contiguous")
    T = int(count.max())
    first_pos = np.argmax(mask, axis=1)
    rows = np.arange(N)
    # ascending source order per output (the reference's gather order)
    if asc:
        start = j[rows, first_pos]
    else:
        start = j[rows, first_pos + count - 1]
    idx = np.zeros((N, T), np.int64)
    w = np.zeros((N, T), float)
    tt = np.arange(T)[None, :]
    valid = tt < count[:, None]
    src = start[:, None] + tt
    idx[valid] = src[valid]
    w[valid] = np.sinc(nq[:, None] - n[np.clip(src, 0, N - 1)])[valid]
    # T-wide window that never leaves [0, N): shift it left at the top edge; the shifted-in
    # positions carry weight 0 and are summed first, which leaves every sum unchanged.
    bstart = np.minimum(start, N - T).astype(np.int32)
    off = start - bstart
    band_w = np.zeros((N, T), float)
    cols = off[:, None] + tt
    keep = cols < T
    band_w[np.broadcast_to(rows[:, None], (N, T))[valid & keep], cols[valid & keep]] = w[valid & keep]
    return idx, w, bstart, band_w


def build_spectral_geometry(lambda_nm, dispersion_quadratic_term: float, n_medium: float,
                            interp_method: str = "sinc5") -> SpectralGeometry:
    """Reference: octrecon/core/geometry.py:84-128."""
    lam = np.asarray(lambda_nm, dtype=float).ravel()
    N = len(lam)
    k = 2 * np.pi / lam
    k_lin = matlab_linspace(k.max(), k.min(), N)
    m = interp_method.lower()
    if m.startswith("sinc"):
        taps = int(m[4:])
        idx, w, bstart, band_w = _sinc_band(k, k_lin, taps)
    elif m == "pchip":
        raise NotImplementedError("pchip k-linearisation is not implemented (sinc5 is the "
                                  "reconstruction method of the path)")
    else:
        raise ValueError(interp_method)

    lambda_eq = pchip(k, lam, k_lin)
    window = dispersion_window(lambda_eq, dispersion_quadratic_term)

    ends = np.sort([lambda_eq[0], lambda_eq[-1]])
    lambda0_um = ends.mean() / 1e3
    dlambda_um = (ends[1] - ends[0]) / 1e3
    z_step = 0.5 * lambda0_um ** 2 / dlambda_um / n_medium
    z_um = matlab_linspace(0.0, z_step * N / 2, N // 2)
    return SpectralGeometry(lam, k, k_lin, lambda_eq, idx, w, window, z_um, interp_method,
                            bstart, band_w, float(dispersion_quadratic_term), float(n_medium))


def dispersion_window(lambda_eq_nm, dispersion_quadratic_term):
    """Hann/rms x exp(-i beta (k_eq - <k_eq>)^2) on the equispaced-k grid (complex128 (N,)),
    octrecon/core/geometry.py:115-119; beta may be an array (-> (len(beta), N))."""
    lambda_eq_nm = np.asarray(lambda_eq_nm, float)
    N = len(lambda_eq_nm)
    k_eq = 2 * np.pi / lambda_eq_nm
    hann = 0.5 * (1 - np.cos(2 * np.pi * np.arange(N) / (N - 1)))     # symmetric Hann
    hann = hann / np.sqrt(np.mean(hann ** 2))
    beta = np.asarray(dispersion_quadratic_term, float)
    phase = -beta[..., None] * (k_eq - k_eq.mean()) ** 2 if beta.ndim else \
        -dispersion_quadratic_term * (k_eq - k_eq.mean()) ** 2
    return hann * np.exp(1j * phase)


def beta_from_index_units(beta_idx, lambda_nm):
    """Dispersion given as phase per squared equispaced-k sample index,
    phi = beta_idx * ((k - k0)/dk)^2, converted to the reference's nm^2/rad term
    beta = beta_idx / dk^2 with dk the equispaced-k step (float in, float out; arrays elementwise)."""
    k = 2 * np.pi / np.asarray(lambda_nm, float)
    dk = (k.max() - k.min()) / (len(k) - 1)
    if np.ndim(beta_idx) == 0:
        return float(beta_idx) / dk ** 2
    return np.asarray(beta_idx, float) / dk ** 2


# --------------------------------------------------------------------------------------
# Spatial geometry
# --------------------------------------------------------------------------------------
@dataclass
class ScanGrid:
    """The scan-description fields tiled stitching needs (the reference reads them from
    ScanInfo.json, octrecon/io/scaninfo.py:26-65)."""
    x_centers_mm: np.ndarray
    y_centers_mm: np.ndarray
    z_depths_mm: np.ndarray
    tile_range_x_mm: float
    tile_range_y_mm: float
    n_x_px: int
    n_y_px: int
    x_offset: float = 0.0
    y_offset: float = 0.0
    pixel_size_um: float = float("nan")
    galvo_phase_delay_asamples: float = 0.0    # used only when no explicit x correction
    galvo_x_correction_given: bool = True

    def __post_init__(self):
        self.x_centers_mm = np.asarray(self.x_centers_mm, float).ravel()
        self.y_centers_mm = np.asarray(self.y_centers_mm, float).ravel()
        self.z_depths_mm = np.asarray(self.z_depths_mm, float).ravel()

    def tiles_in_row(self, yi: int):
        """(xi, zi) pairs of one y row, x-major / z-minor (legacy processing order)."""
        return [(xi, zi) for xi in range(len(self.x_centers_mm)) for zi in range(len(self.z_depths_mm))]


@dataclass
class TiledGeometry:
    tile_x_mm: np.ndarray
    tile_y_mm: np.ndarray
    tile_z_mm: np.ndarray
    out_x_mm: np.ndarray
    out_y_mm: np.ndarray
    out_z_mm: np.ndarray
    out_z_full_mm: np.ndarray
    focus_pix: np.ndarray


def _tile_axis(offset: float, span: float, n: int) -> np.ndarray:
    """Tile-local positions of n samples over `span`, centred on `offset`: the first n nodes of
    MATLAB linspace(-1/2, 1/2, n + 1) scaled (a sample sits at the start of its pixel)."""
    return offset + span * matlab_linspace(-0.5, 0.5, n + 1)[:-1]


def _mosaic_axis(centres, tile: np.ndarray) -> np.ndarray:
    """Output positions of one lateral axis: MATLAB colon over every tile's samples on the tile
    pitch (min centre + first sample ... max centre + last sample, half a pitch of slack); one
    tile keeps its own axis."""
    if len(centres) == 1:
        return tile.copy()
    step = tile[1] - tile[0] if len(tile) > 1 else 0.0
    return matlab_colon(centres.min() + tile[0], step, centres.max() + tile[-1] + step / 2)


def _depth_axes(z_um_tile, focus: np.ndarray, z_depths: np.ndarray):
    """(tile-local z in mm, 0 at the reference focus row; output z before resampling).  The
    reference focus row is the common one, or the one of the depth nearest z = 0; without
    gating (any NaN focus) the tile axis stays at the spectrometer's.  The output axis covers
    every depth's tile on the tile pitch and is re-anchored so that its node nearest 0 is 0."""
    tz = np.asarray(z_um_tile, float) / 1e3
    if not np.isnan(focus).any():
        row = focus[0] if np.all(focus == focus[0]) else focus[int(np.argmin(np.abs(z_depths)))]
        tz = tz - tz[int(np.round(row)) - 1]
    pitch = tz[1] - tz[0]
    span = matlab_colon(z_depths.min() + tz[0], pitch, z_depths.max() + tz[-1] + pitch / 2)
    anchor = int(np.argmin(np.abs(span)))
    return tz, pitch * (np.arange(1, span.size + 1) - (anchor + 1))


def build_tiled_geometry(sg: ScanGrid, z_um_tile, focus_pix, output_pixel_size_um,
                         crop_z_range_mm=None) -> TiledGeometry:
    """Tile and output grids of a tiled scan (the model of octrecon/core/geometry.py:146-202):
    lateral tile axes from the scan's ranges (with the galvo phase-delay shift of x when the scan
    states no explicit x correction), the depth axes of _depth_axes, the output axes of
    _mosaic_axis; an output pixel size must equal the scan's lateral pitch (rounded to 0.01 um)
    and resamples z to it; crop_z_range_mm = [lo, hi] keeps the output depths inside it.  The
    grids' float64 expressions are the reference's, so the grid nodes (and every interpolation
    weight built on them) come out identical."""
    nd = len(sg.z_depths_mm)
    focus = np.asarray(focus_pix, float).ravel()
    if focus.size == 1:
        focus = np.full(nd, focus[0])
    if focus.size != nd:
        raise ValueError("focus_pix must be scalar or one value per zDepth")
    tx = _tile_axis(sg.x_offset, sg.tile_range_x_mm, sg.n_x_px)
    ty = _tile_axis(sg.y_offset, sg.tile_range_y_mm, sg.n_y_px)
    delay = sg.galvo_phase_delay_asamples
    if not sg.galvo_x_correction_given and delay != 0 and np.isfinite(sg.pixel_size_um):
        tx = tx - delay * (sg.pixel_size_um - 1.0) * 1e-3
    tz, z_full = _depth_axes(z_um_tile, focus, sg.z_depths_mm)
    out_x = _mosaic_axis(sg.x_centers_mm, tx)
    out_y = _mosaic_axis(sg.y_centers_mm, ty)
    out_z = z_full
    if output_pixel_size_um is not None:
        px = np.round(np.mean(np.diff(out_x)) * 1e3 * 100) / 100
        py = np.round(np.mean(np.diff(out_y)) * 1e3 * 100) / 100 if len(out_y) >= 2 else px
        if not (px == py == output_pixel_size_um):
            raise ValueError(f"output pixel size {output_pixel_size_um} um must equal the scan pixel size "
                             f"({px}, {py}) um")
        out_z = matlab_colon(z_full[0], output_pixel_size_um * 1e-3, z_full.max())
    if crop_z_range_mm is not None and len(crop_z_range_mm):
        lo, hi = crop_z_range_mm
        out_z = out_z[(out_z >= lo) & (out_z <= hi)]
        if out_z.size == 0:
            raise ValueError("crop z range does not overlap the available z range")
    return TiledGeometry(tx, ty, tz, out_x, out_y, out_z, z_full.copy(), focus)
