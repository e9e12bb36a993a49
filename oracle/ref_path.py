"""ORACLE - test infrastructure only (see oracle/__init__.py).

Restatement of the reference's CPU reconstruction path, written from its algorithm:

  geometry      octrecon/core/geometry.py:84-128   (dense N x N sinc mask, as the reference)
  pre-FFT       octrecon/core/cpu_kernels.py:12-41  (C restatement, oracle/fused_pre_fft.c)
  FFT + crop    octrecon/core/spectral.py:136-165   (scipy.fft.ifft, pocketfft, keep N/2, abs)
  frame_mean    octrecon/core/spectral.py:49-71
  stitching     octrecon/core/stitching.py:20-140   (dense interpolation matrices, matmul)
  accumulate    octrecon/pipeline.py:231-232, 282-284
  finalise      octrecon/pipeline.py:296-300
  clim          octrecon/io/tiff_writer.py:48-50, octrecon/pipeline.py:376-378, 388
  quantise      octrecon/io/tiff_writer.py:26-45, 62-65

Third-party arithmetic used as in the reference: scipy.fft (pocketfft; scipy 1.15.3 in this
image) for the single-precision inverse FFT, numpy matmul (OpenBLAS) for the stitching
products, scipy.interpolate.PchipInterpolator for lambda_eq.

Pins (tests/test_oracle.py): the MATLAB plane of the reference's tests for the Telesto synthetic
scan (tests/golden/legacy_telesto.npz, 1e-4 dB), a brute-force DFT, closed forms of the window and
quantisation, MATLAB filter2 binning values, the reference's own dispersion-window identity, and
the dense float64 operator against the float64 spectral path.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np
from scipy.interpolate import PchipInterpolator

_HERE = Path(__file__).resolve().parent
_LIB = None

MAXBIT = 2 ** 16 - 1
MIN_WEIGHT = float(np.exp(-3.0 ** 2 / 2))


# ----------------------------------------------------------------------------- C library
def lib():
    """Load oracle/liboct_oracle.so (built by oracle/Makefile)."""
    global _LIB
    if _LIB is None:
        path = _HERE / "liboct_oracle.so"
        if not path.exists():
            raise RuntimeError(f"{path} is not built; run `make -C oracle`")
        L = ctypes.CDLL(str(path))
        for name in ("oracle_fused_pre_fft_i16", "oracle_fused_pre_fft_u16", "oracle_fused_pre_fft_f32"):
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_num_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


# ----------------------------------------------------------------------------- geometry
def linspace_m(a, b, n):
    """MATLAB linspace (octrecon/core/geometry.py:30-35)."""
    y = a + (np.arange(n) * (b - a)) / (n - 1)
    y[0] = a
    y[-1] = b
    return y


def colon_m(a, d, b):
    """MATLAB colon (octrecon/core/geometry.py:38-51)."""
    tol = 2.0 * np.finfo(float).eps * max(abs(a), abs(b))
    n = int(np.floor((b - a) / d + tol / abs(d))) if d != 0 else -1
    if n < 0:
        return np.zeros(0)
    k = np.arange(n + 1, dtype=float)
    out = np.empty(n + 1)
    half = n // 2
    end = a + n * d
    out[: half + 1] = a + k[: half + 1] * d
    out[half + 1:] = end - (n - k[half + 1:]) * d
    return out


@dataclass
class OracleGeometry:
    lam: np.ndarray
    idx: np.ndarray       # (N, T) int64
    w: np.ndarray         # (N, T) float64
    window: np.ndarray    # (N,) complex128
    z_um: np.ndarray
    lambda_eq: np.ndarray


def spectral_geometry(lambda_nm, beta, n_medium, taps=5) -> OracleGeometry:
    """Dense restatement of build_spectral_geometry (octrecon/core/geometry.py:84-128):
    the full (N, N) distance matrix, mask |dist| < taps, np.sinc weights."""
    lam = np.asarray(lambda_nm, float).ravel()
    N = len(lam)
    k = 2 * np.pi / lam
    k_lin = linspace_m(k.max(), k.min(), N)
    nq = (k_lin - k_lin.min()) / (k_lin.max() - k_lin.min()) * (N - 1)
    n = (k - k.min()) / (k.max() - k.min()) * (N - 1)
    dist = nq[:, None] - n[None, :]
    mask = np.abs(dist) < taps
    T = int(mask.sum(1).max())
    idx = np.zeros((N, T), np.int64)
    w = np.zeros((N, T))
    for i in range(N):
        jj = np.nonzero(mask[i])[0]
        idx[i, : len(jj)] = jj
        w[i, : len(jj)] = np.sinc(dist[i, jj])
    kk, ll = (k[::-1], lam[::-1]) if k[0] > k[-1] else (k, lam)
    lambda_eq = PchipInterpolator(kk, ll, extrapolate=True)(k_lin)
    k_eq = 2 * np.pi / lambda_eq
    hann = 0.5 * (1 - np.cos(2 * np.pi * np.arange(N) / (N - 1)))
    hann = hann / np.sqrt(np.mean(hann ** 2))
    window = hann * np.exp(1j * (-beta * (k_eq - k_eq.mean()) ** 2))
    ls = np.sort([lambda_eq[0], lambda_eq[-1]])
    z_step = 0.5 * (ls.mean() / 1e3) ** 2 / (np.diff(ls)[0] / 1e3) / n_medium
    z_um = linspace_m(0.0, z_step * N / 2, N // 2)
    return OracleGeometry(lam, idx, w, window, z_um, lambda_eq)


# ----------------------------------------------------------------------------- spectral
def frame_mean_prepare(raw, group=1, dtype=np.float32):
    """apod_mode 'frame_mean' (octrecon/core/spectral.py:54-66): background = mean over all
    A-lines of each group of `group` files, prepended as one A-line."""
    x = raw.astype(dtype)
    B, nL, N = x.shape
    g = max(1, int(group))
    apod = x.reshape(B // g, g, nL, N).mean(axis=2).mean(axis=1)
    apod = np.repeat(apod, g, axis=0)[:, None, :]
    return np.concatenate([apod, x], axis=1)


def binned_columns(n_lines, binning):
    """0-based scan lines kept by AScanBinning (octrecon/core/spectral.py:20-24): every B-th line
    from max(1, B // 2) - 1."""
    return np.arange(max(1, binning // 2) - 1, n_lines, binning)


def bin_ascans(x, binning, dtype=np.float32):
    """AScanBinning along the A-line axis (octrecon/core/spectral.py:27-46): filter2 with
    ones(1, B) / B, 'same', zero padded, then decimation; each product x * (1/B) is rounded in
    `dtype` and the products are summed in order.  x (..., nLines, N) -> (..., nKept, N)."""
    x = np.asarray(x, dtype)
    n = x.shape[-2]
    cols = binned_columns(n, binning)
    left = (binning + 1) // 2 - 1
    right = binning // 2
    pad = [(0, 0)] * x.ndim
    pad[-2] = (left, right)
    xp_ = np.pad(x, pad)
    h = np.dtype(dtype).type(1.0 / binning)
    acc = None
    for t in range(binning):
        sl = xp_[..., cols[0] + t: cols[0] + t + binning * (len(cols) - 1) + 1: binning, :]
        term = (sl * h).astype(dtype)
        acc = term if acc is None else (acc + term).astype(dtype)
    return acc


def prepare_raw(raw, apod_size, binning=1, apod_mode="lines", apod_group=1, dtype=np.float32):
    """Raw files -> the array the pre-FFT kernel sees (octrecon/core/spectral.py:49-71): the
    background lines (or, for 'frame_mean', the frame-mean line computed from the unbinned
    A-lines) followed by the binned scan lines.  Returns (array, background lines)."""
    if apod_mode == "frame_mean":
        pre = frame_mean_prepare(raw, apod_group, dtype)
        if binning > 1:
            pre = np.concatenate([pre[:, :1], bin_ascans(pre[:, 1:], binning, dtype)], axis=1)
        return pre, 1
    if binning > 1:
        x = np.asarray(raw, dtype)
        return np.concatenate([x[:, :apod_size], bin_ascans(x[:, apod_size:], binning, dtype)], axis=1), apod_size
    return raw, apod_size


def average_repeats(mag, n_frames, bscan_avg, size_x, ascan_avg):
    """octrecon/core/spectral.py:168-177: (n_frames nb, size_x na, nZ) -> (n_frames, size_x, nZ),
    the mean over B-scan repeats, then over A-scan repeats (numpy, the array's dtype)."""
    nb, na = max(1, int(bscan_avg)), max(1, int(ascan_avg))
    if nb == 1 and na == 1:
        return mag
    m5 = mag.reshape(n_frames, nb, size_x, na, mag.shape[-1])
    m4 = m5.mean(1) if nb > 1 else m5[:, 0]
    return m4.mean(2) if na > 1 else m4[:, :, 0]


def fused_pre_fft(raw, apod_size, geo: OracleGeometry):
    """Reference CPU pre-FFT (Numba typing) via the C restatement -> complex64 (B, nX, N)."""
    raw = np.ascontiguousarray(raw)
    B, interf, N = raw.shape
    nX = interf - apod_size
    fn = {np.dtype(np.int16): lib().oracle_fused_pre_fft_i16,
          np.dtype(np.uint16): lib().oracle_fused_pre_fft_u16,
          np.dtype(np.float32): lib().oracle_fused_pre_fft_f32}[raw.dtype]
    idxT = np.ascontiguousarray(geo.idx.T)
    wT = np.ascontiguousarray(geo.w.T.astype(np.float32))
    win = np.ascontiguousarray(np.stack([geo.window.real, geo.window.imag], -1).astype(np.float32))
    out = np.empty((B, nX, N), np.complex64)
    rc = fn(raw.ctypes.data, B, interf, N, apod_size, idxT.ctypes.data, wT.ctypes.data,
            idxT.shape[0], win.ctypes.data, out.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"oracle fused_pre_fft failed ({rc})")
    return out


def magnitude_cpu(raw, apod_size, geo: OracleGeometry, apod_mode="lines", apod_group=1, workers=None, binning=1):
    """The reference's float32 CPU path (SpectralProcessor.magnitude with Numba,
    octrecon/core/spectral.py:112-116, 136-153): prepare (binning / frame mean), pre-FFT, scipy
    ifft, keep N/2, abs -> float32."""
    import scipy.fft
    raw, apod_size = prepare_raw(raw, apod_size, binning, apod_mode, apod_group, np.float32)
    spec = fused_pre_fft(raw, apod_size, geo)
    ft = scipy.fft.ifft(spec, axis=-1, workers=workers or os.cpu_count(), overwrite_x=True)
    return np.abs(ft[..., : spec.shape[-1] // 2]).astype(np.float32)


def complex_fp64(raw, apod_size, geo: OracleGeometry, apod_mode="lines", apod_group=1):
    """Float64 oracle of the spectral path up to the cropped complex A-scan
    (octrecon/core/spectral.py:123-134, 151-158 with dtype float64)."""
    if apod_mode == "frame_mean":
        raw = frame_mean_prepare(raw, apod_group, np.float64)
        apod_size = 1
    x = raw.astype(np.float64)
    apod = x[:, :apod_size, :].mean(axis=1, keepdims=True)
    interf = x[:, apod_size:, :] - apod
    mv = interf.mean(axis=-1, keepdims=True)
    d = interf - mv
    out = np.zeros_like(d)
    for t in range(geo.idx.shape[1]):
        out += geo.w[:, t] * np.take(d, geo.idx[:, t], axis=-1)
    out += mv
    ft = np.fft.ifft(out * geo.window, axis=-1)
    return ft[..., : d.shape[-1] // 2]


def magnitude_fp64(raw, apod_size, geo: OracleGeometry, apod_mode="lines", apod_group=1):
    return np.abs(complex_fp64(raw, apod_size, geo, apod_mode, apod_group))


def dft_brute(x):
    """Textbook inverse DFT with 1/N (float64) - used to pin the FFT at small N."""
    x = np.asarray(x, complex)
    N = x.shape[-1]
    n = np.arange(N)
    M = np.exp(2j * np.pi * np.outer(n, n) / N) / N
    return x @ M.T


# ----------------------------------------------------------------------------- stitching
def factor_z(n_rows, focus_pix, sigma):
    """octrecon/core/stitching.py:31-33."""
    zI = np.arange(1, n_rows + 1, dtype=float)
    return np.exp(-(zI - focus_pix) ** 2 / (2 * sigma) ** 2) + np.exp(-3.0 ** 2 / 2)


def min_weight_threshold(dtype):
    """octrecon/core/stitching.py:24-28."""
    return MIN_WEIGHT * (1.0 - 8 * float(np.finfo(dtype).eps))


def interp_matrix(src, dst):
    """Dense linear interpolation matrix (octrecon/core/stitching.py:36-51)."""
    W = np.zeros((len(dst), len(src)))
    inside = (dst >= src[0]) & (dst <= src[-1])
    q = dst[inside]
    j = np.clip(np.searchsorted(src, q, side="right") - 1, 0, len(src) - 2)
    t = (q - src[j]) / (src[j + 1] - src[j])
    t[np.abs(t) < 1e-9] = 0.0
    t[np.abs(1.0 - t) < 1e-9] = 1.0
    rows = np.nonzero(inside)[0]
    W[rows, j] += 1 - t
    W[rows, j + 1] += t
    return W


def opc_build(poly, tile_x_mm, tile_y_mm, tile_z_mm):
    """OpticalPathCorrection.build (octrecon/core/stitching.py:67-78) -> (shift, pdz, nZ)."""
    x, y, z = tile_x_mm * 1e3, tile_y_mm * 1e3, tile_z_mm * 1e3
    dz = (z[-1] - z[0]) / (len(z) - 1)
    X, Y = np.meshgrid(x, y)
    P = poly[0] * X + poly[1] * Y + poly[2] * X ** 2 + poly[3] * Y ** 2 + poly[4] * X * Y
    pdz = P / dz
    return np.floor(pdz + 0.5).astype(np.int64), pdz, len(z)


class Stitcher:
    """TileStitcher (octrecon/core/stitching.py:84-140), dense matrices, numpy matmul."""

    def __init__(self, tile_x_mm, tile_z_mm, out_x_mm, out_z_mm, opc, xc, zd, focus_pix, sigma,
                 dtype=np.float32):
        nz = len(tile_z_mm)
        x = tile_x_mm + xc
        z = tile_z_mm + zd
        if not np.isnan(focus_pix):
            z = z - tile_z_mm[int(np.round(focus_pix)) - 1]
        x = x.copy(); z = z.copy()
        x[0] -= 1e-10; x[-1] += 1e-10
        z[0] -= 1e-10; z[-1] += 1e-10
        Wz = interp_matrix(z, out_z_mm)
        Wx = interp_matrix(x, out_x_mm)
        rows = np.nonzero(Wz.any(axis=0))[0]
        cols = np.nonzero(Wx.any(axis=1))[0]
        self.empty = rows.size == 0 or cols.size == 0
        if self.empty:
            return
        self.r0, self.r1, self.c0, self.c1 = rows[0], rows[-1] + 1, cols[0], cols[-1] + 1
        self.Wz = Wz[:, self.r0:self.r1].astype(dtype)
        self.WxT = Wx[self.c0:self.c1, :].T.astype(dtype)
        fz = factor_z(nz, focus_pix, sigma) if not np.isnan(focus_pix) else np.ones(nz)
        self.fz = fz[self.r0:self.r1].astype(dtype)
        self.rr = np.arange(self.r0, self.r1)
        self.nz = nz
        self.opc = opc

    def contribute(self, mag, frame_idx):
        if self.opc is not None:
            shift, pdz, _ = self.opc
            src = self.rr + shift[frame_idx][..., None]
            t = self.rr + pdz[frame_idx][..., None]
            valid = (t >= 0) & (t <= self.nz - 1)
            src = np.clip(src, 0, self.nz - 1)
        else:
            src = np.broadcast_to(self.rr, (len(frame_idx), mag.shape[1], len(self.rr)))
            valid = np.ones(src.shape, bool)
        a = np.take_along_axis(mag, src, axis=-1)
        f = valid * self.fz
        a = a * f
        num = np.matmul(np.matmul(a, self.Wz.T).transpose(0, 2, 1), self.WxT)
        den = np.matmul(np.matmul(f, self.Wz.T).transpose(0, 2, 1), self.WxT)
        return num, den


def finalise(num, den, dtype=np.float32):
    """octrecon/pipeline.py:296-300: weight below threshold -> NaN; 20 log10(num/den)."""
    den = den.copy()
    den[den < min_weight_threshold(dtype)] = np.nan
    with np.errstate(divide="ignore", invalid="ignore"):
        return (20 * np.log10(num / den)).astype(np.float32)


# ----------------------------------------------------------------------------- quantisation
def matlab_uint16(v):
    """MATLAB uint16(): round half away from zero, saturate, NaN -> 0 (tiff_writer.py:26-30)."""
    out = np.floor(np.abs(v) + 0.5) * np.sign(v)
    out = np.nan_to_num(out, nan=0.0)
    return np.clip(out, 0, MAXBIT).astype(np.uint16)


def db_to_bits(db, clim):
    """octrecon/io/tiff_writer.py:33-38."""
    c1, c2 = sorted(clim)
    with np.errstate(invalid="ignore"):
        bits = matlab_uint16((np.asarray(db, np.float64) - c1) / (c2 - c1) * (MAXBIT - 1)).astype(np.uint32) + 1
    bits = np.minimum(bits, MAXBIT).astype(np.uint16)
    bits[np.isnan(db)] = 0
    return bits


def bits_to_db(bits, clim):
    """octrecon/io/tiff_writer.py:41-45."""
    c1, c2 = sorted(clim)
    out = (bits.astype(np.float64) - 1) * (c2 - c1) / (MAXBIT - 1) + c1
    out[bits == 0] = np.nan
    return out


def plane_clim(db):
    """octrecon/io/tiff_writer.py:48-50."""
    fin = db[np.isfinite(db)]
    return (float(fin.min()), float(fin.max())) if fin.size else (np.nan, np.nan)


def volume_clim(db_planes):
    """Global clim = nanmin/nanmax of the per-plane clims (octrecon/pipeline.py:376-378, 388)."""
    pcs = np.array([plane_clim(p) for p in db_planes])
    return float(np.nanmin(pcs[:, 0])), float(np.nanmax(pcs[:, 1])), pcs


def quantise_volume(db_planes, clim, legacy_double=False, plane_clims=None):
    """Per-plane quantisation as write_legacy_tiff does (octrecon/io/tiff_writer.py:55-62)."""
    out = np.empty(np.shape(db_planes), np.uint16)
    for yi in range(len(db_planes)):
        plane = np.asarray(db_planes[yi], dtype=np.float64)
        if legacy_double:
            pc = plane_clims[yi] if plane_clims is not None else plane_clim(plane)
            plane = bits_to_db(db_to_bits(plane, pc), pc)
        out[yi] = db_to_bits(plane, clim)
    return out


# ----------------------------------------------------------------------------- tiled path
@dataclass
class OracleTiled:
    tx: np.ndarray
    ty: np.ndarray
    tz: np.ndarray
    out_x: np.ndarray
    out_y: np.ndarray
    out_z: np.ndarray


def tiled_geometry(xc, yc, zd, tile_range_x_mm, tile_range_y_mm, n_x_px, n_y_px, z_um_tile, focus,
                   pixel_um, crop=None, x_offset=0.0, y_offset=0.0) -> OracleTiled:
    """Restatement of build_tiled_geometry (octrecon/core/geometry.py:146-202), without the
    galvo-delay correction (the scans used here state it as 0)."""
    xc, yc, zd = (np.asarray(v, float).ravel() for v in (xc, yc, zd))
    focus = np.asarray(focus, float).ravel()
    if focus.size == 1:
        focus = np.full(len(zd), focus[0])
    tx = (x_offset + tile_range_x_mm * linspace_m(-0.5, 0.5, n_x_px + 1))[:-1]
    ty = (y_offset + tile_range_y_mm * linspace_m(-0.5, 0.5, n_y_px + 1))[:-1]
    tz = np.asarray(z_um_tile, float) / 1e3
    if not np.any(np.isnan(focus)):
        ref = focus[0] if np.all(focus == focus[0]) else focus[int(np.argmin(np.abs(zd)))]
        tz = tz - tz[int(np.round(ref)) - 1]
    dx = tx[1] - tx[0]
    dy = ty[1] - ty[0] if len(ty) > 1 else 0.0
    dz = tz[1] - tz[0]
    x_all = colon_m(xc.min() + tx[0], dx, xc.max() + tx[-1] + dx / 2)
    y_all = colon_m(yc.min() + ty[0], dy, yc.max() + ty[-1] + dy / 2)
    z_all = colon_m(zd.min() + tz[0], dz, zd.max() + tz[-1] + dz / 2)
    zero = int(np.argmin(np.abs(z_all)))
    z_all = dz * (np.arange(1, len(z_all) + 1) - (zero + 1))
    if len(xc) == 1:
        x_all = tx.copy()
    if len(yc) == 1:
        y_all = ty.copy()
    if pixel_um is not None:
        z_all = colon_m(z_all[0], pixel_um * 1e-3, z_all.max())
    if crop is not None:
        z_all = z_all[~((z_all < crop[0]) | (z_all > crop[1]))]
    return OracleTiled(tx, ty, tz, x_all, y_all, z_all)


def reconstruct_row(tiles_raw, tile_ids, frames, geo: OracleGeometry, tg: OracleTiled, xc, zd, focus,
                    sigma, apod_size, opc_poly=None, precision="float32", apod_mode="lines"):
    """Restatement of Reconstructor.process_row (octrecon/pipeline.py:223-302) for in-memory
    tiles.  tiles_raw[k]: (nFrames, interf, N) raw frames of tile tile_ids[k] = (xi, zi);
    frames: tile-local y frame indices of those raw frames.  Returns float32 dB
    (nFrames, nOutZ, nOutX).  precision 'float64' runs the float64 oracle, 'float32' the
    reference's float32 CPU path."""
    dtype = np.float64 if precision == "float64" else np.float32
    focus = np.asarray(focus, float).ravel()
    if focus.size == 1:
        focus = np.full(len(zd), focus[0])
    opc = opc_build(np.asarray(opc_poly, float), tg.tx, tg.ty, tg.tz) if opc_poly is not None else None
    nF = len(frames)
    num = np.zeros((nF, len(tg.out_z), len(tg.out_x)), dtype)
    den = np.zeros_like(num)
    fr = np.asarray(frames)
    for raw, (xi, zi) in zip(tiles_raw, tile_ids):
        st = Stitcher(tg.tx, tg.tz, tg.out_x, tg.out_z, opc, xc[xi], zd[zi], focus[zi], sigma, dtype)
        if st.empty:
            continue
        if precision == "float64":
            mag = magnitude_fp64(raw, apod_size, geo, apod_mode)
        else:
            mag = magnitude_cpu(raw, apod_size, geo, apod_mode)
        n_, d_ = st.contribute(mag.astype(dtype), fr)
        num[:, :, st.c0:st.c1] += n_
        den[:, :, st.c0:st.c1] += d_
    den[den < min_weight_threshold(dtype)] = np.nan
    with np.errstate(divide="ignore", invalid="ignore"):
        return (20 * np.log10(num / den)).astype(np.float32)


def linearise_fp64(raw, apod_size, geo: OracleGeometry, apod_mode="lines", apod_group=1):
    """Equispaced interferogram (octrecon/core/spectral.py:118-134) in float64."""
    if apod_mode == "frame_mean":
        raw = frame_mean_prepare(raw, apod_group, np.float64)
        apod_size = 1
    x = raw.astype(np.float64)
    interf = x[:, apod_size:, :] - x[:, :apod_size, :].mean(axis=1, keepdims=True)
    mv = interf.mean(axis=-1, keepdims=True)
    d = interf - mv
    out = np.zeros_like(d)
    for t in range(geo.idx.shape[1]):
        out += geo.w[:, t] * np.take(d, geo.idx[:, t], axis=-1)
    return out + mv


def lsb_agreement(mag_test, mag_ref, clim=None):
    """Quantise 20 log10 of two magnitude volumes with one clim (the reference's, by default:
    its finite dB min/max) and return (fraction of voxels within 1 code, max code difference,
    clim)."""
    with np.errstate(divide="ignore"):
        d_ref = 20 * np.log10(np.asarray(mag_ref, np.float32)).astype(np.float32)
        d_t = 20 * np.log10(np.asarray(mag_test, np.float32)).astype(np.float32)
    if clim is None:
        fin = d_ref[np.isfinite(d_ref)]
        clim = (float(fin.min()), float(fin.max()))
    a = db_to_bits(d_t, clim).astype(np.int64)
    b = db_to_bits(d_ref, clim).astype(np.int64)
    diff = np.abs(a - b)
    return float(np.mean(diff <= 1)), int(diff.max()), clim


def operator_fp64(geo: OracleGeometry, z0: int, nzc: int):
    """Dense float64 operator of the linear spectral map for bins [z0, z0 + nzc):
    X = A d with d = raw - background, A = F_crop diag(window / N) (S + (1/N)(1 - S1) 1^T),
    assembled from the dense restatement's sinc rows (octrecon/core/spectral.py:123-134, 151-153)."""
    N = len(geo.lam)
    S = np.zeros((N, N))
    for t in range(geo.idx.shape[1]):
        np.add.at(S, (np.arange(N), geo.idx[:, t]), geo.w[:, t])
    u = 1.0 - S.sum(1)
    Sp = S + np.outer(u, np.ones(N)) / N
    z = np.arange(z0, z0 + nzc)
    F = np.exp(2j * np.pi * ((np.outer(z, np.arange(N))) % N) / N) * (geo.window / N)[None, :]
    return F @ Sp


def top_percentile_metric(A, z_min=20, frac=0.01):
    """Sharpness metric 'top_percentile' (octrecon/estimation/dispersion.py:46-63): per B-scan,
    log of the mean of the k = max(1, int(M * frac)) largest |A|^2 (A float32, squares cast to
    float64) over all x and z >= z_min."""
    A = np.asarray(A, np.float32)
    B = A.shape[0]
    I = (A[:, :, z_min:] ** 2).reshape(B, -1).astype(np.float64)
    M = I.shape[1]
    k = max(1, int(M * frac))
    top = np.partition(I, M - k, axis=1)[:, M - k:]
    return np.log(top.mean(axis=1))
