"""Optical-path correction + focus-weighted tile stitching on the hip device.

Host-side (float64) operator builders with the reference's semantics, and the device
``HipTileStitcher`` that applies them through liboct_hip.

Reference: octrecon/core/stitching.py
  MIN_WEIGHT / min_weight_threshold   :20-28
  factor_z                            :31-33
  linear_interp_matrix                :36-51
  OpticalPathCorrection               :54-78
  TileStitcher                        :81-140
and the integer validity bounds of octrecon/core/torch_backend.py:141-147.

Each interpolation row has at most two non-zero weights, so the device receives them as
(row, weight) pairs instead of dense matrices; the kernel forms the same two-term sums the
dense products reduce to.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib

CUTOFF_SIGMA = 3.0
MIN_WEIGHT = float(np.exp(-CUTOFF_SIGMA ** 2 / 2))


def min_weight_threshold(dtype) -> float:
    """Weight below which a voxel is NaN: exp(-4.5) less a few ulps of the accumulation type,
    so that a voxel covered by one tile at its floor weight stays finite."""
    return MIN_WEIGHT * (1.0 - 8 * float(np.finfo(dtype).eps))


def factor_z(n_rows: int, focus_pix_1based: float, focus_sigma: float) -> np.ndarray:
    """Focus weight exp(-(z - f)^2 / (2 sigma)^2) + exp(-4.5) over 1-based rows."""
    z = np.arange(1, n_rows + 1, dtype=float)
    return np.exp(-(z - focus_pix_1based) ** 2 / (2 * focus_sigma) ** 2) + np.exp(-CUTOFF_SIGMA ** 2 / 2)


def interp_taps(src: np.ndarray, dst: np.ndarray):
    """Two-tap form of MATLAB griddedInterpolant 'linear' with 0 outside, for strictly
    increasing src: dst[i] = w0[i] v[j[i]] + w1[i] v[j[i] + 1].  Queries within 1e-9 (in
    units of the cell) of a node take the node value exactly.  Outside: weights 0."""
    src = np.asarray(src, float)
    dst = np.asarray(dst, float)
    inside = (dst >= src[0]) & (dst <= src[-1])
    j = np.zeros(len(dst), np.int64)
    w0 = np.zeros(len(dst))
    w1 = np.zeros(len(dst))
    q = dst[inside]
    jj = np.clip(np.searchsorted(src, q, side="right") - 1, 0, len(src) - 2)
    t = (q - src[jj]) / (src[jj + 1] - src[jj])
    t[np.abs(t) < 1e-9] = 0.0
    t[np.abs(1.0 - t) < 1e-9] = 1.0
    j[inside] = jj
    w0[inside] = 1 - t
    w1[inside] = t
    return j, w0, w1, inside


def linear_interp_matrix(src: np.ndarray, dst: np.ndarray) -> np.ndarray:
    """Dense W (len(dst), len(src)) with W @ v == interp1(src, v, dst, 'linear', 0)."""
    j, w0, w1, inside = interp_taps(src, dst)
    W = np.zeros((len(dst), len(src)))
    rows = np.nonzero(inside)[0]
    W[rows, j[rows]] += w0[rows]
    W[rows, j[rows] + 1] += w1[rows]
    return W


@dataclass
class OpticalPathCorrection:
    """Per (frame, column) integer row shift round(P/dz) and the validity of r + P/dz in
    [0, nZ-1], P(x, y) = p1 x + p2 y + p3 x^2 + p4 y^2 + p5 x y in um."""
    shift: np.ndarray    # (nY, nX) int64
    pdz: np.ndarray      # (nY, nX) float64
    n_rows: int

    @classmethod
    def build(cls, poly, tile_x_mm, tile_y_mm, tile_z_mm):
        """Tables for the polynomial poly = (p1..p5) over the tile's (frame, A-scan) grid: the
        path in depth pixels P / dz (dz the mean pitch of the tile's depth axis) and its nearest
        row offset (halves round up).  The polynomial is evaluated term by term as written above,
        in float64, so the offsets and the validity edges match the reference's bit for bit."""
        c = np.asarray(poly, float).ravel()
        if c.size != 5:
            raise ValueError("the optical-path polynomial has 5 coefficients")
        xu = np.asarray(tile_x_mm, float)[None, :] * 1e3          # (1, nX) um
        yu = np.asarray(tile_y_mm, float)[:, None] * 1e3          # (nY, 1) um
        zu = np.asarray(tile_z_mm, float) * 1e3
        pitch = (zu[-1] - zu[0]) / (zu.size - 1)
        path = c[0] * xu + c[1] * yu + c[2] * xu ** 2 + c[3] * yu ** 2 + c[4] * xu * yu
        rows = path / pitch
        return cls(np.floor(rows + 0.5).astype(np.int64), rows, zu.size)

    def valid_bounds(self):
        """Integer (lo, hi): row r is valid iff lo <= r <= hi, decided with the same float64
        expressions r + pdz >= 0 and r + pdz <= nZ - 1 (monotonic in r)."""
        n1 = self.n_rows - 1
        lo = np.ceil(-self.pdz).astype(np.int64)
        hi = np.floor(n1 - self.pdz).astype(np.int64)
        # settle the rounding edge of the float expressions exactly
        lo = np.where((lo - 1) + self.pdz >= 0, lo - 1, lo)
        lo = np.where(lo + self.pdz >= 0, lo, lo + 1)
        hi = np.where((hi + 1) + self.pdz <= n1, hi + 1, hi)
        hi = np.where(hi + self.pdz <= n1, hi, hi - 1)
        return lo, hi


class TileStitcher:
    """Host-built operators for one tile position (xi, zi) of the grid (float64)."""

    def __init__(self, tg, oc: OpticalPathCorrection | None, xi_center_mm, z_depth_mm, focus_pix,
                 focus_sigma):
        nz = len(tg.tile_z_mm)
        x = tg.tile_x_mm + xi_center_mm
        z = tg.tile_z_mm + z_depth_mm
        if not np.isnan(focus_pix):
            z = z - tg.tile_z_mm[int(np.round(focus_pix)) - 1]
        x = x.copy()
        z = z.copy()
        x[0] -= 1e-10
        x[-1] += 1e-10
        z[0] -= 1e-10
        z[-1] += 1e-10
        zj, zw0, zw1, zin = interp_taps(z, tg.out_z_mm)
        xj, xw0, xw1, xin = interp_taps(x, tg.out_x_mm)
        # a row/column of the dense operator is "used" when any weight in it is non-zero
        used_rows = np.zeros(nz, bool)
        used_rows[zj[zin & (zw0 != 0)]] = True
        used_rows[zj[zin & (zw1 != 0)] + 1] = True
        used_cols = (xin & ((xw0 != 0) | (xw1 != 0)))
        rows = np.nonzero(used_rows)[0]
        cols = np.nonzero(used_cols)[0]
        self.empty = rows.size == 0 or cols.size == 0
        self.n_rows = nz
        if self.empty:
            return
        self.r0, self.r1 = int(rows[0]), int(rows[-1] + 1)
        self.c0, self.c1 = int(cols[0]), int(cols[-1] + 1)
        # z taps per output row (absolute tile rows), x taps per output column in [c0, c1)
        self.z_rows = np.stack([zj, zj + 1], -1).astype(np.int32)
        self.z_w = np.stack([zw0, zw1], -1)
        self.z_rows[~zin] = self.r0
        c = slice(self.c0, self.c1)
        self.x_cols = np.stack([xj[c], xj[c] + 1], -1).astype(np.int32)
        self.x_w = np.stack([xw0[c], xw1[c]], -1)
        self.x_cols[~xin[c]] = 0
        fz = factor_z(nz, focus_pix, focus_sigma) if not np.isnan(focus_pix) else np.ones(nz)
        self.fz = fz
        self.oc = oc

    @property
    def C(self):
        return self.c1 - self.c0

    def needed_rows(self):
        """[lo, hi): magnitude rows the contribution reads (with optical-path shifts)."""
        lo, hi = self.r0, self.r1 - 1
        if self.oc is not None:
            lo += int(self.oc.shift.min())
            hi += int(self.oc.shift.max())
        lo = min(max(lo, 0), self.n_rows - 1)
        hi = min(max(hi, 0), self.n_rows - 1)
        return lo, hi + 1


class HipTileStitcher:
    """Applies a host-built TileStitcher on the hip device.  Same interface as the reference's
    TileStitcher/TorchTileStitcher: attributes empty, r0, r1, c0, c1; contribute(mag, frame_idx)
    -> (num, den) of shape (B, nOutZ, c1 - c0)."""

    def __init__(self, base: TileStitcher, xp, n_out_z: int):
        self.xp = xp
        torch = xp.torch
        self.empty = base.empty
        self.base = base
        if self.empty:
            return
        self.r0, self.r1, self.c0, self.c1 = base.r0, base.r1, base.c0, base.c1
        self.n_out_z = int(n_out_z)
        self.n_rows = base.n_rows
        dev = xp.device

        def t(a, dtype):
            return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to(dev)

        self.fz = t(base.fz, np.float32)
        self.z_rows = t(base.z_rows, np.int32)
        self.z_w = t(base.z_w, np.float32)
        self.x_cols = t(base.x_cols, np.int32)
        self.x_w = t(base.x_w, np.float32)
        if base.oc is not None:
            lo, hi = base.oc.valid_bounds()
            self.shift = t(base.oc.shift, np.int32)
            self.vlo = t(lo, np.int32)
            self.vhi = t(hi, np.int32)
            self.n_y = base.oc.shift.shape[0]
            self.n_x_opc = base.oc.shift.shape[1]
        else:
            self.shift = self.vlo = self.vhi = None

    def needed_rows(self):
        return self.base.needed_rows()

    def _desc(self, mag, mag_row0, frame_idx, num, den, ldx, out_b0, out_c0, accumulate):
        torch = self.xp.torch
        B, nX, mrows = mag.shape
        if mag.dtype != torch.float32 or not mag.is_contiguous():
            raise ValueError("mag must be a contiguous float32 device tensor")
        lo, hi = self.needed_rows()
        if mag_row0 > lo or mag_row0 + mrows < hi:
            raise ValueError(f"magnitude rows [{mag_row0}, {mag_row0 + mrows}) do not cover [{lo}, {hi})")
        if self.shift is not None and nX != self.n_x_opc:
            raise ValueError("optical-path correction grid does not match the tile's A-scans")
        if nX < int(self.base.x_cols.max()) + 1:
            raise ValueError("magnitude has fewer A-scans than the tile grid")
        fi = frame_idx.to(device=self.xp.device, dtype=torch.int32).contiguous()
        if fi.numel() != B:
            raise ValueError("frame_idx must have one entry per batch frame")
        if self.shift is not None and B:
            fmin, fmax = int(fi.min()), int(fi.max())
            if fmin < 0 or fmax >= self.n_y:
                raise ValueError("frame index outside the optical-path correction grid")
        d = _lib.StitchDesc()
        d.mag = mag.data_ptr()
        d.mag_stride = mrows
        d.mag_row0 = mag_row0
        d.mag_rows = mrows
        d.n_x = nX
        d.n_z = self.n_rows
        d.frame_idx = fi.data_ptr()
        d.shift = self.shift.data_ptr() if self.shift is not None else None
        d.valid_lo = self.vlo.data_ptr() if self.shift is not None else None
        d.valid_hi = self.vhi.data_ptr() if self.shift is not None else None
        d.fz = self.fz.data_ptr()
        d.z_rows = self.z_rows.data_ptr()
        d.z_w = self.z_w.data_ptr()
        d.x_cols = self.x_cols.data_ptr()
        d.x_w = self.x_w.data_ptr()
        d.num = num.data_ptr()
        d.den = den.data_ptr()
        d.B = B
        d.n_out_z = self.n_out_z
        d.C = self.c1 - self.c0
        d.ldx = ldx
        d.out_b0 = out_b0
        d.out_c0 = out_c0
        d.accumulate = 1 if accumulate else 0
        return d, fi

    def contribute(self, mag, frame_idx, mag_row0: int = 0):
        """mag: (B, nTileX, rows) float32 holding depth rows [mag_row0, mag_row0 + rows);
        frame_idx: (B,) tile-local y frames.  Returns (num, den), each (B, nOutZ, c1 - c0)."""
        torch = self.xp.torch
        B = mag.shape[0]
        C = self.c1 - self.c0
        num = torch.empty((B, self.n_out_z, C), dtype=torch.float32, device=self.xp.device)
        den = torch.empty_like(num)
        d, keep = self._desc(mag, mag_row0, frame_idx, num, den, C, 0, 0, False)
        _lib.check(_lib.lib().oct_stitch(ctypes.byref(d), self.xp.stream()), "oct_stitch")
        del keep
        return num, den

    def contribute_into(self, mag, frame_idx, num, den, b0: int = 0, mag_row0: int = 0):
        """Add this tile's contribution into row accumulators num/den (nF, nOutZ, nOutX) at
        frames b0.. and columns c0:c1 (octrecon/pipeline.py:283-284)."""
        nF, nz_out, nx_out = num.shape
        if nz_out != self.n_out_z or den.shape != num.shape:
            raise ValueError("accumulator shape mismatch")
        if b0 < 0 or b0 + mag.shape[0] > nF or self.c1 > nx_out:
            raise ValueError("contribution outside the accumulators")
        if not (num.is_contiguous() and den.is_contiguous()):
            raise ValueError("accumulators must be contiguous")
        d, keep = self._desc(mag, mag_row0, frame_idx, num, den, nx_out, b0, self.c0, True)
        _lib.check(_lib.lib().oct_stitch(ctypes.byref(d), self.xp.stream()), "oct_stitch")
        del keep
