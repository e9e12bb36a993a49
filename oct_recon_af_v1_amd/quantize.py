"""Finalisation, volume clim and legacy uint16 quantisation.

Host formulas (float64 numpy, used for small planes and for the TIFF metadata) and the device
path (liboct_hip finalize / clim / quantize kernels) that the pipeline uses.

Reference: octrecon/io/tiff_writer.py (_matlab_uint16 :26-30, db_to_bits :33-38,
bits_to_db :41-45, plane_clim :48-50, legacy double quantisation :62-65) and
octrecon/pipeline.py:296-300 (finalise), :376-378, 388 (clim).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .stitching import min_weight_threshold

MAXBIT = 2 ** 16 - 1


def matlab_uint16(v):
    """MATLAB uint16(): round half away from zero, saturate to [0, 65535], NaN -> 0."""
    v = np.asarray(v, dtype=np.float64)
    out = np.floor(np.abs(v) + 0.5) * np.sign(v)
    out = np.nan_to_num(out, nan=0.0)
    return np.clip(out, 0, MAXBIT).astype(np.uint16)


def db_to_bits(db, clim):
    c1, c2 = sorted(clim)
    with np.errstate(invalid="ignore", divide="ignore"):
        bits = matlab_uint16((np.asarray(db, np.float64) - c1) / (c2 - c1) * (MAXBIT - 1)).astype(np.uint32) + 1
    bits = np.minimum(bits, MAXBIT).astype(np.uint16)
    bits[np.isnan(db)] = 0
    return bits


def bits_to_db(bits, clim):
    c1, c2 = sorted(clim)
    out = (np.asarray(bits).astype(np.float64) - 1) * (c2 - c1) / (MAXBIT - 1) + c1
    out[np.asarray(bits) == 0] = np.nan
    return out


def plane_clim(db):
    fin = np.asarray(db)[np.isfinite(db)]
    return (float(fin.min()), float(fin.max())) if fin.size else (np.nan, np.nan)


class DeviceFinalizer:
    """num/den -> dB (+ plane and global clim) and dB -> uint16 codes on the hip device."""

    def __init__(self, xp):
        self.xp = xp
        self.torch = xp.torch
        self._ws = None

    def finalize(self, num, den, db=None, dtype=np.float32):
        """num/den (nF, nOutZ, nOutX) -> (db, plane_clim (nF, 2), clim (2,)) device tensors.
        den None: db = 20 log10(num) (a magnitude volume without stitching)."""
        torch = self.torch
        if num.dtype != torch.float32 or not num.is_contiguous() or (
                den is not None and (den.shape != num.shape or den.dtype != torch.float32 or not den.is_contiguous())):
            raise ValueError("num/den must be contiguous float32 tensors of equal shape")
        nF = num.shape[0]
        pe = int(np.prod(num.shape[1:]))
        if db is None:
            db = torch.empty_like(num)
        pcl = torch.empty((nF, 2), dtype=torch.float32, device=num.device)
        cl = torch.empty(2, dtype=torch.float32, device=num.device)
        L = _lib.lib()
        need = L.oct_finalize_workspace_bytes(nF, pe)
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(max(need, 1), dtype=torch.uint8, device=num.device)
        thr = float(np.float32(min_weight_threshold(dtype)))
        _lib.check(L.oct_finalize(num.data_ptr(), den.data_ptr() if den is not None else None, db.data_ptr(), nF,
                                  pe, thr, pcl.data_ptr(),
                                  cl.data_ptr(), self._ws.data_ptr(), self._ws.numel(), self.xp.stream()),
                   "oct_finalize")
        return db, pcl, cl

    def quantize(self, db, clim, plane_clims=None, legacy_double=False, out=None):
        """db (nF, ...) float32 -> uint16 codes with device clim (2,) [and plane clims (nF, 2)]."""
        torch = self.torch
        nF = db.shape[0]
        pe = int(np.prod(db.shape[1:]))
        if out is None:
            out = torch.empty(db.shape, dtype=torch.uint16, device=db.device)
        clim = self._dev2(clim)
        pcl = self._dev2(plane_clims) if legacy_double else None
        _lib.check(_lib.lib().oct_quantize(db.data_ptr(), out.data_ptr(), nF, pe, clim.data_ptr(),
                                           pcl.data_ptr() if pcl is not None else None, 1 if legacy_double else 0,
                                           self.xp.stream()), "oct_quantize")
        return out

    def _dev2(self, v):
        torch = self.torch
        if isinstance(v, torch.Tensor):
            return v.to(device=self.xp.device, dtype=torch.float32).contiguous()
        return torch.as_tensor(np.asarray(v, np.float32), device=self.xp.device).contiguous()
