"""Raw spectral B-scans -> depth-resolved magnitude on the MI355X (``hip`` device).

Drop-in for the reference's SpectralProcessor / TorchSpectralProcessor
(octrecon/core/spectral.py:84-165, octrecon/core/torch_backend.py:66-119): same constructor
meaning, same ``apod_size``/``n``/``nz`` attributes, same ``magnitude`` / ``linearise`` /
``complex_scan`` results, computed by the fused kernel of liboct_hip (background subtraction,
mean removal, banded sinc, complex window, 2048-point inverse FFT, crop, |.|).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .geometry import SpectralGeometry

_RAW_CODE = {np.dtype(np.int16): _lib.OCT_RAW_I16, np.dtype(np.uint16): _lib.OCT_RAW_U16,
             np.dtype(np.float32): _lib.OCT_RAW_F32}


def pair_positions(n: int = 2048) -> np.ndarray:
    """Complex position of each element of a 2048-point table (window, twiddles) in the pair order
    that the spectral kernels read.  The transform's lane owns elements i = lane + 64 q (q < 32);
    its elements q and q + 16 are adjacent, so one 16-byte load gives both: element i sits at
    2 ((q mod 16) 64 + lane) + q // 16 (oct_common.hpp pair_pos)."""
    if n != 2048:
        raise ValueError("the pair order is defined for 2048-point tables")
    i = np.arange(n)
    q, lane = i >> 6, i & 63
    return 2 * ((q & 15) * 64 + lane) + (q >> 4)


def pair_order(w: np.ndarray) -> np.ndarray:
    """(..., 2048, 2) windows in natural order -> the pair order (oct_plan_use_window(s))."""
    w = np.asarray(w)
    out = np.empty_like(w)
    out[..., pair_positions(w.shape[-2]), :] = w
    return out


def natural_order(w: np.ndarray) -> np.ndarray:
    """Inverse of pair_order (windows read back from the device, e.g. oct_dispersion_windows)."""
    w = np.asarray(w)
    return w[..., pair_positions(w.shape[-2]), :]


def _padded_band(sg: SpectralGeometry, dtype=np.float32):
    """Band operator padded to the kernel's tap window (12 or 16): weights are appended as
    zeros and the window start is pulled left where it would run past N-1, with the leading
    positions weighted 0 (zero terms leave the tap sums unchanged)."""
    N, T = sg.band_w.shape
    TP = 12 if T <= 12 else 16
    if T > 16:
        raise ValueError(f"sinc band of {T} taps exceeds the kernel's 16-tap window")
    start = sg.band_start.astype(np.int64) + 0
    w = np.zeros((N, TP), np.float64)
    newstart = np.minimum(start, N - TP)
    off = start - newstart
    rows = np.arange(N)
    for t in range(T):
        w[rows, off + t] = sg.band_w[:, t]
    return newstart.astype(np.int32), np.ascontiguousarray(w.astype(dtype)), TP


class SpectralPlan:
    """Device copy of the spectral operators (owns an oct_plan)."""

    def __init__(self, sg: SpectralGeometry, device_index: int = 0):
        if sg.n != 2048:
            raise ValueError(f"the hip spectral kernel supports N = 2048 spectral pixels (got {sg.n})")
        self.n = sg.n
        start, w, TP = _padded_band(sg)
        win = np.ascontiguousarray(np.stack([sg.window.real, sg.window.imag], -1).astype(np.float32))
        self.taps = TP
        self._start, self._w, self._win = start, w, win
        h = ctypes.c_void_p()
        L = _lib.lib()
        _lib.check(L.oct_plan_create(ctypes.byref(h), device_index, self.n, TP, start.ctypes.data,
                                     w.ctypes.data, win.ctypes.data), "oct_plan_create")
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.lib().oct_plan_destroy(h)
            except Exception:  # noqa: BLE001
                pass
            self.handle = None


class OperatorPlan:
    """Device copy of the fused spectral operator A for depth bins [z0, z0 + nzc) (owns an
    oct_opplan): built on the device in float64, applied as a split-precision fp16 GEMM."""

    def __init__(self, sg: SpectralGeometry, z0: int, nzc: int, device_index: int = 0, want_fp64: bool = False):
        if sg.n != 2048:
            raise ValueError(f"the operator form supports N = 2048 spectral pixels (got {sg.n})")
        start, w, TP = _padded_band(sg, np.float64)
        win = np.ascontiguousarray(np.stack([sg.window.real, sg.window.imag], -1).astype(np.float64))
        self.z0, self.nzc, self.n = int(z0), int(nzc), sg.n
        self.op64 = np.empty((nzc, sg.n, 2), np.float64) if want_fp64 else None
        sa = ctypes.c_int(0)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().oct_opplan_create(ctypes.byref(h), device_index, sg.n, TP, start.ctypes.data,
                                                w.ctypes.data, win.ctypes.data, self.z0, self.nzc, ctypes.byref(sa),
                                                self.op64.ctypes.data if want_fp64 else None),
                   "oct_opplan_create")
        self.handle = h
        self.scale_log2 = sa.value

    def max_window_abs(self) -> float:
        v = ctypes.c_double(0.0)
        _lib.check(_lib.lib().oct_opplan_scale_bound(self.handle, ctypes.byref(v)), "oct_opplan_scale_bound")
        return v.value

    def set_window(self, window_ptr, stream):
        """Rebuild the operator on `stream` for the (N, 2) float64 device window at window_ptr
        (None: the creation window); |window| must stay within max_window_abs()."""
        _lib.check(_lib.lib().oct_opplan_set_window(self.handle, window_ptr, stream), "oct_opplan_set_window")

    def operator_fp64(self):
        """Complex float64 operator (nzc, N) as built on the device (when requested)."""
        return None if self.op64 is None else self.op64[..., 0] + 1j * self.op64[..., 1]

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.lib().oct_opplan_destroy(h)
            except Exception:  # noqa: BLE001
                pass
            self.handle = None


class HipSpectralProcessor:
    """Same interface/results as SpectralProcessor.magnitude, on the hip device (float32).

    sg: SpectralGeometry; apod_size: background A-lines at the start of each raw file
    (apod_mode 'lines') or ignored (apod_mode 'frame_mean': background = mean of all A-lines
    of each group of apod_group files).  ascan_binning > 1: the scan lines are binned on the
    device first (oct_bin_lines; the background is taken from the unbinned lines, as
    octrecon/core/spectral.py:49-71 does)."""

    def __init__(self, sg: SpectralGeometry, apod_size: int, xp, ascan_binning: int = 1,
                 apod_mode: str = "lines", apod_group: int = 1):
        if apod_mode not in ("lines", "frame_mean"):
            raise ValueError(apod_mode)
        if int(ascan_binning) < 1:
            raise ValueError(f"ascan_binning {ascan_binning}")
        self.xp = xp
        self.torch = xp.torch
        self.dtype = np.dtype(np.float32)
        self.apod_mode = apod_mode
        self.apod_group = max(1, int(apod_group))
        self.ascan_binning = int(ascan_binning)
        self._binned = None
        self.raw_apod_size = apod_size
        self.apod_size = 1 if apod_mode == "frame_mean" else apod_size
        self.n = sg.n
        self.nz = self.n // 2
        dev = xp.device
        self.sg = sg
        self.plan = SpectralPlan(sg, dev.index if dev.index is not None else 0)
        self._ops = {}
        self._ws = None
        # optional list collecting (start, end) events around each fused-kernel launch on the
        # current stream (the stream the kernel is enqueued on)
        self.timing_events = None

    # ------------------------------------------------------------------ helpers
    def _device_raw(self, raw):
        torch = self.torch
        t = raw if isinstance(raw, torch.Tensor) else self.xp.asarray(np.ascontiguousarray(raw))
        if t.device != self.xp.device:
            t = t.to(self.xp.device)
        if t.dtype == torch.int32:       # the reference's torch backend widens uint16 to int32
            t = t.to(torch.float32)
        if t.dtype not in (torch.int16, torch.uint16, torch.float32):
            t = t.to(torch.float32)
        t = t.contiguous()
        if t.data_ptr() % 16:
            t = t.clone()
        return t

    def _code(self, t):
        torch = self.torch
        return {torch.int16: _lib.OCT_RAW_I16, torch.uint16: _lib.OCT_RAW_U16,
                torch.float32: _lib.OCT_RAW_F32}[t.dtype]

    def _workspace(self, n_files):
        """Background-row workspace, one per stream (launches on different streams may overlap)."""
        need = _lib.lib().oct_spectral_workspace_bytes(self.plan.handle, n_files)
        if not isinstance(self._ws, dict):
            self._ws = {}
        key = self.xp.stream()
        ws = self._ws.get(key)
        if ws is None or ws.numel() < need:
            ws = self.torch.empty(max(need, 1), dtype=self.torch.uint8, device=self.xp.device)
            self._ws[key] = ws
        return ws

    def n_scan_lines(self, lines: int) -> int:
        """Scan A-lines per raw file of `lines` lines, after binning."""
        n = lines - self.raw_apod_size if self.apod_mode == "lines" else lines
        return _lib.lib().oct_binned_lines(n, self.ascan_binning) if self.ascan_binning > 1 else n

    def _bin(self, t):
        """Binned float32 copy of a raw batch: the background lines, then the binned scan lines."""
        B, lines, N = t.shape
        keep = self.raw_apod_size if self.apod_mode == "lines" else 0
        shape = (B, keep + self.n_scan_lines(lines), N)
        if self._binned is None or tuple(self._binned.shape) != shape:
            self._binned = self.torch.empty(shape, dtype=self.torch.float32, device=self.xp.device)
        _lib.check(_lib.lib().oct_bin_lines(t.data_ptr(), self._code(t), B, lines, keep, keep, self.ascan_binning,
                                            self._binned.data_ptr(), self.xp.stream()), "oct_bin_lines")
        return self._binned

    def average_repeats(self, mag, n_frames: int, bscan_avg: int, size_x: int, ascan_avg: int, out=None):
        """octrecon/core/spectral.py:168-177 on the device: (n_frames nb, size_x na, rows) ->
        (n_frames, size_x, rows) (oct_average_repeats)."""
        nb, na = max(1, int(bscan_avg)), max(1, int(ascan_avg))
        if nb == 1 and na == 1 and out is None:
            return mag
        rows = mag.shape[-1]
        if tuple(mag.shape) != (n_frames * nb, size_x * na, rows) or mag.stride(2) != 1 or \
                mag.stride(0) != mag.shape[1] * mag.stride(1):
            raise ValueError(f"magnitudes {tuple(mag.shape)} do not hold {n_frames} x {nb} files of {size_x} x {na} lines")
        if out is None:
            out = self.torch.empty((n_frames, size_x, rows), dtype=self.torch.float32, device=self.xp.device)
        _lib.check(_lib.lib().oct_average_repeats(mag.data_ptr(), out.data_ptr(), n_frames, nb, size_x, na, rows,
                                                  mag.stride(1), out.stride(1), self.xp.stream()),
                   "oct_average_repeats")
        return out

    def run(self, raw, kind: int, row0: int = 0, rows: int | None = None, out=None, out_stride=None):
        """Launch the fused kernel on the current stream; returns the output tensor."""
        torch = self.torch
        t = self._device_raw(raw)
        if t.dim() != 3 or t.shape[-1] != self.n:
            raise ValueError(f"raw must be (B, lines, {self.n}), got {tuple(t.shape)}")
        B, lines, _ = t.shape
        nX = self.n_scan_lines(lines)
        if kind == _lib.OCT_OUT_LINEARISED:
            rows = self.n
            row0 = 0
        elif rows is None:
            rows = self.nz - row0
        if out is None:
            if kind == _lib.OCT_OUT_COMPLEX:
                out = torch.empty((B, nX, rows), dtype=torch.complex64, device=self.xp.device)
            else:
                out = torch.empty((B, nX, rows), dtype=torch.float32, device=self.xp.device)
            out_stride = rows
        elif out_stride is None:
            out_stride = rows
        if B == 0 or nX <= 0:
            if nX <= 0:
                raise ValueError(f"raw has {lines} lines: no scan A-lines after {self.raw_apod_size} background lines")
            return out
        ws = self._workspace(B)
        mode = _lib.OCT_APOD_LINES if self.apod_mode == "lines" else _lib.OCT_APOD_FRAME_MEAN
        L = _lib.lib()
        code = self._code(t)
        _lib.check(L.oct_background(self.plan.handle, t.data_ptr(), code, B, lines, mode, self.raw_apod_size,
                                    self.apod_group, ws.data_ptr(), self.xp.stream()), "oct_background")
        if self.ascan_binning > 1:             # background from the unbinned lines, then bin
            t = self._bin(t)
            code = _lib.OCT_RAW_F32
            lines = t.shape[1]
        ev = self.timing_events
        if ev is not None:
            e0 = torch.cuda.Event(enable_timing=True)
            e0.record()
        _lib.check(L.oct_spectral(self.plan.handle, t.data_ptr(), code, B, lines, mode, self.raw_apod_size, kind,
                                  row0, rows, out.data_ptr(), out_stride, ws.data_ptr(), self.xp.stream()),
                   "oct_spectral")
        if ev is not None:
            e1 = torch.cuda.Event(enable_timing=True)
            e1.record()
            ev.append((e0, e1))
        return out

    def background(self, raw, out=None):
        """Background rows (B, N) float32 of a raw batch (oct_background), for run_repeat."""
        t = self._device_raw(raw)
        B, lines, _ = t.shape
        if out is None:
            out = self.torch.empty((B, self.n), dtype=self.torch.float32, device=self.xp.device)
        mode = _lib.OCT_APOD_LINES if self.apod_mode == "lines" else _lib.OCT_APOD_FRAME_MEAN
        _lib.check(_lib.lib().oct_background(self.plan.handle, t.data_ptr(), self._code(t), B, lines, mode,
                                             self.raw_apod_size, self.apod_group, out.data_ptr(), self.xp.stream()),
                   "oct_background")
        return out

    def run_repeat(self, raw, n_repeat: int, bg, out, kind: int = _lib.OCT_OUT_MAG, row0: int = 0,
                   rows: int | None = None, x0: int = 0, nx: int | None = None):
        """The batch reconstructed n_repeat times in one launch (windows per repeat / file from
        oct_plan_use_windows), scan lines [x0, x0 + nx) of every file; out (n_repeat B, nx, rows);
        bg from background()."""
        t = self._device_raw(raw)
        B, lines, _ = t.shape
        nX = self.n_scan_lines(lines) - x0 if nx is None else nx
        rows = self.nz - row0 if rows is None else rows
        if tuple(out.shape) != (n_repeat * B, nX, rows) or not out.is_contiguous():
            raise ValueError(f"out must be contiguous ({n_repeat * B}, {nX}, {rows}), got {tuple(out.shape)}")
        if tuple(bg.shape) != (B, self.n):
            raise ValueError(f"bg must be ({B}, {self.n})")
        mode = _lib.OCT_APOD_LINES if self.apod_mode == "lines" else _lib.OCT_APOD_FRAME_MEAN
        code = self._code(t)
        if self.ascan_binning > 1:
            t = self._bin(t)
            code = _lib.OCT_RAW_F32
            lines = t.shape[1]
        _lib.check(_lib.lib().oct_spectral_repeat(self.plan.handle, t.data_ptr(), code, B, lines, mode,
                                                  self.raw_apod_size, n_repeat, x0, nX, kind, row0, rows,
                                                  out.data_ptr(), rows, bg.data_ptr(), self.xp.stream()),
                   "oct_spectral_repeat")
        return out

    def operator_plan(self, z0: int, nzc: int, want_fp64: bool = False) -> OperatorPlan:
        key = (int(z0), int(nzc))
        op = self._ops.get(key)
        if op is None or (want_fp64 and op.op64 is None):
            dev = self.xp.device
            op = OperatorPlan(self.sg, z0, nzc, dev.index if dev.index is not None else 0, want_fp64)
            self._ops[key] = op
        return op

    def run_operator(self, raw, z0: int = 0, nzc: int | None = None, out=None, out_stride=None):
        """Magnitudes of bins [z0, z0 + nzc) through the operator form (MFMA, fp16 3-product)."""
        torch = self.torch
        nzc = self.nz - z0 if nzc is None else int(nzc)
        t = self._device_raw(raw)
        if t.dim() != 3 or t.shape[-1] != self.n:
            raise ValueError(f"raw must be (B, lines, {self.n}), got {tuple(t.shape)}")
        if t.dtype == torch.float32:
            raise ValueError("the operator form takes integer raw counts (int16 / uint16)")
        B, lines, _ = t.shape
        nX = self.n_scan_lines(lines)
        if out is None:
            out = torch.empty((B, nX, nzc), dtype=torch.float32, device=self.xp.device)
            out_stride = nzc
        elif out_stride is None:
            out_stride = nzc
        if B == 0 or nX <= 0:
            if nX <= 0:
                raise ValueError(f"raw has {lines} lines: no scan A-lines after {self.raw_apod_size} background lines")
            return out
        op = self.operator_plan(z0, nzc)
        ws = self._workspace(B)
        mode = _lib.OCT_APOD_LINES if self.apod_mode == "lines" else _lib.OCT_APOD_FRAME_MEAN
        L = _lib.lib()
        code = self._code(t)
        _lib.check(L.oct_background(self.plan.handle, t.data_ptr(), code, B, lines, mode, self.raw_apod_size,
                                    self.apod_group, ws.data_ptr(), self.xp.stream()), "oct_background")
        if self.ascan_binning > 1:             # background from the unbinned lines, then bin
            t = self._bin(t)
            code = _lib.OCT_RAW_F32
            lines = t.shape[1]
        ev = self.timing_events
        if ev is not None:
            e0 = torch.cuda.Event(enable_timing=True)
            e0.record()
        _lib.check(L.oct_opspectral(op.handle, t.data_ptr(), code, B, lines, mode, self.raw_apod_size,
                                    out.data_ptr(), out_stride, ws.data_ptr(), self.xp.stream()), "oct_opspectral")
        if ev is not None:
            e1 = torch.cuda.Event(enable_timing=True)
            e1.record()
            ev.append((e0, e1))
        return out

    # ------------------------------------------------------------------ reference API
    def magnitude(self, raw):
        """raw (B, interfSize, N) int16/uint16/float32 -> |scan| (B, nX, N/2) float32 device tensor."""
        return self.run(raw, _lib.OCT_OUT_MAG)

    def magnitude_rows(self, raw, row0: int, rows: int):
        """Depth rows [row0, row0 + rows) of magnitude(raw) only (what a stitcher reads)."""
        return self.run(raw, _lib.OCT_OUT_MAG, row0, rows)

    def linearise(self, raw):
        """raw -> equispaced interferogram (B, nX, N) float32 (no window)."""
        return self.run(raw, _lib.OCT_OUT_LINEARISED)

    def complex_scan(self, raw):
        """raw -> complex A-scans (B, nX, N/2) complex64 (window, inverse FFT, crop)."""
        return self.run(raw, _lib.OCT_OUT_COMPLEX)
