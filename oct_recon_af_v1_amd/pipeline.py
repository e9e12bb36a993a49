"""Tiled-scan reconstruction driver on the hip device, for tiles resident in device memory.

Mirrors octrecon/pipeline.py: ReconConfig (:49-92, the reconstruction and execution fields),
Reconstructor.__init__ operator build (:155-190), process_row (:223-302), process_planes
(:307-317) and the clim/quantisation part of run (:376-399).  File reading, TIFF writing and
the web/CLI layers of the reference are not part of this package: tiles come from a mapping
(xi, yi, zi) -> raw tensor (nYPixels, lines, N), or a callable returning the frames asked for.

Per y-tile row: for every tile of the row (x-major, z-minor) and batch of frames, the fused
spectral kernel produces the depth rows the tile's stitcher reads, the stitch kernel adds the
tile's weighted, interpolated contribution into num/den, and the finalise kernel turns the row
into dB with per-plane finite min/max.  ``run`` accumulates every row into one volume, so one
finalise pass yields the global clim, and one quantise pass the uint16 codes.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np

from . import _lib
from .backend import get_xp, synchronize
from .geometry import ScanGrid, build_spectral_geometry, build_tiled_geometry, dispersion_window
from .quantize import DeviceFinalizer
from .spectral import HipSpectralProcessor, pair_order
from .stitching import HipTileStitcher, OpticalPathCorrection, TileStitcher, min_weight_threshold


@dataclass
class ReconConfig:
    dispersion_quadratic_term: float
    focus_positions: object                 # number | list (1-based focus pixel per depth) | "none"
    focus_sigma: float                      # pixels (effective Gaussian std sqrt(2) * sigma)
    crop_z_range_mm: object = None          # None | [zmin, zmax]
    output_pixel_size_um: object = None     # None | number (must equal the scan pixel size)
    n_medium: float = 1.33
    interp_method: str = "sinc5"
    apply_path_length_correction: bool = True
    device: str = "hip"
    precision: str = "float32"
    batch_frames: int = 50
    legacy_double_quantization: bool = False
    spectral_rows: str = "needed"           # "needed": only the depth rows stitching reads; "all"
    form: str = "plain"                     # "plain": fused FP32 kernel; "mfma": fp16 3-product operator GEMM
    fused_rows: bool = True                 # one stitch+finalise pass per y row (False: per-tile contribute)


_ROW_ARG_TILES = 40        # tiles a row launch carries in its argument block (oct_stitch_row)


class Cancelled(Exception):
    pass


class Reconstructor:
    def __init__(self, cfg: ReconConfig, scan: ScanGrid, lambda_nm, tiles, apod_size: int,
                 apod_mode: str = "lines", opc_poly=None, xp=None, log=None, ascan_binning: int = 1,
                 bscan_avg: int = 1, ascan_avg: int = 1, tile_betas=None):
        """ascan_binning, bscan_avg, ascan_avg: the tile header's spectra_avg (A-line binning),
        B-scan repeats (consecutive raw files per frame) and A-scan repeats (consecutive A-lines
        per scan position), octrecon/pipeline.py:138-141,177-183,216-219,276-279.
        tile_betas: optional {(xi, yi, zi): dispersion} (reference units) reconstructing each
        listed tile with its own dispersion window (plain form: the plan's window pointer per
        launch; MFMA form: the operator rebuilt on the device per tile); the reference applies
        cfg.dispersion_quadratic_term to every tile."""
        if cfg.precision != "float32":
            raise ValueError("the hip device computes in float32")
        self.cfg = cfg
        self.scan = scan
        self.log = log or (lambda *a: None)
        self.xp, self.device = (xp, "hip") if xp is not None else get_xp(cfg.device)
        self.torch = self.xp.torch
        self.tiles = tiles
        self.sg = build_spectral_geometry(lambda_nm, cfg.dispersion_quadratic_term, cfg.n_medium,
                                          cfg.interp_method)
        nd = len(scan.z_depths_mm)
        if isinstance(cfg.focus_positions, str) and cfg.focus_positions == "none":
            focus = np.full(nd, np.nan)
        else:
            focus = np.atleast_1d(np.asarray(cfg.focus_positions, float)).ravel()
            if focus.size == 1:
                focus = np.full(nd, focus[0])
        self.focus = focus
        self.tg = build_tiled_geometry(scan, self.sg.z_um, focus, cfg.output_pixel_size_um, cfg.crop_z_range_mm)
        poly = opc_poly if cfg.apply_path_length_correction else None
        self.oc = (OpticalPathCorrection.build(poly, self.tg.tile_x_mm, self.tg.tile_y_mm, self.tg.tile_z_mm)
                   if poly is not None else None)
        self.nb, self.na = max(1, int(bscan_avg)), max(1, int(ascan_avg))
        self.sp = HipSpectralProcessor(self.sg, apod_size, self.xp, ascan_binning=ascan_binning,
                                       apod_mode=apod_mode, apod_group=self.nb)
        self._rep = None
        self.n_out = (len(self.tg.out_y_mm), len(self.tg.out_z_mm), len(self.tg.out_x_mm))
        self.stitchers = {}
        for xi, xc in enumerate(scan.x_centers_mm):
            for zi, zd in enumerate(scan.z_depths_mm):
                base = TileStitcher(self.tg, self.oc, xc, zd, focus[zi], cfg.focus_sigma)
                self.stitchers[(xi, zi)] = HipTileStitcher(base, self.xp, self.n_out[1])
        self.fin = DeviceFinalizer(self.xp)
        self._tile_win = None
        self._tile_betas = {}
        self._streams = {}
        self.concurrent_windows = True
        self.concurrent_forms = ("mfma", "plain")
        self.max_side_streams = 4
        if tile_betas is not None:
            self.set_tile_betas(tile_betas)
        self._mag = None
        self._tile_mag = {}
        self._partial = None
        self._keep = None
        # volume rows: each row's stitch runs on its own stream so the next row's spectral work
        # overlaps it; the tile magnitudes alternate between two buffer sets by row parity
        self.overlap_rows = True
        self._stitch_stream = None
        self._stitch_done = {}              # parity -> event after the last stitch reading that set
        self._row_seq = 0
        self._frames_dev = {}               # frame lists on the device (no upload per row)
        self._table_pin = [None, None]      # pinned staging of device tile tables, alternating
        self.timers = {"spectral": 0.0, "stitch": 0.0, "finalise": 0.0}
        self.prepare_operators()

    def prepare_operators(self):
        """Create the operator (MFMA) plan of every row window the tiles use, now rather than at a
        tile's first launch: plan creation clears and builds on the device and synchronises, so
        it must not run inside a row (where side streams are busy) or under graph capture.
        Called at construction and again at the start of every volume and row (a no-op unless
        spectral_rows changed)."""
        for st in self.stitchers.values():
            if not st.empty:
                lo, hi = self._tile_rows(st)
                self.sp.operator_plan(lo, hi - lo)

    # ------------------------------------------------------------------ data access
    def set_tile_betas(self, tile_betas, update: bool = False):
        """Per-tile dispersion windows on the device: float32 (already times 1/N) for the plain
        kernel, float64 for the operator build; both from geometry.dispersion_window, as the
        plan's own window.  update=True adds to / replaces the tiles already set."""
        if update and self._tile_betas:
            tile_betas = {**self._tile_betas, **tile_betas}
        self._tile_betas = dict(tile_betas)
        keys = sorted(tile_betas)
        if not keys:
            self._tile_win = None
            return
        w = dispersion_window(self.sg.lambda_eq_nm, np.array([tile_betas[k] for k in keys], float))
        w64 = np.ascontiguousarray(np.stack([w.real, w.imag], -1))
        # / N is exact (2^-11); the plain kernels read the pair order
        w32 = np.ascontiguousarray(pair_order(w64.astype(np.float32) / np.float32(self.sg.n)))
        torch = self.torch
        self._tile_win = {
            "index": {k: i for i, k in enumerate(keys)},
            "f32": torch.from_numpy(w32).to(self.xp.device),
            "f64": torch.from_numpy(w64).to(self.xp.device),
            "max_abs": float(np.abs(w).max()),
        }

    def _use_tile_window(self, key, lo, hi):
        """Point the next spectral launch at tile `key`'s window (None: the volume's)."""
        tw = self._tile_win
        if tw is None:
            return
        i = tw["index"].get(key)
        if self.cfg.form == "mfma":
            op = self.sp.operator_plan(lo, hi - lo)
            if i is not None and tw["max_abs"] > op.max_window_abs() * (1 + 1e-12):
                raise ValueError("tile window exceeds the operator's scale bound")
            ptr = None
            if i is not None:
                ptr = tw["f64"][i].data_ptr()
                # the operator build may run on a side stream that the current stream never
                # waits for: keep the windows' memory until it has read them
                tw["f64"].record_stream(self.torch.cuda.current_stream(self.xp.device))
            op.set_window(ptr, self.xp.stream())
        else:
            ptr = None
            if i is not None:
                ptr = tw["f32"][i].data_ptr()
                tw["f32"].record_stream(self.torch.cuda.current_stream(self.xp.device))
            _lib.check(_lib.lib().oct_plan_use_window(self.sp.plan.handle, ptr), "oct_plan_use_window")

    def _restore_windows(self):
        if self._tile_win is None:
            return
        _lib.check(_lib.lib().oct_plan_use_window(self.sp.plan.handle, None), "oct_plan_use_window")
        for op in self.sp._ops.values():
            op.set_window(None, self.xp.stream())

    def frame_files(self, frames):
        """Raw file indices of tile-local frames: the bscan_avg repeats of frame f are files
        f * bscan_avg + a (octrecon/pipeline.py:216-219)."""
        f = np.asarray(frames, np.int64)
        return (f[:, None] * self.nb + np.arange(self.nb)).ravel() if self.nb > 1 else f

    def tile_frames(self, xi, yi, zi, frames):
        """Raw files (len(frames) * bscan_avg, lines, N) of the frames; a callable tile source is
        given the raw file indices."""
        src = self.tiles
        f = self.frame_files(frames)
        if callable(src):
            return src(xi, yi, zi, f)
        t = src[(xi, yi, zi)]
        if f.size and np.all(np.diff(f) == 1):
            return t[int(f[0]):int(f[-1]) + 1]
        idx = self.torch.as_tensor(f, device=t.device)
        if t.dtype == self.torch.uint16:          # no uint16 gather on the device: move the bits as int16
            return t.view(self.torch.int16)[idx].view(self.torch.uint16)
        return t[idx]

    def _positions(self, lines):
        """Scan positions per frame of raw files with `lines` A-lines (after binning, repeats)."""
        n = self.sp.n_scan_lines(lines)
        if n % self.na:
            raise ValueError(f"{n} A-lines per file after binning is not a multiple of ascan_avg {self.na}")
        return n // self.na

    def _magnitudes(self, raw, n_frames, lo, hi, dst):
        """|scan| rows [lo, hi) of the frames' raw files into dst (n_frames, positions, hi - lo),
        averaged over the B-scan and A-scan repeats (oct_average_repeats)."""
        mfma = self.cfg.form == "mfma"
        if self.nb == 1 and self.na == 1:
            if mfma:
                return self.sp.run_operator(raw, lo, hi - lo, out=dst)
            return self.sp.run(raw, _lib.OCT_OUT_MAG, lo, hi - lo, out=dst)
        shape = (raw.shape[0], self.sp.n_scan_lines(raw.shape[1]), hi - lo)
        n = shape[0] * shape[1] * shape[2]
        if self._rep is None or self._rep.numel() < n:
            self._rep = self.torch.empty(n, dtype=self.torch.float32, device=self.xp.device)
        tmp = self._rep.narrow(0, 0, n).view(shape)
        if mfma:
            self.sp.run_operator(raw, lo, hi - lo, out=tmp)
        else:
            self.sp.run(raw, _lib.OCT_OUT_MAG, lo, hi - lo, out=tmp)
        return self.sp.average_repeats(tmp, n_frames, self.nb, dst.shape[1], self.na, out=dst)

    def _mag_buffer(self, n):
        if self._mag is None or self._mag.numel() < n:
            self._mag = self.torch.empty(n, dtype=self.torch.float32, device=self.xp.device)
        return self._mag

    # ------------------------------------------------------------------ core
    def union_rows(self):
        """[lo, hi): depth rows read by any stitcher."""
        rows = [st.needed_rows() for st in self.stitchers.values() if not st.empty]
        return min(r[0] for r in rows), max(r[1] for r in rows)

    def _accumulate_row(self, yi, frames, num, den, b_off=0, cancel=None):
        cfg = self.cfg
        if cfg.form not in ("plain", "mfma"):
            raise ValueError(f"form must be plain|mfma, got {cfg.form!r}")
        frames = np.asarray(frames)
        nF = len(frames)
        B = max(1, int(cfg.batch_frames))
        fr_dev = self.torch.as_tensor(frames.astype(np.int32), device=self.xp.device)
        for (xi, zi) in self.scan.tiles_in_row(yi):
            st = self.stitchers[(xi, zi)]
            if st.empty:
                continue
            lo, hi = self._tile_rows(st)
            self._use_tile_window((xi, yi, zi), lo, hi)
            for s in range(0, nF, B):
                if cancel is not None and cancel.is_set():
                    raise Cancelled()
                fr = frames[s:s + B]
                raw = self.tile_frames(xi, yi, zi, fr)
                nX = self._positions(raw.shape[1])
                mbuf = self._mag_buffer(len(fr) * nX * (hi - lo)).narrow(0, 0, len(fr) * nX * (hi - lo))
                mag = self._magnitudes(raw, len(fr), lo, hi, mbuf.view(len(fr), nX, hi - lo))
                st.contribute_into(mag, fr_dev[s:s + len(fr)], num, den, b0=b_off + s, mag_row0=lo)

    # ------------------------------------------------------------------ fused row path
    def _tile_rows(self, st):
        """Depth rows computed for a tile: the rows its stitcher reads ("needed"), or all.  The
        operator form keeps one operator per distinct row window (tiles of one focus depth
        share it)."""
        if self.cfg.spectral_rows == "needed":
            return st.needed_rows()
        return 0, self.sp.nz

    def _side_streams(self):
        """Launches of different row windows (focus depths) run on their own streams, one set per
        form.  Operator form: one tile's GEMM (a few hundred workgroups) does not fill the chip,
        tiles of other focus depths do.  Plain form: the persistent spectral kernel's last groups
        leave CUs idle, which the next depth's launch fills (2.4 % per tiled step).  Tiles of one
        window share a stream (their operator may be rebuilt per tile); windows share
        max_side_streams streams in turn (2, 3 and 4 measured the same on the tiled workload).
        Forms outside concurrent_forms keep every launch on the caller's stream.  Off when binning
        or repeats share the processor's staging buffers."""
        return (self.cfg.form in self.concurrent_forms and self.concurrent_windows and self.nb == 1 and self.na == 1
                and self.sp.ascan_binning == 1)

    def _row_magnitudes(self, yi, frames, cancel=None, parity=None, events=None):
        """Spectral stage for every tile of row yi into per-tile buffers (nF, nX, rows).  With a
        parity (0 / 1) the buffers are that parity's set, written only after the last stitch
        that read the set has finished.  Tiles on side streams end in events that the current
        stream waits for, or, given the list `events`, that are appended to it for the consumer
        to wait for: then the next row's side work on a stream starts as soon as that stream's
        tiles of this row are done, not when the slowest stream's are."""
        torch = self.torch
        frames = np.asarray(frames)
        out = []
        side = self._side_streams()
        main = torch.cuda.current_stream(self.xp.device) if (side or parity is not None) else None
        reuse = self._stitch_done.get(parity) if parity is not None else None
        if reuse is not None and not side:
            main.wait_event(reuse)
        start = None
        used = {}
        for (xi, zi) in self.scan.tiles_in_row(yi):
            st = self.stitchers[(xi, zi)]
            if st.empty:
                continue
            lo, hi = self._tile_rows(st)
            key = (xi, zi) if parity is None else (xi, zi, parity)
            buf = self._tile_mag.get(key)
            ctx = None
            if side:
                if start is None:
                    start = torch.cuda.Event()
                    start.record(main)
                streams = self._streams.setdefault(self.cfg.form, {})   # each form its own set
                s_ = streams.get((lo, hi))
                if s_ is None:                      # windows share max_side_streams streams, in turn
                    k = len(streams)
                    if k < self.max_side_streams:
                        s_ = torch.cuda.Stream(device=self.xp.device)
                    else:
                        s_ = list(streams.values())[k % self.max_side_streams]
                    streams[(lo, hi)] = s_
                if (lo, hi) not in used:
                    s_.wait_event(start)
                    if reuse is not None:
                        s_.wait_event(reuse)
                    used[(lo, hi)] = s_
                ctx = torch.cuda.stream(s_)
                ctx.__enter__()
            try:
                buf = self._tile_spectral(xi, yi, zi, frames, lo, hi, buf, key, cancel, main)
            finally:
                if ctx is not None:
                    ctx.__exit__(None, None, None)
            out.append((st, buf, lo))
        for s_ in used.values():
            ev = torch.cuda.Event()
            ev.record(s_)
            if events is None:
                main.wait_event(ev)
            else:
                events.append(ev)
        return out

    def _tile_spectral(self, xi, yi, zi, frames, lo, hi, buf, key, cancel, main=None):
        """One tile's rows [lo, hi) of every frame into its buffer, on the current stream; a
        buffer made on a side stream is marked as used by `main` (the stitch reads it there)."""
        nF = len(frames)
        B = max(1, int(self.cfg.batch_frames))
        self._use_tile_window((xi, yi, zi), lo, hi)
        for s in range(0, nF, B):
            if cancel is not None and cancel.is_set():
                raise Cancelled()
            fr = frames[s:s + B]
            raw = self.tile_frames(xi, yi, zi, fr)
            nX = self._positions(raw.shape[1])
            shape = (nF, nX, hi - lo)
            if buf is None or tuple(buf.shape) != shape:
                old = self._tile_mag.get(key)
                if old is not None and self._stitch_stream is not None:
                    old.record_stream(self._stitch_stream)   # a stitch may still read it
                buf = self.torch.empty(shape, dtype=self.torch.float32, device=self.xp.device)
                if main is not None:
                    buf.record_stream(main)
                self._tile_mag[key] = buf
            self._magnitudes(raw, len(fr), lo, hi, buf[s:s + len(fr)])
            if main is not None and raw.is_cuda:
                # a side stream reads it and the caller's stream may never wait for that stream
                raw.record_stream(self.torch.cuda.current_stream(self.xp.device))
        return buf

    def _stitch_row(self, tiles, frames, db, plane0, partial, stream=None, after=()):
        """One fused stitch + finalise launch for a row (oct_stitch_row), on the current stream
        or, given `stream`, on that stream once the current stream's work so far (descriptors)
        and the events `after` (the row's magnitudes) are done; returns the event recorded
        after it there."""
        _, nz_out, nx_out = self.n_out
        arr = (_lib.RowTile * max(1, len(tiles)))()
        nX = None
        for i, (st, buf, lo) in enumerate(tiles):
            nX = buf.shape[1]
            if nX < int(st.base.x_cols.max()) + 1:
                raise ValueError("magnitude has fewer A-scans than the tile grid")
            nlo, nhi = st.needed_rows()
            if lo > nlo or lo + buf.shape[2] < nhi:
                raise ValueError("tile magnitude rows do not cover the stitcher's rows")
            t = arr[i]
            t.mag = buf.data_ptr()
            t.mag_stride = buf.shape[2]
            t.mag_row0 = lo
            t.c0 = st.c0
            t.C = st.c1 - st.c0
            t.z_rows = st.z_rows.data_ptr()
            t.z_w = st.z_w.data_ptr()
            t.x_cols = st.x_cols.data_ptr()
            t.x_w = st.x_w.data_ptr()
            t.fz = st.fz.data_ptr()
        fkey = np.asarray(frames, np.int32).tobytes()
        fr = self._frames_dev.get(fkey)
        if fr is None:
            fr = self._frames_dev[fkey] = self.torch.as_tensor(np.asarray(frames, np.int32), device=self.xp.device)
        st0 = tiles[0][0] if tiles else None
        shift = vlo = vhi = None
        if st0 is not None and st0.shift is not None:
            if nX != st0.n_x_opc:
                raise ValueError("optical-path correction grid does not match the tile's A-scans")
            fa = np.asarray(frames)
            if fa.size and (int(fa.min()) < 0 or int(fa.max()) >= st0.n_y):
                raise ValueError("frame index outside the optical-path correction grid")
            shift, vlo, vhi = st0.shift.data_ptr(), st0.vlo.data_ptr(), st0.vhi.data_ptr()
        thr = float(np.float32(min_weight_threshold(np.float32)))
        table = None
        if len(tiles) > _ROW_ARG_TILES:              # descriptors through a device table
            table = self._upload_table(arr)
        sh = self.xp.stream()
        done = None
        if stream is not None:
            ready = self.torch.cuda.Event()
            ready.record(self.torch.cuda.current_stream(self.xp.device))
            stream.wait_event(ready)
            for ev in after:
                stream.wait_event(ev)
            sh = stream.cuda_stream
            for t in (fr, table, db, partial):
                if t is not None:
                    t.record_stream(stream)
        if table is not None:
            _lib.check(_lib.lib().oct_stitch_row_table(arr, table.data_ptr(), len(tiles), len(frames), nX or 1,
                                                       self.sp.nz, fr.data_ptr(), shift, vlo, vhi, nz_out, nx_out,
                                                       thr, db.data_ptr(), plane0, partial.data_ptr(), sh),
                       "oct_stitch_row_table")
        else:
            _lib.check(_lib.lib().oct_stitch_row(arr, len(tiles), len(frames), nX or 1, self.sp.nz, fr.data_ptr(),
                                                 shift, vlo, vhi, nz_out, nx_out, thr, db.data_ptr(), plane0,
                                                 partial.data_ptr(), sh), "oct_stitch_row")
        if stream is not None:
            done = self.torch.cuda.Event()
            done.record(stream)
        self._keep = (arr, fr, table)
        return done

    def _upload_table(self, arr):
        """Tile descriptors -> a new device table, copied asynchronously on the current stream
        from one of two pinned staging buffers (one is refilled only after its last copy ended)."""
        torch = self.torch
        nbytes = ctypes.sizeof(arr)
        slot = self._row_seq % 2
        pin, ev = self._table_pin[slot] or (None, None)
        if pin is None or pin.numel() < nbytes:
            pin = torch.empty(max(nbytes, 4096), dtype=torch.uint8, pin_memory=True)
        elif ev is not None:
            ev.synchronize()
        ctypes.memmove(pin.data_ptr(), ctypes.addressof(arr), nbytes)
        table = torch.empty(nbytes, dtype=torch.uint8, device=self.xp.device)
        table.copy_(pin[:nbytes], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(self.xp.device))
        self._table_pin[slot] = (pin, ev)
        return table

    def _join_stitches(self):
        """The current stream waits for every row stitch issued so far (the stitch stream runs
        them in order, so for its last one)."""
        if self._stitch_stream is not None:
            last = self.torch.cuda.Event()
            last.record(self._stitch_stream)
            self.torch.cuda.current_stream(self.xp.device).wait_event(last)

    def _row_partials(self, n_planes):
        _, nz_out, nx_out = self.n_out
        nb = _lib.lib().oct_stitch_row_blocks(nz_out, nx_out)
        need = n_planes * nb * 2
        if self._partial is None or self._partial.numel() < need:
            self._partial = self.torch.empty(need, dtype=self.torch.int32, device=self.xp.device)
        return self._partial, nb

    def _clim(self, partial, n_planes, nb):
        pcl = self.torch.empty((n_planes, 2), dtype=self.torch.float32, device=self.xp.device)
        cl = self.torch.empty(2, dtype=self.torch.float32, device=self.xp.device)
        _lib.check(_lib.lib().oct_volume_clim(partial.data_ptr(), n_planes, nb, pcl.data_ptr(), cl.data_ptr(),
                                              self.xp.stream()), "oct_volume_clim")
        return pcl, cl

    def process_row(self, yi: int, frames=None, cancel=None):
        """Reconstruct the output planes of y-tile row `yi` (tile-local frames, default all).
        Returns (db, plane_clim) device tensors: (nF, nOutZ, nOutX) float32 and (nF, 2)."""
        frames = np.arange(self.scan.n_y_px) if frames is None else np.asarray(frames)
        _, nz_out, nx_out = self.n_out
        self.prepare_operators()
        try:
            if self.cfg.fused_rows:
                self._join_stitches()              # volume rows may still use the partials
                db = self.xp.empty((len(frames), nz_out, nx_out))
                partial, nb = self._row_partials(len(frames))
                self._stitch_row(self._row_magnitudes(yi, frames, cancel), frames, db, 0, partial)
                pcl, _ = self._clim(partial, len(frames), nb)
                return db, pcl
            num = self.xp.zeros((len(frames), nz_out, nx_out))
            den = self.xp.zeros((len(frames), nz_out, nx_out))
            self._accumulate_row(yi, frames, num, den, cancel=cancel)
            db, pcl, _ = self.fin.finalize(num, den)
            return db, pcl
        finally:
            self._restore_windows()

    def process_planes(self, y_indices0):
        """Arbitrary output y planes (0-based) -> host float32 (n, nOutZ, nOutX)."""
        out, by_row = {}, {}
        for y in y_indices0:
            r, f = divmod(int(y), self.scan.n_y_px)
            by_row.setdefault(r, []).append((y, f))
        for r, lst in by_row.items():
            db, _ = self.process_row(r, [f for _, f in lst])
            host = db.cpu().numpy()
            for (y, _), p in zip(lst, host):
                out[y] = p
        return np.stack([out[y] for y in y_indices0])

    def allocate_volume(self, rows=None):
        rows = list(range(len(self.scan.y_centers_mm))) if rows is None else list(rows)
        ny = self.scan.n_y_px
        _, nz_out, nx_out = self.n_out
        shape = (len(rows) * ny, nz_out, nx_out)
        vol = {k: (self.xp.zeros(shape) if not self.cfg.fused_rows else None) for k in ("num", "den")}
        vol["db"] = self.xp.empty(shape)
        vol["bits"] = self.torch.empty(shape, dtype=self.torch.uint16, device=self.xp.device)
        vol["rows"] = rows
        return vol

    def volume_row(self, vol, k, cancel=None):
        """The k-th requested row of a volume (fused path): its planes of vol["db"] and its clim
        partials.  Rows may be given in any order and at any time before finish_volume; used to
        reconstruct a streamed volume row by row as its tiles arrive.  Returns an event that
        completes once the device has finished with the row's raw tiles (its stitch, which
        follows every spectral launch of the row)."""
        if not self.cfg.fused_rows:
            raise ValueError("row-by-row reconstruction needs the fused row path")
        ny = self.scan.n_y_px
        n_planes = len(vol["rows"]) * ny
        partial, _ = self._row_partials(n_planes)
        frames = np.arange(ny)
        if not self.overlap_rows:
            self._stitch_row(self._row_magnitudes(vol["rows"][k], frames, cancel), frames, vol["db"], k * ny, partial)
            done = self.torch.cuda.Event()
            done.record(self.torch.cuda.current_stream(self.xp.device))
            return done
        parity = self._row_seq % 2
        self._row_seq += 1
        if self._stitch_stream is None:
            self._stitch_stream = self.torch.cuda.Stream(device=self.xp.device)
        ends = []
        tiles = self._row_magnitudes(vol["rows"][k], frames, cancel, parity=parity, events=ends)
        self._stitch_done[parity] = self._stitch_row(tiles, frames, vol["db"], k * ny, partial,
                                                     stream=self._stitch_stream, after=ends)
        return self._stitch_done[parity]

    def finish_volume(self, vol):
        """Clims and codes of a volume whose rows were all made by volume_row (vol["db"] is
        complete, on the current stream, from here on)."""
        ny = self.scan.n_y_px
        n_planes = len(vol["rows"]) * ny
        partial, nb = self._row_partials(n_planes)
        self._join_stitches()
        self._restore_windows()
        pcl, cl = self._clim(partial, n_planes, nb)
        self.fin.quantize(vol["db"], cl, pcl, legacy_double=self.cfg.legacy_double_quantization, out=vol["bits"])
        vol["plane_clim"], vol["clim"] = pcl, cl
        return vol

    def reconstruct_volume(self, vol, cancel=None):
        """All requested y-tile rows into one device volume: dB, plane clims, global clim and
        uint16 codes (single finalise and quantise passes).  No host synchronisation."""
        ny = self.scan.n_y_px
        self.prepare_operators()
        try:
            if self.cfg.fused_rows:
                for k in range(len(vol["rows"])):
                    self.volume_row(vol, k, cancel)
                return self.finish_volume(vol)
            else:
                vol["num"].zero_()
                vol["den"].zero_()
                for k, r in enumerate(vol["rows"]):
                    self._accumulate_row(r, np.arange(ny), vol["num"], vol["den"], b_off=k * ny, cancel=cancel)
                db, pcl, cl = self.fin.finalize(vol["num"], vol["den"], db=vol["db"])
        finally:
            self._restore_windows()
        self.fin.quantize(db, cl, pcl, legacy_double=self.cfg.legacy_double_quantization, out=vol["bits"])
        vol["plane_clim"], vol["clim"] = pcl, cl
        return vol

    def bscan_magnitudes(self, xi, yi, zi, frames):
        """|A| of every depth row of tile (xi, yi, zi)'s frames at the volume's dispersion (repeats
        averaged, binning applied): a new device tensor (len(frames), nX, N/2) float32."""
        nz = self.sp.nz
        raw = self.tile_frames(xi, yi, zi, frames)
        dst = self.torch.empty((len(frames), self._positions(raw.shape[1]), nz), dtype=self.torch.float32,
                               device=self.xp.device)
        return self._magnitudes(raw, len(frames), 0, nz, dst)

    def opc_tables(self):
        """The optical-path correction's device tables (shift, valid_lo, valid_hi), each (nY, nX)
        int32 per tile frame and A-scan (the same for every tile), or None without a correction."""
        st = next(iter(self.stitchers.values()))
        return (st.shift, st.vlo, st.vhi) if st.shift is not None else None

    def measure_focus_drift(self, prior_pix, n_frames: int = 5, tissue_ri: float = 1.4, log=None, **kw):
        """Focus pixel per focus depth measured on the central tile (octrecon/estimation/focus.py:
        221-279): magnitudes at the volume's dispersion (repeats averaged), optical-path corrected
        depth profiles on the device, confocal peaks and the drift fit on the host.  Returns
        (focus per depth, fit diagnostics, per-depth measurements)."""
        from .drift import central_frames, measure_focus_drift
        xi0 = int(np.argmin(np.abs(self.scan.x_centers_mm)))
        yi0 = int(np.argmin(np.abs(self.scan.y_centers_mm)))
        frames = central_frames(self.scan.n_y_px, n_frames)
        dz_um = float(np.median(np.diff(self.sg.z_um)))
        return measure_focus_drift(self.xp, lambda zi: self.bscan_magnitudes(xi0, yi0, zi, frames), frames,
                                   self.scan.z_depths_mm, prior_pix, dz_um, self.sp.nz, opc=self.opc_tables(),
                                   tissue_ri=tissue_ri, log=log, **kw)

    def run(self, rows=None):
        t0 = time.perf_counter()
        vol = self.reconstruct_volume(self.allocate_volume(rows))
        synchronize(self.xp)
        clim = vol["clim"].cpu().numpy().astype(float)
        return {"device": self.device, "output_shape_yzx": list(vol["db"].shape), "clim_dB": tuple(clim),
                "reconstruction_seconds": time.perf_counter() - t0, "volume": vol}
