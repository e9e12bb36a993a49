"""Streamed tiled volume: tiles that are not in device memory when the job starts.

One process drives one GPU with three threads:
  producer  the host generator (liboct_hostgen, OpenMP, GIL released) fills a ring of pinned
            host tiles;
  copier    each pinned tile goes to its row's slot of the raw row ring in HBM with an
            asynchronous copy on its own stream; the pinned slot is handed back when the copy is
            done; a row slot is refilled only after the device has finished reading the row
            that held it (the copy stream waits for that row's events);
  main      as soon as a y-tile row has landed, the dispersion and focus of each of its tiles
            come from the learned estimator on `est_bscans` B-scans per tile (the sharpness
            search and confocal detector for the B-scans it is not confident about), and the
            row is reconstructed at once, each tile at its own dispersion (MFMA operator rebuilt
            on the device per tile), while the next rows stream; the focus per depth is the
            median of the first row's tiles of that depth (every row holds all depths).  After
            the last row: one clim and one quantisation pass over the volume.  rowwise=False
            reconstructs after the stream instead, with the focus per depth from every row.

The reference reads raw files with io_threads reader threads and a bounded read-ahead
(octrecon/pipeline.py:197-212, 240-288), copies each batch with a pageable default-stream copy
and runs spectral + stitching per batch; dispersion and focus are measured before the run
(octrecon/estimation/dispersion.py:154-267, octrecon/estimation/focus.py:221-279) and one
dispersion is applied to the whole volume.  Here the raw tiles live in a ring of
`resident_rows` y-tile rows in HBM (default: every row, so that reconstructing after the stream
reads HBM, not the host link; a 12 x 9 x 8 tile volume of 40 x 1000 A-scans is 145 GB of uint16).
With fewer resident rows the device holds that many rows of raw data at a time, so a volume of
any size streams through; the plain FP32 validation then runs row by row beside the MFMA
reconstruction, before the row's slot is refilled.

    python -m oct_recon_af_v1_amd.stream [--tiles-x 12 --tiles-y 9 --depths 8 --bscans 40 --ascans 1000]
"""
from __future__ import annotations

import json
import queue
import threading
import time
from dataclasses import asdict, dataclass

import numpy as np

from . import synthetic as SY
from .geometry import ScanGrid, beta_from_index_units, build_spectral_geometry
from .quantize import DeviceFinalizer


@dataclass
class StreamSpec:
    tiles_x: int = 12
    tiles_y: int = 9
    depths: int = 8
    n_y: int = 40                   # B-scans per tile
    n_x: int = 1000                 # A-scans per B-scan
    overlap: float = 0.1            # lateral (x) overlap of neighbouring tiles
    depth_step_mm: float = 0.05     # focus planes
    beta_idx: float = 1e-4          # mean dispersion (phase per squared equispaced-k index)
    beta_spread: float = 0.2        # tile dispersion = beta_idx (1 + spread u), u ~ U(-1, 1)
    focus0: float = 420.0           # focus depth bin (0-based) of the first focus plane
    focus_drift: float = -12.0      # focus bins per focus plane
    focus_jitter: float = 4.0       # per tile, uniform +-
    focus_sigma_um: float = 30.0
    crop_z_mm: tuple = (-0.05, 0.4)
    seed: int = 1

    @property
    def n_tiles(self) -> int:
        return self.tiles_x * self.tiles_y * self.depths


class StreamedVolume:
    def __init__(self, xp, ss: StreamSpec = StreamSpec(), spec: SY.SpectrumSpec = SY.SpectrumSpec(),
                 fspec: SY.FocusSpec = SY.FocusSpec(), hspec: SY.HostTileSpec = SY.HostTileSpec(), ring: int = 8,
                 gen_threads: int | None = None, estimator: str = "learned", est_bscans: int = 8,
                 form: str = "mfma", log=None, readers: dict | None = None, io_threads: int = 8,
                 rowwise: bool = True, resident_rows: int | None = None):
        """readers: optional {(xi, yi, zi): rawio.OcityTile | rawio.SrrTile} streaming the tiles
        from their raw files (n_y files each) instead of the host generator; the generator's
        per-tile dispersion / focus stay the 'given' estimates.  resident_rows: y-tile rows of
        raw data held in HBM at once (default all; at least 1; rowwise=False needs all)."""
        if estimator not in ("learned", "given"):
            raise ValueError(f"estimator must be learned|given, got {estimator!r}")
        if ring < 2:
            raise ValueError("the ring needs at least 2 pinned tiles")
        self.xp, self.torch = xp, xp.torch
        self.ss, self.spec = ss, spec
        self.form = form
        self.log = log or (lambda *a: None)
        self.lam = spec.lambda_nm()
        self.to_ref = beta_from_index_units(1.0, self.lam)
        px_mm = 1e-3
        tx, ty = ss.n_x * px_mm, ss.n_y * px_mm
        xc = (np.arange(ss.tiles_x) - (ss.tiles_x - 1) / 2) * tx * (1 - ss.overlap)
        yc = (np.arange(ss.tiles_y) - (ss.tiles_y - 1) / 2) * ty
        zd = np.arange(ss.depths) * ss.depth_step_mm
        self.grid = ScanGrid(xc, yc, zd, tx, ty, ss.n_x, ss.n_y, pixel_size_um=1.0)
        self.keys = [(xi, yi, zi) for yi in range(ss.tiles_y) for (xi, zi) in self.grid.tiles_in_row(yi)]
        rng = np.random.default_rng(ss.seed)
        n = len(self.keys)
        beta = ss.beta_idx * (1 + ss.beta_spread * rng.uniform(-1, 1, n))
        focus = np.array([ss.focus0 + ss.focus_drift * k[2] for k in self.keys]) + rng.uniform(-1, 1, n) * ss.focus_jitter
        seeds = rng.integers(1, 2 ** 62, n)
        self.truth = {"beta_idx": beta, "focus": focus, "seed": seeds}
        if gen_threads is None:               # one host core stays with the copier and main threads
            import os
            gen_threads = max(1, min(16, os.cpu_count() or 1) - 1)
        self.gen = SY.HostTileGenerator(spec, fspec, hspec, threads=gen_threads)
        self.lines = spec.apod + ss.n_x
        self.n = spec.n
        torch = self.torch
        self.readers = readers
        self.io_pool = None
        slot_lines = self.lines
        if readers is not None:
            from concurrent.futures import ThreadPoolExecutor
            missing = [k for k in self.keys if k not in readers]
            if missing:
                raise ValueError(f"no reader for tiles {missing[:4]}")
            for k in self.keys:
                r = readers[k]
                if r.interf_size != self.lines or r.n_lambda != self.n:
                    raise ValueError(f"tile {k}: files of {r.interf_size} x {r.n_lambda}, the volume needs "
                                     f"{self.lines} x {self.n}")
                if r.kind == "srr":
                    slot_lines = max(slot_lines, r.file_lines)
            self.io_pool = ThreadPoolExecutor(max(1, int(io_threads)))
        # files hold int16 counts (Thorlabs); the generator writes uint16
        self.raw_dtype = torch.int16 if readers is not None else torch.uint16
        self.per_row = len(self.grid.tiles_in_row(0))
        R = ss.tiles_y if resident_rows is None else int(resident_rows)
        if not 1 <= R <= ss.tiles_y:
            raise ValueError(f"resident_rows must lie in [1, {ss.tiles_y}]")
        if not rowwise and R < ss.tiles_y:
            raise ValueError("reconstructing after the stream needs every row resident")
        self.resident_rows = R
        self.slots = torch.empty((R, self.per_row, ss.n_y, self.lines, self.n), dtype=self.raw_dtype,
                                 device=xp.device)
        self.tile_map = {}                     # key -> the device tile of a row being reconstructed
        self.ring = [torch.empty(ss.n_y * slot_lines * self.n, dtype=self.raw_dtype, pin_memory=True)
                     for _ in range(ring)]
        self._payload = None
        self.estimator = estimator
        self.est_bscans = int(est_bscans)
        self.est = None
        self.fallback = None
        if estimator == "learned":
            from .estimation import HipDispersionScorer
            from .learned import LearnedEstimator
            self.est = LearnedEstimator(xp, self.lam, spec.apod)
            self.est_x = 256                            # the network's crop width
            if ss.n_x < self.est_x:
                raise ValueError(f"the estimator needs B-scans of at least {self.est_x} A-scans")
            self.fallback = HipDispersionScorer(self.lam, self.slots[0, 0, :1, :spec.apod + self.est_x], spec.apod,
                                                xp)
        self._est_buf = None
        self._err = None
        self.rowwise = bool(rowwise)
        self._release = {}                     # row -> (threading.Event, [device events]) of its slot
        self._release_lock = threading.Lock()

    def tile_raw(self, i: int):
        """Device raw tile i (n_y, lines, N): its row's slot of the ring (valid while that row is
        resident)."""
        return self.slots[(i // self.per_row) % self.resident_rows, i % self.per_row]

    # ------------------------------------------------------------------ threads
    def _guard(self, fn, *args):
        try:
            fn(*args)
        except BaseException as e:  # noqa: BLE001 - re-raised on the main thread
            self._err = e
            for q in args:
                if isinstance(q, queue.Queue):
                    q.put(None)

    def _produce(self, free_q, ready_q, stats):
        for i in range(len(self.keys)):
            slot = free_q.get()
            if slot is None or self._err is not None:
                break
            t0 = time.perf_counter()
            self._fill(i, self.ring[slot])
            stats["gen_s"] += time.perf_counter() - t0
            ready_q.put((i, slot))
        ready_q.put(None)

    def _slot_view(self, i, flat):
        """The ring slot as tile i's host layout: (n_y, lines, N), or the SRR payload."""
        r = self.readers[self.keys[i]] if self.readers is not None else None
        rows = r.file_lines if (r is not None and r.kind == "srr") else self.lines
        return flat[:self.ss.n_y * rows * self.n].view(self.ss.n_y, rows, self.n)

    def _fill(self, i, flat):
        view = self._slot_view(i, flat)
        if self.readers is None:
            self.gen.generate(view, self.ss.n_x, self.truth["beta_idx"][i], self.truth["focus"][i],
                              int(self.truth["seed"][i]))
        else:
            self.readers[self.keys[i]].read(range(self.ss.n_y), view, self.io_pool)

    def _land(self, i, flat, stream):
        """Host slot -> tile i in HBM, on `stream` (SRR payloads unpacked on the device)."""
        torch = self.torch
        src = self._slot_view(i, flat)
        r = self.readers[self.keys[i]] if self.readers is not None else None
        if r is not None and r.kind == "srr":
            from . import _lib
            if self._payload is None or tuple(self._payload.shape) != tuple(src.shape):
                self._payload = torch.empty(src.shape, dtype=torch.int16, device=self.xp.device)
            self._payload.copy_(src, non_blocking=True)
            _lib.check(_lib.lib().oct_srr_unpack(self._payload.data_ptr(), self.ss.n_y, r.file_lines, self.n, r.apod[0],
                                                 r.apod[1], r.scan[0], r.scan[1], self.tile_raw(i).data_ptr(),
                                                 stream.cuda_stream), "oct_srr_unpack")
        else:
            self.tile_raw(i).copy_(src, non_blocking=True)

    def _copy(self, ready_q, free_q, landed_q, stats):
        """Issue each ready tile's copy at once (copies queue back to back on the copy stream),
        and hand slots back / report tiles as their copies complete."""
        from collections import deque
        torch = self.torch
        stream = torch.cuda.Stream(device=self.xp.device)
        pending = deque()                     # (tile, slot, start event, end event, queued behind)
        done = False
        def report():
            i, slot, e0, e1, prev = pending.popleft()
            t = e0.elapsed_time(e1)                 # from the later of its record and the
            if prev is not None:                    # previous copy's end: its own time
                t = min(t, prev.elapsed_time(e1))
            stats["h2d_s"] += t / 1e3
            free_q.put(slot)
            landed_q.put(i)

        with torch.cuda.stream(stream):
            while not done or pending:
                while pending and pending[0][3].query():
                    report()
                if done:
                    time.sleep(1e-4)          # the GIL stays free for the others
                    continue
                try:
                    item = ready_q.get(timeout=1e-4) if pending else ready_q.get()
                except queue.Empty:
                    continue
                if item is None or self._err is not None:
                    done = True
                    continue
                i, slot = item
                row = i // self.per_row
                if i % self.per_row == 0 and row >= self.resident_rows:
                    # the row slot is refilled only after the device has read the row that held
                    # it; the copies in flight (earlier rows) are reported first, since the main
                    # thread queues a row (and frees its slot) only once all its tiles have landed
                    while pending:
                        pending[0][3].synchronize()
                        report()
                    flag, evs = self._release_of(row - self.resident_rows)
                    flag.wait()
                    for ev in evs:
                        stream.wait_event(ev)
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                prev = pending[-1][3] if pending else None
                e0.record(stream)
                self._land(i, self.ring[slot], stream)
                e1.record(stream)
                pending.append((i, slot, e0, e1, prev))
        landed_q.put(None)

    def _release_of(self, row):
        """(host flag, device events) that free row `row`'s slot: set by the main thread once the
        row's reconstruction is queued (or on failure, with no events)."""
        with self._release_lock:
            return self._release.setdefault(row, (threading.Event(), []))

    def _release_row(self, row, events):
        flag, evs = self._release_of(row)
        evs.extend(events)
        flag.set()

    # ------------------------------------------------------------------ estimation
    def _gather(self, idx):
        """The estimator's B-scans of tiles `idx`: est_bscans frames spread over each tile, the
        background lines and the central est_x scan lines -> (len(idx) E, apod + est_x, N)."""
        torch = self.torch
        E, a, w = self.est_bscans, self.spec.apod, self.est_x
        shape = (len(idx) * E, a + w, self.n)
        if self._est_buf is None or tuple(self._est_buf.shape) != shape:
            self._est_buf = torch.empty(shape, dtype=self.raw_dtype, device=self.xp.device)
        buf = self._est_buf.view(torch.int16)             # uint16 bits moved as int16
        fr = torch.as_tensor(np.round(np.linspace(0, self.ss.n_y - 1, E)).astype(np.int64), device=self.xp.device)
        x0 = a + (self.ss.n_x - w) // 2
        for k, i in enumerate(idx):
            rv = self.tile_raw(i).view(torch.int16)
            buf[k * E:(k + 1) * E, :a] = rv[:, :a].index_select(0, fr)
            buf[k * E:(k + 1) * E, a:] = rv[:, x0:x0 + w].index_select(0, fr)
        return self._est_buf

    def _estimate(self, idx, stats):
        """Per tile of `idx`: median dispersion (index units) and focus over its B-scans, and the
        number of B-scans that fell back to the search."""
        from .learned import make_labels, take_bscans
        E = self.est_bscans
        buf = self._gather(idx)
        t0 = time.perf_counter()
        p = self.est.predict(buf, graph=True)
        t1 = time.perf_counter()
        beta, focus = p["beta_idx"].copy(), p["focus"].copy()
        bad = np.nonzero(~p["confident"])[0]
        if bad.size:
            self.fallback.set_raw(take_bscans(buf, bad) if bad.size < buf.shape[0] else buf)
            lab = make_labels(self.fallback)
            beta[bad] = lab["beta_idx"]
            focus[bad] = np.where(np.isfinite(lab["focus"]), lab["focus"], focus[bad])
        stats["net_s"] += t1 - t0
        stats["fallback_s"] += time.perf_counter() - t1
        n = len(idx)
        nb = np.bincount(bad // E, minlength=n) if bad.size else np.zeros(n, int)
        return np.median(beta.reshape(n, E), 1), np.median(focus.reshape(n, E), 1), nb

    def warmup(self):
        """Capture the estimator's graph for the row batch shape (outside any timed region)."""
        if self.est is None:
            return
        per_row = len(self.grid.tiles_in_row(0))
        self._fill(0, self.ring[0])
        self._land(0, self.ring[0], self.torch.cuda.current_stream())
        self.est.predict(self._gather([0] * per_row), graph=True)
        self.xp.synchronize()

    # ------------------------------------------------------------------ reconstruction
    def reconstructor(self, tile_beta_idx, focus_per_depth, form=None, idx=None):
        from .pipeline import ReconConfig, Reconstructor
        ss = self.ss
        b0 = beta_from_index_units(ss.beta_idx, self.lam)
        sg = build_spectral_geometry(self.lam, b0, 1.33)
        dz_um = float(sg.z_um[1] - sg.z_um[0])
        cfg = ReconConfig(dispersion_quadratic_term=b0, focus_positions=[float(f) for f in focus_per_depth],
                          focus_sigma=ss.focus_sigma_um / dz_um, crop_z_range_mm=list(ss.crop_z_mm),
                          output_pixel_size_um=1.0, n_medium=1.33, batch_frames=ss.n_y, form=form or self.form)
        use = range(len(self.keys)) if idx is None else idx
        for i in use:
            self.tile_map[self.keys[i]] = self.tile_raw(i)
        betas = {self.keys[i]: float(tile_beta_idx[i]) * self.to_ref for i in use}
        return Reconstructor(cfg, self.grid, self.lam, self.tile_map, self.spec.apod, xp=self.xp, tile_betas=betas)

    def focus_per_depth(self, tile_focus, idx=None):
        """1-based focus pixel per focus plane: median over the plane's tiles (0-based bins + 1),
        of the tiles idx (default all)."""
        idx = np.arange(len(self.keys)) if idx is None else np.asarray(idx)
        z = np.array([self.keys[i][2] for i in idx])
        f = np.asarray(tile_focus)[idx]
        return [float(np.median(f[z == d])) + 1.0 for d in range(self.ss.depths)]

    # ------------------------------------------------------------------ run
    def run(self, validate: bool = True):
        """Stream, estimate and reconstruct the volume; returns measurements (and the volume)."""
        torch = self.torch
        ss = self.ss
        n = len(self.keys)
        stats = {"gen_s": 0.0, "h2d_s": 0.0, "net_s": 0.0, "fallback_s": 0.0}
        free_q, ready_q, landed_q = queue.Queue(), queue.Queue(), queue.Queue()
        for s in range(len(self.ring)):
            free_q.put(s)
        rows = {}
        for i, k in enumerate(self.keys):
            rows.setdefault(k[1], []).append(i)
        tile_beta = np.full(n, ss.beta_idx)
        tile_focus = np.full(n, np.nan)
        tile_bad = np.zeros(n, int)
        self.xp.synchronize()
        t0 = time.perf_counter()
        threads = [threading.Thread(target=self._guard, args=(self._produce, free_q, ready_q, stats), daemon=True),
                   threading.Thread(target=self._guard, args=(self._copy, ready_q, free_q, landed_q, stats),
                                    daemon=True)]
        for t in threads:
            t.start()
        landed = np.zeros(n, bool)
        done_rows = set()
        t_est = t_rows = 0.0
        rec = vol = fpd = ref = vref = None
        # with part of the rows resident the plain validation must run row by row, before the
        # row's slot is refilled
        inline_check = validate and self.resident_rows < ss.tiles_y
        try:
            while True:
                i = landed_q.get()
                if i is None:
                    break
                landed[i] = True
                yi = self.keys[i][1]
                if yi not in done_rows and landed[rows[yi]].all():
                    done_rows.add(yi)
                    te = time.perf_counter()
                    idx = rows[yi]
                    if self.est is not None:
                        b, f, nb = self._estimate(idx, stats)
                    else:
                        b, f, nb = self.truth["beta_idx"][idx], self.truth["focus"][idx], np.zeros(len(idx), int)
                    tile_beta[idx], tile_focus[idx], tile_bad[idx] = b, f, nb
                    t_est += time.perf_counter() - te
                    if self.rowwise:
                        # reconstruct this row now, while the next rows stream; the focus per
                        # depth comes from the first row (every row holds all depths)
                        tr = time.perf_counter()
                        for i in idx:
                            self.tile_map[self.keys[i]] = self.tile_raw(i)
                        betas = {self.keys[i]: float(tile_beta[i]) * self.to_ref for i in idx}
                        if rec is None:
                            fpd = self.focus_per_depth(tile_focus, idx)
                            rec = self.reconstructor(tile_beta, fpd, idx=idx)
                            vol = rec.allocate_volume()
                            if inline_check:
                                ref = self.reconstructor(tile_beta, fpd, form="plain", idx=idx)
                                vref = ref.allocate_volume()
                        else:
                            rec.set_tile_betas(betas, update=True)
                            if ref is not None:
                                ref.set_tile_betas(betas, update=True)
                        evs = [rec.volume_row(vol, yi)]
                        if ref is not None:
                            evs.append(ref.volume_row(vref, yi))
                        main_ev = torch.cuda.Event()          # the estimator's reads of the row
                        main_ev.record(torch.cuda.current_stream(self.xp.device))
                        self._release_row(yi, evs + [main_ev])
                        t_rows += time.perf_counter() - tr
        finally:
            free_q.put(None)
            for yi in range(ss.tiles_y):                  # a failed run never leaves the copier waiting
                self._release_of(yi)[0].set()
            for t in threads:
                t.join(timeout=60)
        if self._err is not None:
            raise RuntimeError("streaming thread failed") from self._err
        if not landed.all():
            raise RuntimeError(f"{int((~landed).sum())} tiles never landed")
        t_stream = time.perf_counter()
        if self.rowwise:
            vol = rec.finish_volume(vol)
        else:
            fpd = self.focus_per_depth(tile_focus)
            rec = self.reconstructor(tile_beta, fpd)
            vol = rec.reconstruct_volume(rec.allocate_volume())
        self.xp.synchronize()
        t1 = time.perf_counter()
        total = n * ss.n_y * ss.n_x
        raw_bytes = float(n * ss.n_y * self.lines * self.n * 2)
        out = {
            "ascans": total, "tiles": n, "raw_GB": raw_bytes / 1e9,
            "seconds": t1 - t0, "ascans_per_s": total / (t1 - t0),
            "stream_seconds": t_stream - t0, "after_stream_seconds": t1 - t_stream,
            "row_reconstruction": "as rows land" if self.rowwise else "after the stream",
            "row_enqueue_seconds": t_rows,
            "host_gen_ascans_per_s": total / stats["gen_s"] * 1.0 if stats["gen_s"] else None,
            "host_gen_threads": self.gen.threads,
            "h2d_GB_per_s": raw_bytes / stats["h2d_s"] / 1e9 if stats["h2d_s"] else None,
            "estimation_seconds": t_est, "network_seconds": stats["net_s"], "fallback_seconds": stats["fallback_s"],
            "fallback_rate_bscans": float(tile_bad.sum() / (n * self.est_bscans)) if self.est is not None else 0.0,
            "fallback_rate_tiles": float((tile_bad > 0).mean()) if self.est is not None else 0.0,
            "beta_err_idx_median": float(np.median(np.abs(tile_beta - self.truth["beta_idx"]))),
            "focus_err_bins_median": float(np.median(np.abs(tile_focus - self.truth["focus"]))),
            "focus_per_depth": fpd, "output_shape_yzx": list(vol["db"].shape), "form": self.form,
            "source": "files" if self.readers is not None else "host generator",
            "resident_rows": self.resident_rows, "raw_resident_GB": float(self.slots.numel() * 2) / 1e9,
            "validation": ("plain FP32 row by row beside the reconstruction (inside the timed run)" if inline_check
                           else "plain FP32 volume after the run" if validate else None),
        }
        if validate:
            if inline_check:
                vref = ref.finish_volume(vref)
            else:
                ref = self.reconstructor(tile_beta, fpd, form="plain")
                vref = ref.reconstruct_volume(ref.allocate_volume())
            frac, mx = lsb_agreement(self.xp, vol["db"], vref["db"], vref["clim"])
            out["within_1lsb_of_plain"] = frac
            out["max_lsb_vs_plain"] = mx
            out["plain_clim_dB"] = [float(v) for v in vref["clim"].cpu().numpy()]
            del vref, ref
        out["volume"] = vol
        return out


def lsb_agreement(xp, db, db_ref, clim_ref, planes: int = 16):
    """Fraction of voxels whose uint16 codes (quantised with the reference volume's clim) are
    within 1 LSB of the reference volume's, and the largest difference; on the device."""
    torch = xp.torch
    fin = DeviceFinalizer(xp)
    ok = 0
    mx = 0
    for p0 in range(0, db.shape[0], planes):
        a = fin.quantize(db[p0:p0 + planes], clim_ref).view(torch.int16).to(torch.int32) & 0xFFFF
        b = fin.quantize(db_ref[p0:p0 + planes], clim_ref).view(torch.int16).to(torch.int32) & 0xFFFF
        d = (a - b).abs_()
        ok += int((d <= 1).sum())
        mx = max(mx, int(d.max()))
    return ok / db.numel(), mx


def main(argv=None):
    import argparse

    from .backend import get_xp
    ap = argparse.ArgumentParser(description="streamed tiled volume: host generation, pinned async copies, "
                                             "per-tile estimation, MFMA reconstruction, stitching")
    ap.add_argument("--tiles-x", type=int, default=12)
    ap.add_argument("--tiles-y", type=int, default=9)
    ap.add_argument("--depths", type=int, default=8)
    ap.add_argument("--bscans", type=int, default=40)
    ap.add_argument("--ascans", type=int, default=1000)
    ap.add_argument("--ring", type=int, default=8)
    ap.add_argument("--gen-threads", type=int, default=None)
    ap.add_argument("--estimator", default="learned", choices=["learned", "given"])
    ap.add_argument("--form", default="mfma", choices=["mfma", "plain"])
    ap.add_argument("--no-validate", action="store_true")
    ap.add_argument("--after-stream", action="store_true",
                    help="reconstruct once every tile has landed (focus per depth from all rows)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--resident-rows", type=int, default=None, help="y-tile rows of raw data held in HBM")
    ap.add_argument("--tiff", default=None, help="also write the volume as the legacy BigTIFF (+ .json)")
    a = ap.parse_args(argv)
    xp = get_xp("hip")[0]
    ss = StreamSpec(tiles_x=a.tiles_x, tiles_y=a.tiles_y, depths=a.depths, n_y=a.bscans, n_x=a.ascans, seed=a.seed)
    sv = StreamedVolume(xp, ss, ring=a.ring, gen_threads=a.gen_threads, estimator=a.estimator, form=a.form,
                        log=lambda *m: print(*m, flush=True), rowwise=not a.after_stream,
                        resident_rows=a.resident_rows)
    print(f"streamed volume: {ss.n_tiles} tiles, {sv.slots.numel() * 2 / 1e9:.1f} GB raw resident "
          f"({sv.resident_rows} of {ss.tiles_y} rows)", flush=True)
    sv.warmup()
    r = sv.run(validate=not a.no_validate)
    vol = r.pop("volume")
    if a.tiff:
        from .tiffio import write_legacy_tiff
        t0 = time.perf_counter()
        size = write_legacy_tiff(a.tiff, xp, vol["bits"], vol["clim"], {"spec": asdict(ss)})
        r["tiff_seconds"] = time.perf_counter() - t0
        r["tiff_bytes"] = size
    r["spec"] = asdict(ss)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
