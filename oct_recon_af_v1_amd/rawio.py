"""Raw B-scan files of a tile, read into caller memory (pinned host buffers) and ingested on the
device.

Layouts (the reference's readers):
  * Thorlabs OCITY, unzipped (octrecon/io/volume.py:197-215): <tile>/data/Spectral{i}.data,
    little-endian int16 (interf_size, n_lambda) with the background A-lines first.  Read with
    readinto straight into the destination buffer, no intermediate copy.
  * Thorlabs SRR (octrecon/io/srr.py): <tile>/Data_Y%04d_YTotal%d_B%04d_BTotal%d_<System>.srr,
    a text header ('headersize=<n>' line, then n bytes of key=value tokens) followed by uint16
    samples (file_lines, n_lambda) of which the low 12 bits are data.  The payload is read as it
    is and unpacked on the device (oct_srr_unpack: rem 4096, background region, then the scan
    region), so the host only moves bytes.
Header XML parsing, zipped .oct archives and the other vendors' formats stay with the caller:
the readers take the sizes the caller's header gives.
"""
from __future__ import annotations

import os
import re
import threading
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

from . import _lib

_SRR_NAME = re.compile(r"^Data_Y(\d{1,4})_YTotal(\d+)_B(\d{1,4})_BTotal(\d+)_(.+)\.srr$")


def _readinto(path: Path, dst: memoryview, nbytes: int, offset: int = 0, whole: bool = False):
    """nbytes of the file from `offset` into dst; IOError if the file holds fewer.  whole=True:
    the file must hold exactly nbytes after the offset, counted in 2-byte samples (an SRR payload,
    octrecon/io/srr.py read_file), so a longer or truncated file is an error too."""
    with open(path, "rb", buffering=0) as f:
        if whole:
            samples = (os.fstat(f.fileno()).st_size - offset) // 2
            if samples != nbytes // 2:
                raise IOError(f"{path}: expected {nbytes // 2} samples, got {samples}")
        if offset:
            f.seek(offset)
        got = f.readinto(dst)
    if got != nbytes:
        raise IOError(f"{path}: expected {nbytes} bytes, got {got}")


def _host_view(out) -> memoryview:
    """Writable byte view of a contiguous numpy array or CPU torch tensor (pinned or not)."""
    if hasattr(out, "data_ptr"):
        if out.device.type != "cpu" or not out.is_contiguous():
            raise ValueError("destination must be a contiguous host tensor")
        arr = out.numpy()
    else:
        arr = out
        if not arr.flags.c_contiguous:
            raise ValueError("destination must be contiguous")
    return memoryview(arr.reshape(-1).view(np.uint8))


class OcityTile:
    """An unzipped Thorlabs OCITY tile folder (data/Spectral{i}.data)."""
    kind = "ocity"

    def __init__(self, folder, interf_size: int, n_lambda: int):
        self.dir = Path(folder)
        self.interf_size, self.n_lambda = int(interf_size), int(n_lambda)
        self.file_bytes = 2 * self.interf_size * self.n_lambda

    def path(self, file_index: int) -> Path:
        return self.dir / "data" / f"Spectral{int(file_index)}.data"

    def read(self, file_indices, out, pool: ThreadPoolExecutor | None = None):
        """Files -> out (len(file_indices), interf_size, n_lambda) int16 / uint16, host memory."""
        mv = _host_view(out)
        n = len(file_indices)
        if mv.nbytes != n * self.file_bytes:
            raise ValueError(f"destination holds {mv.nbytes} bytes, {n} files need {n * self.file_bytes}")
        jobs = [(self.path(fi), mv[k * self.file_bytes:(k + 1) * self.file_bytes], self.file_bytes)
                for k, fi in enumerate(file_indices)]
        if pool is None:
            for j in jobs:
                _readinto(*j)
        else:
            list(pool.map(lambda j: _readinto(*j), jobs))
        return out


def parse_srr_header(b: bytes) -> dict:
    """The SRR text header: 'headersize=<n>' on the first line, then n bytes of whitespace
    separated key=value tokens; scanregions / aporegions = <count>,<start>,<end> (0-based start,
    exclusive end, in A-lines).  headerTotalBytes is where the samples start
    (octrecon/io/srr.py:29-53)."""
    nl = b.index(b"\n")
    m = re.match(r"headersize=(\d+)", b[:nl].decode("ascii", errors="replace").strip())
    if not m:
        raise ValueError("SRR file does not start with 'headersize=<n>'")
    total = nl + 1 + int(m.group(1))
    h = {"headerTotalBytes": total}
    for tok in b[nl + 1:total].decode("ascii", errors="replace").split():
        key, sep, val = tok.partition("=")
        if not sep:
            continue
        if key in ("scanregions", "aporegions"):
            parts = val.split(",")
            lo, hi = int(float(parts[1])), int(float(parts[2]))
            if key == "scanregions":
                h["scanstart"], h["scanend"] = lo, hi
            else:
                h["apodstart"], h["apodend"] = lo, hi
        else:
            try:
                h[key] = float(val) if any(c in val for c in ".eE") else int(val)
            except ValueError:
                h[key] = val.strip("'\"")
    return h


def srr_file_name(y1: int, y_total: int, b1: int, b_total: int, system: str) -> str:
    """SRR file name of 1-based frame y1 and repeat b1 (octrecon/io/srr.py:56-59)."""
    sysname = system.replace("_SRR", "").replace("_srr", "")
    return f"Data_Y{y1:04d}_YTotal{y_total}_B{b1:04d}_BTotal{b_total}_{sysname}.srr"


class SrrTile:
    """A Thorlabs SRR tile folder; the header of the first file (sorted) describes all files,
    as the reference reads it (octrecon/io/srr.py:62-93)."""
    kind = "srr"

    def __init__(self, folder):
        self.dir = Path(folder)
        names = sorted(p.name for p in self.dir.glob("*.srr"))
        if not names:
            raise FileNotFoundError(f"{self.dir}: no .srr files")
        m = _SRR_NAME.match(names[0])
        if not m:
            raise ValueError(f"{names[0]}: not an SRR file name (Data_Y%04d_YTotal%d_B%04d_BTotal%d_<System>.srr)")
        self.y_total, self.b_total, self.system = int(m.group(2)), int(m.group(4)), m.group(5)
        with open(self.dir / names[0], "rb") as f:
            self.header = h = parse_srr_header(f.read(1 << 16))
        self.n_lambda = int(h["size1"])
        self.apod = (int(h["apodstart"]), int(h["apodend"]))
        self.scan = (int(h["scanstart"]), int(h["scanend"]))
        if self.apod[1] <= self.apod[0]:
            raise ValueError(f"{self.dir}: SRR file has no background region")
        self.file_lines = self.scan[1]                 # the payload holds exactly size1 * scanend samples
        if self.apod[1] > self.file_lines:
            raise ValueError(f"{self.dir}: background region beyond the samples")
        self.interf_size = (self.apod[1] - self.apod[0]) + (self.scan[1] - self.scan[0])
        self.payload_bytes = 2 * self.file_lines * self.n_lambda
        self.offset = int(h["headerTotalBytes"])

    def path(self, file_index: int) -> Path:
        y0, b0 = divmod(int(file_index), self.b_total)
        return self.dir / srr_file_name(y0 + 1, self.y_total, b0 + 1, self.b_total, self.system)

    def read(self, file_indices, out, pool: ThreadPoolExecutor | None = None):
        """Payloads -> out (len(file_indices), file_lines, n_lambda) uint16, host memory (as
        stored; oct_srr_unpack turns them into the (interf_size, n_lambda) int16 layout)."""
        mv = _host_view(out)
        n = len(file_indices)
        if mv.nbytes != n * self.payload_bytes:
            raise ValueError(f"destination holds {mv.nbytes} bytes, {n} files need {n * self.payload_bytes}")
        jobs = [(self.path(fi), mv[k * self.payload_bytes:(k + 1) * self.payload_bytes], self.payload_bytes,
                 self.offset, True) for k, fi in enumerate(file_indices)]
        if pool is None:
            for j in jobs:
                _readinto(*j)
        else:
            list(pool.map(lambda j: _readinto(*j), jobs))
        return out


class FileTileSource:
    """Tile source for Reconstructor (``tiles=source``): (xi, yi, zi, file_indices) -> device
    int16 tensor (files, interf_size, n_lambda).  Files are read by io_threads threads into a
    pinned staging buffer and copied asynchronously on the caller's stream; SRR payloads are
    unpacked on the device.  A staging buffer is reused only after its copy has completed."""

    def __init__(self, tiles: dict, xp, io_threads: int = 8, stage_depth: int = 2):
        self.tiles = tiles
        self.xp, self.torch = xp, xp.torch
        self.pool = ThreadPoolExecutor(max(1, int(io_threads)))
        self._stages = [None] * max(1, int(stage_depth))
        self._events = [None] * len(self._stages)
        self._next = 0
        self._lock = threading.Lock()

    def _stage(self, shape, dtype):
        with self._lock:
            k = self._next
            self._next = (k + 1) % len(self._stages)
        ev = self._events[k]
        if ev is not None:
            ev.synchronize()                           # its previous copy has finished
        buf = self._stages[k]
        n = int(np.prod(shape))
        if buf is None or buf.numel() < n or buf.dtype != dtype:
            buf = self.torch.empty(n, dtype=dtype, pin_memory=True)
            self._stages[k] = buf
        return k, buf[:n].view(shape)

    def __call__(self, xi, yi, zi, file_indices):
        torch = self.torch
        t = self.tiles[(xi, yi, zi)]
        files = [int(f) for f in np.asarray(file_indices).ravel()]
        n = len(files)
        if t.kind == "ocity":
            k, host = self._stage((n, t.interf_size, t.n_lambda), torch.int16)
            t.read(files, host, self.pool)
            dev = torch.empty(host.shape, dtype=torch.int16, device=self.xp.device)
            dev.copy_(host, non_blocking=True)
        else:
            k, host = self._stage((n, t.file_lines, t.n_lambda), torch.int16)   # uint16 bits as int16
            t.read(files, host, self.pool)
            payload = torch.empty(host.shape, dtype=torch.int16, device=self.xp.device)
            payload.copy_(host, non_blocking=True)
            dev = torch.empty((n, t.interf_size, t.n_lambda), dtype=torch.int16, device=self.xp.device)
            _lib.check(_lib.lib().oct_srr_unpack(payload.data_ptr(), n, t.file_lines, t.n_lambda, t.apod[0],
                                                 t.apod[1], t.scan[0], t.scan[1], dev.data_ptr(),
                                                 self.xp.stream()), "oct_srr_unpack")
        ev = torch.cuda.Event()
        ev.record()
        self._events[k] = ev
        return dev

    def close(self):
        self.pool.shutdown(wait=True)
