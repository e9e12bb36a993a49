"""The legacy BigTIFF output (octrecon/io/tiff_writer.py:53-71, yOCT2Tif.m): one page per output
y plane holding the (nZ rows, nX columns) uint16 codes, PackBits compression, TIFF tag 305
"Software" = JSON {"metadata", "clim", "version": 3} on the first page only, and the same JSON
beside the file as <path>.json.

The codes come quantised from the device (DeviceFinalizer.quantize); each row is PackBits-encoded
on the device (oct_packbits_sizes / oct_packbits_write, one strip per row) in batches of planes,
copied to a pinned host buffer and written with the BigTIFF structures built here (TIFF 6.0 with
the BigTIFF extension: 8-byte offsets, one IFD per page).
"""
from __future__ import annotations

import json
import struct
from pathlib import Path

import numpy as np

from . import _lib

# TIFF tag numbers and field types
_WIDTH, _LENGTH, _BITS, _COMPRESSION, _PHOTOMETRIC = 256, 257, 258, 259, 262
_STRIP_OFFSETS, _SAMPLES, _ROWS_PER_STRIP, _STRIP_COUNTS, _SOFTWARE, _SAMPLE_FORMAT = 273, 277, 278, 279, 305, 339
_ASCII, _SHORT, _LONG, _LONG8 = 2, 3, 4, 16
_PACKBITS = 32773


class BigTiffWriter:
    """Pages of PackBits strips into a little-endian BigTIFF file."""

    def __init__(self, path):
        self.path = Path(path)
        self.f = open(self.path, "wb")
        self.f.write(b"II" + struct.pack("<HHHQ", 43, 8, 0, 0))
        self._link = 8                       # where the next IFD offset is stored (header first)

    def _entry(self, tag, typ, count, value):
        if typ in (_SHORT,):
            raw = struct.pack("<H", value).ljust(8, b"\0")
        elif typ == _LONG:
            raw = struct.pack("<I", value).ljust(8, b"\0")
        else:                                # LONG8 inline value or any offset
            raw = struct.pack("<Q", value)
        return struct.pack("<HHQ", tag, typ, count) + raw

    def write_page(self, width: int, height: int, strips: memoryview, strip_counts, software: bytes | None = None):
        """strips: the page's encoded rows back to back; strip_counts: bytes of each row."""
        f = self.f
        counts = np.asarray(strip_counts, np.uint64)
        if counts.size != height or int(counts.sum()) != len(strips):
            raise ValueError("strip byte counts do not add up to the page data")
        data_pos = f.tell()
        f.write(strips)
        offsets = data_pos + np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.uint64)
        if f.tell() % 2:
            f.write(b"\0")                   # word-aligned tables
        if height > 1:
            off_pos = f.tell()
            f.write(offsets.astype("<u8").tobytes())
            cnt_pos = f.tell()
            f.write(counts.astype("<u8").tobytes())
        sw_pos = None
        if software is not None:
            sw = software + b"\0"
            if len(sw) > 8:
                sw_pos = f.tell()
                f.write(sw)
                if f.tell() % 2:
                    f.write(b"\0")
        entries = [
            self._entry(_WIDTH, _LONG, 1, width),
            self._entry(_LENGTH, _LONG, 1, height),
            self._entry(_BITS, _SHORT, 1, 16),
            self._entry(_COMPRESSION, _SHORT, 1, _PACKBITS),
            self._entry(_PHOTOMETRIC, _SHORT, 1, 1),             # min-is-black
            self._entry(_STRIP_OFFSETS, _LONG8, height, int(offsets[0]) if height == 1 else off_pos),
            self._entry(_SAMPLES, _SHORT, 1, 1),
            self._entry(_ROWS_PER_STRIP, _LONG, 1, 1),
            self._entry(_STRIP_COUNTS, _LONG8, height, int(counts[0]) if height == 1 else cnt_pos),
        ]
        if software is not None:
            sw = software + b"\0"
            if sw_pos is None:
                entries.append(struct.pack("<HHQ", _SOFTWARE, _ASCII, len(sw)) + sw.ljust(8, b"\0"))
            else:
                entries.append(self._entry(_SOFTWARE, _ASCII, len(sw), sw_pos))
        entries.append(self._entry(_SAMPLE_FORMAT, _SHORT, 1, 1))   # unsigned integer
        ifd_pos = f.tell()
        f.write(struct.pack("<Q", len(entries)) + b"".join(entries) + struct.pack("<Q", 0))
        end = f.tell()
        f.seek(self._link)
        f.write(struct.pack("<Q", ifd_pos))
        f.seek(end)
        self._link = ifd_pos + 8 + 20 * len(entries)

    def close(self):
        self.f.close()


def packbits_rows(xp, codes):
    """Device PackBits of every row of codes (..., row_elems) uint16 -> (device bytes, per-row
    byte counts as host int64)."""
    torch = xp.torch
    L = _lib.lib()
    row_elems = codes.shape[-1]
    rows = int(codes.numel() // row_elems)
    nch = L.oct_packbits_chunks(row_elems)
    sizes = torch.empty((rows, nch), dtype=torch.int32, device=xp.device)
    _lib.check(L.oct_packbits_sizes(codes.data_ptr(), rows, row_elems, sizes.data_ptr(), xp.stream()),
               "oct_packbits_sizes")
    ends = torch.cumsum(sizes.view(-1).to(torch.int64), 0)
    offsets = ends - sizes.view(-1).to(torch.int64)
    total = int(ends[-1].item()) if rows else 0
    out = torch.empty(max(total, 1), dtype=torch.uint8, device=xp.device)
    _lib.check(L.oct_packbits_write(codes.data_ptr(), rows, row_elems, offsets.data_ptr(), out.data_ptr(), xp.stream()),
               "oct_packbits_write")
    row_bytes = sizes.to(torch.int64).sum(1).cpu().numpy()
    return out[:total], row_bytes


def write_legacy_tiff(path, xp, bits, clim, metadata: dict, planes_per_batch: int = 16):
    """bits: device (nY, nZ, nX) uint16 codes (Reconstructor volume["bits"]); clim: the two dB
    limits they were quantised with.  Writes <path> and <path>.json; returns the file size."""
    torch = xp.torch
    path = Path(path)
    c = [float(v) for v in (clim.cpu().numpy() if hasattr(clim, "cpu") else np.asarray(clim, float))]
    meta = {"metadata": metadata, "clim": c, "version": 3}
    meta_json = json.dumps(meta, separators=(",", ":")).encode()
    if bits.dim() != 3 or bits.dtype != torch.uint16 or not bits.is_contiguous():
        raise ValueError("bits must be a contiguous (nY, nZ, nX) uint16 device tensor")
    ny, nz, nx = bits.shape
    w = BigTiffWriter(path)
    host = None
    try:
        for p0 in range(0, ny, planes_per_batch):
            blk = bits[p0:p0 + planes_per_batch]
            enc, row_bytes = packbits_rows(xp, blk)
            if host is None or host.numel() < enc.numel():
                host = torch.empty(max(enc.numel(), 1), dtype=torch.uint8, pin_memory=True)
            h = host[:enc.numel()]
            h.copy_(enc)
            mv = memoryview(h.numpy())
            rb = row_bytes.reshape(blk.shape[0], nz)
            pos = 0
            for k in range(blk.shape[0]):
                n = int(rb[k].sum())
                w.write_page(nx, nz, mv[pos:pos + n], rb[k], software=meta_json if p0 + k == 0 else None)
                pos += n
    finally:
        w.close()
    with open(path.with_suffix(path.suffix + ".json"), "w") as f:
        json.dump(meta, f)
    return path.stat().st_size
