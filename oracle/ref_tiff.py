"""Oracle (test infrastructure only): PackBits and the BigTIFF container restated from the
published specifications (TIFF 6.0 section 9 PackBits; the BigTIFF extension: 'II', version 43,
8-byte offsets and counts, 20-byte IFD entries), independent of the product's writer and of the
reference's tifffile call (octrecon/io/tiff_writer.py:53-71).  Not used by the product."""
import struct

import numpy as np


def packbits_decode(data: bytes, n_out: int) -> bytes:
    out = bytearray()
    i = 0
    while len(out) < n_out and i < len(data):
        h = data[i] - 256 if data[i] > 127 else data[i]
        i += 1
        if 0 <= h <= 127:
            out += data[i:i + h + 1]
            i += h + 1
        elif -127 <= h <= -1:
            out += bytes([data[i]]) * (1 - h)
            i += 1
        # h == -128: no operation
    return bytes(out[:n_out])


def packbits_encode(data: bytes) -> bytes:
    """Greedy: runs of 3+ equal bytes become replicate runs, the rest literal runs (<= 128)."""
    out = bytearray()
    i, n = 0, len(data)
    lit = bytearray()

    def flush():
        for k in range(0, len(lit), 128):
            part = lit[k:k + 128]
            out.append(len(part) - 1)
            out.extend(part)
        lit.clear()
    while i < n:
        j = i
        while j < n and j - i < 128 and data[j] == data[i]:
            j += 1
        if j - i >= 3:
            flush()
            out.append(256 - (j - i - 1))
            out.append(data[i])
            i = j
        else:
            lit.append(data[i])
            i += 1
    flush()
    return bytes(out)


def read_bigtiff(path):
    """-> list of (tags {tag: values}, pixels (rows, cols) uint16) for a PackBits/uncompressed
    little-endian BigTIFF of 16-bit single-sample pages."""
    b = open(path, "rb").read()
    assert b[:2] == b"II" and struct.unpack_from("<HHH", b, 2) == (43, 8, 0)
    pos = struct.unpack_from("<Q", b, 8)[0]
    size = {1: 1, 2: 1, 3: 2, 4: 4, 16: 8}
    fmt = {1: "B", 2: "B", 3: "H", 4: "I", 16: "Q"}
    pages = []
    while pos:
        n = struct.unpack_from("<Q", b, pos)[0]
        tags = {}
        for e in range(n):
            tag, typ, cnt = struct.unpack_from("<HHQ", b, pos + 8 + 20 * e)
            nbytes = size[typ] * cnt
            at = pos + 8 + 20 * e + 12
            if nbytes > 8:
                at = struct.unpack_from("<Q", b, at)[0]
            vals = struct.unpack_from("<%d%s" % (cnt, fmt[typ]), b, at)
            tags[tag] = bytes(vals) if typ == 2 else list(vals)
        pos = struct.unpack_from("<Q", b, pos + 8 + 20 * n)[0]
        w, h = tags[256][0], tags[257][0]
        rps = tags.get(278, [h])[0]
        raw = bytearray()
        for off, cnt in zip(tags[273], tags[279]):
            seg = b[off:off + cnt]
            nrow = min(rps, h - len(raw) // (2 * w))
            raw += packbits_decode(seg, 2 * w * nrow) if tags[259][0] == 32773 else seg
        pages.append((tags, np.frombuffer(bytes(raw), "<u2").reshape(h, w)))
    return pages
