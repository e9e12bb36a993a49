"""The BigTIFF container of the legacy output (oct_recon_af_v1_amd.tiffio.BigTiffWriter): pages
of PackBits strips (from the oracle's encoder) read back by Pillow and by the oracle's reader;
PackBits pinned by the published example.  CPU only."""
import json

import numpy as np
import pytest

from oracle import ref_tiff as RT
from oct_recon_af_v1_amd.tiffio import BigTiffWriter


def test_packbits_published_example():
    # the example of the PackBits technical note (Apple TN1023), also in the TIFF 6.0 spec
    packed = bytes.fromhex("FEAA0280002AFDAA0380002A22F7AA")
    unpacked = bytes.fromhex("AAAAAA80002AAAAAAAAA80002A22" + "AA" * 10)
    assert RT.packbits_decode(packed, len(unpacked)) == unpacked
    rng = np.random.default_rng(0)
    for data in (unpacked, bytes(300), rng.integers(0, 3, 1000).astype(np.uint8).tobytes(), b"\x01"):
        assert RT.packbits_decode(RT.packbits_encode(data), len(data)) == data


def _write(path, pages, software):
    w = BigTiffWriter(path)
    for k, p in enumerate(pages):
        rows = [RT.packbits_encode(r.astype("<u2").tobytes()) for r in p]
        w.write_page(p.shape[1], p.shape[0], memoryview(b"".join(rows)), [len(r) for r in rows],
                     software=software if k == 0 else None)
    w.close()


def test_writer_pages_read_by_pillow_and_oracle(tmp_path):
    pytest.importorskip("PIL")
    from PIL import Image
    rng = np.random.default_rng(1)
    pages = [rng.integers(0, 65536, size=(7, 33), dtype=np.uint16) for _ in range(3)]
    pages[1][:] = 0                                   # a NaN plane
    pages[2][2:5] = 0x0101
    meta = json.dumps({"metadata": {"x": [1, 2]}, "clim": [-3.5, 61.25], "version": 3},
                      separators=(",", ":")).encode()
    path = tmp_path / "v.tiff"
    _write(path, pages, meta)
    got = RT.read_bigtiff(path)
    assert len(got) == 3
    for (tags, px), p in zip(got, pages):
        assert np.array_equal(px, p)
        assert tags[259] == [32773] and tags[258] == [16] and tags[262] == [1]
    assert got[0][0][305] == meta + b"\0" and 305 not in got[1][0]
    im = Image.open(path)
    for k, p in enumerate(pages):
        im.seek(k)
        assert np.array_equal(np.array(im).astype(np.uint16), p), k
    assert im.n_frames == 3


def test_single_row_and_short_software(tmp_path):
    p = np.arange(5, dtype=np.uint16)[None]
    path = tmp_path / "one.tiff"
    _write(path, [p], b"{}")
    (tags, px), = RT.read_bigtiff(path)
    assert np.array_equal(px, p) and tags[305] == b"{}\0"
