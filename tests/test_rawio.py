"""Raw-file readers (oct_recon_af_v1_amd.rawio) against files written in the reference's
layouts and against the oracle's numpy restatement of the reference's readers.  CPU only."""
import numpy as np
import pytest

import rawfiles
from oracle import ref_io as RI
from oct_recon_af_v1_amd import rawio


def test_srr_file_name_format():
    # the format of the reference's SRR names (octrecon/io/srr.py:4-5, 56-59)
    assert rawio.srr_file_name(1, 4, 2, 3, "Ganymede_SRR") == "Data_Y0001_YTotal4_B0002_BTotal3_Ganymede.srr"
    assert rawio.srr_file_name(12, 100, 1, 1, "Telesto") == "Data_Y0012_YTotal100_B0001_BTotal1_Telesto.srr"


def test_ocity_read(tmp_path):
    rng = np.random.default_rng(1)
    raw = rng.integers(-2000, 4096, size=(5, 30, 64)).astype(np.int16)
    rawfiles.write_ocity(tmp_path, raw)
    t = rawio.OcityTile(tmp_path, 30, 64)
    out = np.empty((3, 30, 64), np.int16)
    t.read([4, 0, 2], out)
    assert np.array_equal(out, raw[[4, 0, 2]])
    assert np.array_equal(out, RI.read_ocity(tmp_path, [4, 0, 2], 30, 64))
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(3) as pool:
        out2 = np.empty_like(out)
        t.read([4, 0, 2], out2, pool)
    assert np.array_equal(out, out2)
    with pytest.raises(IOError):
        rawio.OcityTile(tmp_path, 31, 64).read([0], np.empty((1, 31, 64), np.int16))
    with pytest.raises(ValueError):
        t.read([0, 1], np.empty((1, 30, 64), np.int16))


@pytest.mark.parametrize("apod,scan", [((0, 5), (5, 37)), ((2, 7), (10, 42))])
def test_srr_read_and_oracle(tmp_path, apod, scan):
    rng = np.random.default_rng(2)
    file_lines = scan[1]                      # the payload holds size1 * scanend samples
    payload = rng.integers(0, 65536, size=(6, file_lines, 64)).astype(np.uint16)   # high bits set
    rawfiles.write_srr(tmp_path, payload, y_total=3, b_total=2, apod=apod, scan=scan)
    t = rawio.SrrTile(tmp_path)
    assert (t.y_total, t.b_total, t.system) == (3, 2, "Ganymede")
    assert t.n_lambda == 64 and t.apod == apod and t.scan == scan and t.file_lines == file_lines
    assert t.interf_size == (apod[1] - apod[0]) + (scan[1] - scan[0])
    out = np.empty((2, file_lines, 64), np.uint16)
    t.read([5, 1], out)
    assert np.array_equal(out, payload[[5, 1]])
    # the oracle's restatement: low 12 bits, background region first
    names = [t.path(i).name for i in (5, 1)]
    o = RI.read_srr(tmp_path, names)
    want = np.concatenate([payload[[5, 1], apod[0]:apod[1]], payload[[5, 1], scan[0]:scan[1]]], 1) & 0x0FFF
    assert np.array_equal(o, want.astype(np.int16))
    # a payload one sample longer or shorter than size1 * scanend is an error, as in the
    # reference's read_file (an odd trailing byte is not a sample)
    p5 = t.path(5)
    good = p5.read_bytes()
    for bad in (good + b"\x00\x00", good[:-2], good + b"\x00"):
        p5.write_bytes(bad)
        if len(bad) == len(good) + 1:
            t.read([5], np.empty((1, file_lines, 64), np.uint16))
        else:
            with pytest.raises(IOError):
                t.read([5], np.empty((1, file_lines, 64), np.uint16))
    p5.write_bytes(good)


def test_srr_header_tokens():
    body = b"size1=2048 scanregions=1,25,525 aporegions=1,0,25 a=1.5\n"
    first = b"headersize=%d\n" % len(body)
    h = rawio.parse_srr_header(first + body + b"PAYLOAD")
    assert h["size1"] == 2048 and (h["scanstart"], h["scanend"]) == (25, 525) and (h["apodstart"], h["apodend"]) == (0, 25)
    assert h["a"] == 1.5 and h["headerTotalBytes"] == len(first) + len(body)
