"""The legacy BigTIFF output from the device (oct_recon_af_v1_amd.tiffio): device PackBits rows
decode (oracle decoder, TIFF 6.0) to the codes, and a reconstructed volume written as BigTIFF
reads back (Pillow and the oracle reader) as the device's codes with the metadata JSON."""
import json

import numpy as np
import pytest

from oracle import ref_tiff as RT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def xp():
    from oct_recon_af_v1_amd.backend import get_xp
    return get_xp("hip")[0]


@pytest.mark.parametrize("cols", [1, 63, 64, 65, 1000, 10900])
def test_device_packbits_rows(xp, cols):
    from oct_recon_af_v1_amd.tiffio import packbits_rows
    print(f"case: packbits rows of {cols} codes", flush=True)
    torch = xp.torch
    rng = np.random.default_rng(cols)
    codes = rng.integers(0, 65536, size=(6, cols), dtype=np.uint16)
    codes[1] = 0                                        # uniform rows
    codes[2] = 0x0101
    codes[3, : cols // 2] = 7                           # uniform values whose bytes differ (0x0007)
    codes[4, :: 3] = 0
    dev = torch.from_numpy(codes.view(np.int16)).to(xp.device).view(torch.uint16)
    enc, row_bytes = packbits_rows(xp, dev)
    host = enc.cpu().numpy().tobytes()
    assert int(row_bytes.sum()) == len(host)
    pos = 0
    for r in range(codes.shape[0]):
        n = int(row_bytes[r])
        assert RT.packbits_decode(host[pos:pos + n], 2 * cols) == codes[r].astype("<u2").tobytes(), r
        pos += n
    if cols >= 64:
        assert row_bytes[1] == 2 * ((2 * cols + 127) // 128)     # zero rows: one 2-byte run per chunk


def test_volume_bigtiff(xp, tmp_path):
    pytest.importorskip("PIL")
    from PIL import Image

    from oct_recon_af_v1_amd import synthetic as SY
    from oct_recon_af_v1_amd.geometry import ScanGrid, beta_from_index_units, build_spectral_geometry
    from oct_recon_af_v1_amd.pipeline import ReconConfig, Reconstructor
    from oct_recon_af_v1_amd.tiffio import write_legacy_tiff
    print("case: small volume to BigTIFF", flush=True)
    spec = SY.SpectrumSpec()
    lam = spec.lambda_nm()
    bref = beta_from_index_units(4e-5, lam)
    dz = float(np.diff(build_spectral_geometry(lam, bref, 1.33).z_um[:2])[0])
    ny, nx = 5, 128
    grid = ScanGrid(np.array([-0.058, 0.058]), np.array([0.0]), np.array([0.0, 0.05]), nx * 1e-3, ny * 1e-3, nx,
                    ny, pixel_size_um=1.0)
    tiles = {(xi, 0, zi): SY.make_bscans(ny, nx, spec, beta=4e-5, seed=90 + 2 * xi + zi, device=xp.device)
             for xi in range(2) for zi in range(2)}
    cfg = ReconConfig(dispersion_quadratic_term=bref, focus_positions=512.0, focus_sigma=30.0 / dz,
                      crop_z_range_mm=[0.0, 0.1], output_pixel_size_um=1.0, n_medium=1.33, batch_frames=ny)
    vol = Reconstructor(cfg, grid, lam, tiles, spec.apod, xp=xp).run()["volume"]
    meta = {"x": {"order": 1, "units": "mm"}, "note": "synthetic"}
    path = tmp_path / "vol.tiff"
    write_legacy_tiff(path, xp, vol["bits"], vol["clim"], meta, planes_per_batch=2)
    bits = vol["bits"].view(xp.torch.int16).cpu().numpy().view(np.uint16)
    pages = RT.read_bigtiff(path)
    assert len(pages) == bits.shape[0]
    for (tags, px), b in zip(pages, bits):
        assert np.array_equal(px, b)
    doc = json.loads(pages[0][0][305][:-1].decode())
    assert doc["metadata"] == meta and doc["version"] == 3
    assert doc["clim"] == [float(v) for v in vol["clim"].cpu().numpy()]
    assert json.loads((tmp_path / "vol.tiff.json").read_text()) == doc
    im = Image.open(path)
    im.seek(bits.shape[0] - 1)
    assert np.array_equal(np.array(im).astype(np.uint16), bits[-1])
