"""Raw files straight to the device (oct_recon_af_v1_amd.rawio): the SRR unpack kernel against the
oracle's restatement of the reference's reader (bit-exact), and tiled reconstructions from tile
folders equal to those from the same data already on the device (bit for bit)."""
import numpy as np
import pytest

import rawfiles
from oracle import ref_io as RI

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def xp():
    from oct_recon_af_v1_amd.backend import get_xp
    return get_xp("hip")[0]


@pytest.mark.parametrize("apod,scan", [((0, 25), (25, 281)), ((3, 28), (40, 296))])
def test_srr_unpack_bit_exact(xp, tmp_path, apod, scan):
    from oct_recon_af_v1_amd import _lib, rawio
    print(f"case: srr unpack {apod} {scan}", flush=True)
    torch = xp.torch
    rng = np.random.default_rng(apod[0])
    payload = rng.integers(0, 65536, size=(4, scan[1], 2048)).astype(np.uint16)
    rawfiles.write_srr(tmp_path, payload, y_total=2, b_total=2, apod=apod, scan=scan)
    t = rawio.SrrTile(tmp_path)
    host = np.empty((4, t.file_lines, 2048), np.uint16)
    t.read(range(4), host)
    dev_in = torch.from_numpy(host.view(np.int16)).to(xp.device)
    out = torch.empty((4, t.interf_size, 2048), dtype=torch.int16, device=xp.device)
    _lib.check(_lib.lib().oct_srr_unpack(dev_in.data_ptr(), 4, t.file_lines, 2048, apod[0], apod[1], scan[0],
                                         scan[1], out.data_ptr(), xp.stream()), "oct_srr_unpack")
    want = RI.read_srr(tmp_path, [t.path(i).name for i in range(4)])
    assert np.array_equal(out.cpu().numpy(), want)


def _grid(ny, nx):
    from oct_recon_af_v1_amd.geometry import ScanGrid
    xc = (np.arange(2) - 0.5) * nx * 1e-3 * 0.9
    return ScanGrid(xc, np.array([0.0]), np.array([0.0, 0.05]), nx * 1e-3, ny * 1e-3, nx, ny, pixel_size_um=1.0)


@pytest.mark.parametrize("kind", ["ocity", "srr"])
def test_file_source_equals_device_tiles(xp, tmp_path, kind):
    from oct_recon_af_v1_amd import rawio
    from oct_recon_af_v1_amd import synthetic as SY
    from oct_recon_af_v1_amd.geometry import beta_from_index_units, build_spectral_geometry
    from oct_recon_af_v1_amd.pipeline import ReconConfig, Reconstructor
    print(f"case: tiled volume from {kind} files (2 B-scan repeats)", flush=True)
    torch = xp.torch
    spec = SY.SpectrumSpec()
    lam = spec.lambda_nm()
    bref = beta_from_index_units(4e-5, lam)
    dz = float(np.diff(build_spectral_geometry(lam, bref, 1.33).z_um[:2])[0])
    ny, nx, nb = 4, 128, 2
    grid = _grid(ny, nx)
    rng = np.random.default_rng(3)
    dev_tiles, sources = {}, {}
    for k, key in enumerate([(0, 0, 0), (1, 0, 0), (0, 0, 1), (1, 0, 1)]):
        raw = SY.make_bscans(ny * nb, nx, spec, beta=4e-5, seed=70 + k).numpy()        # (files, 25 + nx, N)
        dev_tiles[key] = torch.from_numpy(raw.view(np.int16)).to(xp.device)
        folder = tmp_path / f"tile{k}"
        if kind == "ocity":
            rawfiles.write_ocity(folder, raw.view(np.int16))
            sources[key] = rawio.OcityTile(folder, raw.shape[1], raw.shape[2])
        else:
            # the 12-bit samples with random bits above them (the unpack keeps the low 12)
            lines = raw.astype(np.uint16) | (rng.integers(0, 16, size=raw.shape).astype(np.uint16) << 12)
            rawfiles.write_srr(folder, lines, y_total=ny, b_total=nb, apod=(0, 25), scan=(25, 25 + nx))
            sources[key] = rawio.SrrTile(folder)
    cfg = ReconConfig(dispersion_quadratic_term=bref, focus_positions=512.0, focus_sigma=30.0 / dz,
                      crop_z_range_mm=[0.0, 0.08], output_pixel_size_um=1.0, n_medium=1.33, batch_frames=3)
    src = rawio.FileTileSource(sources, xp, io_threads=4)
    a = Reconstructor(cfg, grid, lam, src, spec.apod, xp=xp, bscan_avg=nb).run()["volume"]["db"].cpu().numpy()
    b = Reconstructor(cfg, grid, lam, dev_tiles, spec.apod, xp=xp, bscan_avg=nb).run()["volume"]["db"].cpu().numpy()
    src.close()
    assert np.isfinite(a).any()
    assert np.array_equal(a, b, equal_nan=True)
