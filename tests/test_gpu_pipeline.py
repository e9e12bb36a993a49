"""Tiled reconstruction on the hip device (fused per-row stitch + finalise, both spectral forms)
against the oracle's restatement of Reconstructor.process_row.

Tolerances: dB volumes quantised with one clim (the oracle's); against the float64 oracle, at
least 99.99 % of voxels within 1 LSB for the plain FP32 path and 99.9 % for the MFMA form (the
north star's bar), and the plain path also against the reference's float32 CPU path; NaN masks
identical; the fused row path equals the per-tile contribute path to 2e-6 dB; every combination
of overlapped rows and per-depth streams gives the same bits.
"""
import numpy as np
import pytest

import legacy_case
from oracle import ref_path as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def xp():
    from oct_recon_af_v1_amd.backend import get_xp
    return get_xp("hip")[0]


def _tiled(xp, ny, nx, seed=3, bscan_dev=None):
    """3 x 3 tiles (10 % x overlap) x 4 focus depths, 1 um pixels, depth 0-0.15 mm."""
    from oct_recon_af_v1_amd import synthetic as SY
    from oct_recon_af_v1_amd.geometry import ScanGrid, beta_from_index_units, build_spectral_geometry
    from oct_recon_af_v1_amd.pipeline import ReconConfig, Reconstructor
    spec = SY.SpectrumSpec()
    lam = spec.lambda_nm()
    beta = 4e-5
    bref = beta_from_index_units(beta, lam)
    dz_um = float(np.diff(build_spectral_geometry(lam, bref, 1.33).z_um[:2])[0])
    xc = (np.arange(3) - 1) * nx * 1e-3 * 0.9
    yc = (np.arange(3) - 1) * ny * 1e-3
    zd = np.arange(4) * 0.05
    grid = ScanGrid(xc, yc, zd, nx * 1e-3, ny * 1e-3, nx, ny, pixel_size_um=1.0)
    tiles = {}
    k = 0
    for yi in range(3):
        for xi in range(3):
            for zi in range(4):
                tiles[(xi, yi, zi)] = SY.make_bscans(ny, nx, spec, beta=beta, seed=seed * 100 + k,
                                                     device=bscan_dev or xp.device, chunk_files=10)
                k += 1
    cfg = ReconConfig(dispersion_quadratic_term=bref, focus_positions=512.0, focus_sigma=30.0 / dz_um,
                      crop_z_range_mm=[0.0, 0.15], output_pixel_size_um=1.0, n_medium=1.33, batch_frames=ny)
    rec = Reconstructor(cfg, grid, lam, tiles, spec.apod, xp=xp)
    return rec, grid, lam, bref, tiles


def _oracle_rows(rec, grid, lam, bref, tiles, rows, frames, precision="float64"):
    geo = R.spectral_geometry(lam, bref, 1.33)
    ot = R.tiled_geometry(grid.x_centers_mm, grid.y_centers_mm, grid.z_depths_mm, grid.tile_range_x_mm,
                          grid.tile_range_y_mm, grid.n_x_px, grid.n_y_px, geo.z_um, 512.0, 1.0, crop=[0.0, 0.15])
    out = []
    for yi in rows:
        ids = [(xi, zi) for xi in range(3) for zi in range(4)]
        if isinstance(frames, slice):
            raws = [tiles[(xi, yi, zi)][frames].cpu().numpy() for xi, zi in ids]
        else:
            raws = [np.stack([tiles[(xi, yi, zi)][int(i)].cpu().numpy() for i in frames]) for xi, zi in ids]
        fr = np.arange(raws[0].shape[0]) if isinstance(frames, slice) else np.asarray(frames)
        out.append(R.reconstruct_row(raws, ids, fr, geo, ot, grid.x_centers_mm, grid.z_depths_mm, 512.0,
                                     rec.cfg.focus_sigma, 25, None, precision))
    return np.concatenate(out)


def _lsb(db_t, db_ref):
    fin = db_ref[np.isfinite(db_ref)]
    clim = (float(fin.min()), float(fin.max()))
    a = R.db_to_bits(db_t, clim).astype(np.int64)
    b = R.db_to_bits(db_ref, clim).astype(np.int64)
    d = np.abs(a - b)
    return float(np.mean(d <= 1)), int(d.max())


def test_small_tiled_volume_vs_oracle(xp):
    print("case: small tiled volume (3x3x4 tiles, 4 x 64 A-scans)", flush=True)
    rec, grid, lam, bref, tiles = _tiled(xp, 4, 64)
    ref = _oracle_rows(rec, grid, lam, bref, tiles, [0, 1, 2], slice(None))
    ref32 = _oracle_rows(rec, grid, lam, bref, tiles, [0, 1, 2], slice(None), "float32")
    for form, bar in (("plain", 0.9999), ("mfma", 0.999)):
        rec.cfg.form = form
        res = rec.run()
        vol = res["volume"]
        db = vol["db"].cpu().numpy()
        assert db.shape == ref.shape
        assert np.array_equal(np.isfinite(db), np.isfinite(ref))
        frac, worst = _lsb(db, ref)
        frac32, worst32 = _lsb(db, ref32)
        print(f"{form}: within 1 LSB of the float64 oracle {frac:.6%} (max {worst}); of the float32 CPU path "
              f"{frac32:.6%} (max {worst32})", flush=True)
        assert frac >= bar
        if form == "plain":
            assert frac32 >= 0.9999
        # device clim and codes are exactly those of numpy on the device's own dB volume
        gmin, gmax, pcs = R.volume_clim(db)
        assert vol["clim"].cpu().numpy().tolist() == [np.float32(gmin), np.float32(gmax)]
        assert np.array_equal(vol["bits"].cpu().numpy(), R.quantise_volume(db, (float(gmin), float(gmax))))


def test_overlapped_rows_equal_serial_rows(xp):
    """Volume rows whose stitches run on their own stream beside the next row's spectral work
    (tile magnitudes alternating between two buffer sets, each focus depth's tiles on a stream
    of its own) give the same bits as rows run one after the other on the caller's stream alone,
    in both forms; also for a second volume that reuses the buffer sets, for process_row called
    in between, and for tiles that a source hands over as device copies whose last reference the
    reconstructor drops while the side streams still read them."""
    print("case: overlapped vs serial volume rows", flush=True)
    rec, grid, lam, bref, tiles = _tiled(xp, 4, 64, seed=5)
    combos = [(False, ()), (False, ("mfma", "plain")), (True, ()), (True, ("mfma", "plain")),
              (True, ("mfma", "plain"))]
    for form in ("plain", "mfma"):
        rec.cfg.form = form
        got = {}
        for overlap, forms in combos:
            print(f"  {form}: overlap_rows={overlap} concurrent_forms={forms}", flush=True)
            rec.overlap_rows = overlap
            rec.concurrent_forms = forms
            vol = rec.reconstruct_volume(rec.allocate_volume())
            mid_db, _ = rec.process_row(1)
            # the first combination is the serial reference: every launch on the caller's stream
            got.setdefault(bool(overlap or forms), []).append((vol["db"].cpu().numpy(), vol["bits"].cpu().numpy(),
                                                              mid_db.cpu().numpy()))
        # tiles from a source that hands over device copies made beforehand on the caller's
        # stream, whose last reference goes as soon as the tile's launches are issued
        resident = rec.tiles
        files = rec.frame_files(np.arange(grid.n_y_px))
        staged = {k + (files.tobytes(),): v[files[0]:files[-1] + 1].clone() for k, v in resident.items()}

        def source(xi, yi, zi, f):
            return staged.pop((xi, yi, zi, np.asarray(f).tobytes()))

        rec.tiles = source
        rec.overlap_rows = True
        rec.concurrent_forms = ("mfma", "plain")
        try:
            vol = rec.reconstruct_volume(rec.allocate_volume())
            got[True].append((vol["db"].cpu().numpy(), vol["bits"].cpu().numpy(), got[True][0][2]))
        finally:
            rec.tiles = resident
        ref_db, ref_bits, ref_mid = got[False][0]
        for db, bits, mid in got[True]:
            assert np.array_equal(db.view(np.uint32), ref_db.view(np.uint32)), form
            assert np.array_equal(bits, ref_bits), form
            assert np.array_equal(mid.view(np.uint32), ref_mid.view(np.uint32)), form


def test_tile_windows_same_bits_in_every_stream_arrangement(xp):
    """Per-tile dispersion (tile_betas: the plain form points each launch at its tile's window,
    the MFMA form rebuilds the depth's operator per tile) gives the same bits with every
    combination of overlapped rows and per-depth streams as with every launch on the caller's
    stream, in both forms; and a tile given the volume's own dispersion reconstructs as without
    a tile window."""
    print("case: per-tile windows, every stream arrangement", flush=True)
    rec, grid, lam, bref, tiles = _tiled(xp, 4, 64, seed=6)
    rng = np.random.default_rng(11)
    betas = {k: bref * (1 + rng.uniform(-0.2, 0.2)) for k in tiles}
    same = (1, 1, 2)
    combos = [(False, ()), (False, ("mfma", "plain")), (True, ()), (True, ("mfma", "plain"))]
    for form in ("plain", "mfma"):
        rec.cfg.form = form
        rec.set_tile_betas({})
        rec.overlap_rows, rec.concurrent_forms = False, ()
        plain_vol = rec.reconstruct_volume(rec.allocate_volume())["db"].cpu().numpy()
        rec.set_tile_betas(betas)
        outs = []
        for overlap, forms in combos:
            print(f"  {form}: overlap_rows={overlap} concurrent_forms={forms}", flush=True)
            rec.overlap_rows, rec.concurrent_forms = overlap, forms
            vol = rec.reconstruct_volume(rec.allocate_volume())
            outs.append((vol["db"].cpu().numpy(), vol["bits"].cpu().numpy()))
        for db, bits in outs[1:]:
            assert np.array_equal(db.view(np.uint32), outs[0][0].view(np.uint32)), form
            assert np.array_equal(bits, outs[0][1]), form
        assert not np.array_equal(outs[0][0], plain_vol)          # the tile windows took effect
        # one tile given the volume's own dispersion as its window: the same bits as no window
        rec.set_tile_betas({same: bref})
        rec.overlap_rows, rec.concurrent_forms = True, ("mfma", "plain")
        only = rec.reconstruct_volume(rec.allocate_volume())["db"].cpu().numpy()
        assert np.array_equal(only.view(np.uint32), plain_vol.view(np.uint32)), form
    rec.set_tile_betas({})


def test_fused_rows_equal_per_tile_path(xp):
    """Legacy scan with optical-path correction: fused row kernel vs per-tile contribute."""
    from oct_recon_af_v1_amd.geometry import ScanGrid
    from oct_recon_af_v1_amd.pipeline import ReconConfig, Reconstructor
    print("case: fused vs per-tile (legacy OPC scan)", flush=True)
    c = legacy_case.load("telesto")
    s = c["scan"]
    grid = ScanGrid(s["x_centers_mm"], s["y_centers_mm"], s["z_depths_mm"], s["tile_range_x_mm"],
                    s["tile_range_y_mm"], s["n_x_px"], s["n_y_px"], pixel_size_um=s["pixel_size_um"])
    tiles = {}
    rng = np.random.default_rng(0)
    for k, (xi, zi) in enumerate(s["tiles"]):
        full = np.stack([c["raw"][k], c["raw"][k][:, rng.permutation(2048)]])   # two distinct frames
        tiles[(xi, 0, zi)] = xp.asarray(full)
    out = {}
    for fused in (True, False):
        cfg = ReconConfig(dispersion_quadratic_term=s["dispersion"], focus_positions=s["focus_pix"],
                          focus_sigma=s["focus_sigma"], output_pixel_size_um=s["pixel_size_um"],
                          n_medium=s["tissue_ri"], batch_frames=1, fused_rows=fused)
        rec = Reconstructor(cfg, grid, c["lambda_nm"], tiles, s["apod_size"], opc_poly=s["opc_poly"], xp=xp)
        out[fused] = rec.process_planes([0, 1]).astype(float)
    a, b = out[True], out[False]
    assert np.array_equal(np.isfinite(a), np.isfinite(b))
    ok = np.isfinite(a)
    assert np.abs(a[ok] - b[ok]).max() < 2e-6 * max(1.0, np.abs(b[ok]).max())


def test_full_size_tiled_sampled_plane(xp):
    """The benchmark workload at full size (36 tiles x 100 x 1000 A-scans): one output plane of
    every y row against the float64 oracle's reconstruction of that plane, plus volume-wide
    properties, in the plain form (at least 99.99 % of the sampled voxels within 1 LSB) and then
    the MFMA form (at least 99.9 %, and every code within 1 LSB of the plain volume's but for at
    most 1e-4 of them)."""
    print("case: full-size tiled workload, sampled planes", flush=True)
    rec, grid, lam, bref, tiles = _tiled(xp, 100, 1000, seed=5)
    res = rec.run()
    vol = res["volume"]
    assert res["output_shape_yzx"] == [300, 150, 2800]
    db = vol["db"]
    assert not bool(xp.torch.isnan(db).any())            # every voxel sees the focus-weight floor
    frames = [37]
    ref = _oracle_rows(rec, grid, lam, bref, tiles, [0, 1, 2], frames)
    got = np.stack([db[yi * 100 + 37].cpu().numpy() for yi in range(3)])
    frac, worst = _lsb(got, ref)
    print(f"full size plain: sampled planes within 1 LSB of the oracle {frac:.6%} (max {worst})", flush=True)
    assert frac >= 0.9999
    bits = vol["bits"].cpu().numpy()
    assert bits.min() >= 1 and bits.max() == 65535
    del res, vol, db
    rec.cfg.form = "mfma"
    vol = rec.run()["volume"]
    got = np.stack([vol["db"][yi * 100 + 37].cpu().numpy() for yi in range(3)])
    frac, worst = _lsb(got, ref)
    print(f"full size mfma: sampled planes within 1 LSB of the oracle {frac:.6%} (max {worst})", flush=True)
    assert frac >= 0.999
    d = np.abs(vol["bits"].cpu().numpy().astype(np.int32) - bits.astype(np.int32))
    print(f"full size mfma: codes within 1 LSB of plain {np.mean(d <= 1):.6%} (max {int(d.max())})", flush=True)
    assert np.mean(d <= 1) >= 0.9999
