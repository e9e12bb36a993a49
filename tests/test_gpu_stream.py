"""Per-tile dispersion and the streamed volume (oct_recon_af_v1_amd/stream.py) on the device.

Bars:
  * a tile reconstructed with its own dispersion window (plain: per-launch window pointer; MFMA:
    operator rebuilt in place) equals a volume built with that dispersion, bit for bit;
  * streamed tiles land in HBM bit-identical to the host generator's output, and the volume
    does not depend on the ring depth;
  * the MFMA volume has at least 99.9 % of voxels within 1 LSB of the plain FP32 volume (both
    quantised with the plain volume's clim);
  * with the learned estimator (search fallback), per-tile dispersion and per-depth focus land
    within the estimator's tolerances of the generator's values (median over tiles);
  * a volume streamed through a ring of fewer resident y-tile rows (the plain validation then
    row by row) equals the all-resident volume bit for bit;
  * the full configuration-4 volume (12 x 9 tiles x 8 depths of 40 x 1000 A-scans, 145 GB of
    raw data) streamed through 2 resident rows meets the bars above.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def xp():
    from oct_recon_af_v1_amd.backend import get_xp
    return get_xp("hip")[0]


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("form", ["plain", "mfma"])
def test_tile_window_equals_volume_dispersion(xp, form):
    from oct_recon_af_v1_amd import synthetic as SY
    from oct_recon_af_v1_amd.geometry import ScanGrid, beta_from_index_units, build_spectral_geometry
    from oct_recon_af_v1_amd.pipeline import ReconConfig, Reconstructor
    _say(f"case: per-tile window vs volume dispersion, {form}")
    spec = SY.SpectrumSpec()
    lam = spec.lambda_nm()
    b0, b1 = beta_from_index_units(4e-5, lam), beta_from_index_units(1.3e-4, lam)
    dz_um = float(np.diff(build_spectral_geometry(lam, b0, 1.33).z_um[:2])[0])
    ny, nx = 4, 128
    grid = ScanGrid(np.array([-0.058, 0.058]), np.array([0.0]), np.array([0.0]), nx * 1e-3, ny * 1e-3, nx, ny,
                    pixel_size_um=1.0)
    tiles = {(xi, 0, 0): SY.make_bscans(ny, nx, spec, beta=1.3e-4 if xi == 0 else 4e-5, seed=40 + xi,
                                        device=xp.device) for xi in range(2)}

    def cfg(beta):
        return ReconConfig(dispersion_quadratic_term=beta, focus_positions=512.0, focus_sigma=30.0 / dz_um,
                           crop_z_range_mm=[0.0, 0.1], output_pixel_size_um=1.0, n_medium=1.33, batch_frames=ny,
                           form=form)
    # tile (0, 0, 0) at b1, tile (1, 0, 0) at the volume's b0
    a = Reconstructor(cfg(b0), grid, lam, tiles, spec.apod, xp=xp, tile_betas={(0, 0, 0): b1})
    da = a.run()["volume"]["db"].cpu().numpy()
    da2 = a.run()["volume"]["db"].cpu().numpy()                    # rebuilt / restored windows
    assert np.array_equal(da, da2, equal_nan=True)
    # the same, each tile in a volume built with its dispersion
    ref = {}
    for xi, b in ((0, b1), (1, b0)):
        r = Reconstructor(cfg(b), grid, lam, {k: v for k, v in tiles.items() if k[0] == xi}, spec.apod, xp=xp)
        ref[xi] = r
    # compare the per-tile magnitudes that feed the stitcher
    for xi, b in ((0, b1), (1, b0)):
        st = a.stitchers[(xi, 0)]
        lo, hi = a._tile_rows(st)
        a._use_tile_window((xi, 0, 0), lo, hi)
        ma = a._magnitudes(tiles[(xi, 0, 0)], ny, lo, hi, xp.torch.empty((ny, nx, hi - lo), device=xp.device))
        a._restore_windows()
        r = ref[xi]
        mb = r._magnitudes(tiles[(xi, 0, 0)], ny, lo, hi, xp.torch.empty((ny, nx, hi - lo), device=xp.device))
        assert np.array_equal(ma.cpu().numpy(), mb.cpu().numpy()), xi
    # without the per-tile window the first tile differs
    c = Reconstructor(cfg(b0), grid, lam, tiles, spec.apod, xp=xp).run()["volume"]["db"].cpu().numpy()
    assert not np.array_equal(c, da, equal_nan=True)


def _small(**kw):
    from oct_recon_af_v1_amd.stream import StreamSpec
    base = dict(tiles_x=2, tiles_y=2, depths=2, n_y=6, n_x=256, crop_z_mm=(-0.03, 0.08))
    base.update(kw)
    return StreamSpec(**base)


def test_streamed_volume_given_estimates(xp):
    from oct_recon_af_v1_amd.stream import StreamedVolume
    _say("case: streamed volume 2x2x2 tiles, given dispersion/focus, ring 2 and 3")
    torch = xp.torch
    ss = _small()
    sv = StreamedVolume(xp, ss, ring=2, gen_threads=8, estimator="given")
    r = sv.run(validate=True)
    _say({k: v for k, v in r.items() if k != "volume"})
    for i in (0, len(sv.keys) - 1):
        host = np.empty((ss.n_y, sv.lines, sv.n), np.uint16)
        sv.gen.generate(host, ss.n_x, sv.truth["beta_idx"][i], sv.truth["focus"][i], int(sv.truth["seed"][i]))
        dev = sv.tile_raw(i).view(torch.int16).cpu().numpy().view(np.uint16)
        assert np.array_equal(host, dev), i
    assert r["within_1lsb_of_plain"] >= 0.999
    sv3 = StreamedVolume(xp, ss, ring=3, gen_threads=8, estimator="given")
    r3 = sv3.run(validate=False)
    assert np.array_equal(r["volume"]["db"].cpu().numpy(), r3["volume"]["db"].cpu().numpy(), equal_nan=True)


def test_streamed_volume_learned_estimator(xp):
    from oct_recon_af_v1_amd import learned as LN
    from oct_recon_af_v1_amd.stream import StreamedVolume
    _say("case: streamed volume 2x2x2 tiles of 8 x 512, learned estimator")
    ss = _small(n_y=8, n_x=512)
    sv = StreamedVolume(xp, ss, ring=2, gen_threads=8, estimator="learned")
    sv.warmup()
    r = sv.run(validate=True)
    _say({k: v for k, v in r.items() if k != "volume"})
    assert r["beta_err_idx_median"] <= LN.BETA_TOL_IDX
    assert r["focus_err_bins_median"] <= LN.FOCUS_TOL_BINS
    assert r["within_1lsb_of_plain"] >= 0.999


def test_row_table_path_equals_argument_path(xp, monkeypatch):
    """Rows of more tiles than the argument block holds read the descriptors from a device
    table, and a volume made row by row while streaming equals one made after the stream: the
    same volume, bit for bit."""
    from oct_recon_af_v1_amd import pipeline as P
    from oct_recon_af_v1_amd.stream import StreamedVolume
    _say("case: row stitch through the device tile table")
    sv = StreamedVolume(xp, _small(tiles_x=3, depths=2), ring=2, gen_threads=8, estimator="given")
    r = sv.run(validate=False)
    a = r["volume"]["db"].cpu().numpy()
    monkeypatch.setattr(P, "_ROW_ARG_TILES", 2)
    # the whole volume at once (after the stream) with the focus the rows were made with
    rec = sv.reconstructor(sv.truth["beta_idx"], r["focus_per_depth"])
    b = rec.reconstruct_volume(rec.allocate_volume())["db"].cpu().numpy()
    assert np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("kind", ["ocity", "srr"])
def test_streamed_volume_from_files(xp, tmp_path, kind):
    """The same tiles streamed from raw tile folders (read into the pinned ring, SRR payloads
    unpacked on the device) give the generator-streamed volume bit for bit."""
    import rawfiles
    from oct_recon_af_v1_amd import rawio
    from oct_recon_af_v1_amd.stream import StreamedVolume
    _say(f"case: streamed volume from {kind} tile folders")
    ss = _small()
    gen = StreamedVolume(xp, ss, ring=2, gen_threads=8, estimator="given")
    ref = gen.run(validate=False)["volume"]["db"].cpu().numpy()
    rng = np.random.default_rng(9)
    readers = {}
    for i, key in enumerate(gen.keys):
        raw = gen.tile_raw(i).view(xp.torch.int16).cpu().numpy()
        folder = tmp_path / f"t{i}"
        if kind == "ocity":
            rawfiles.write_ocity(folder, raw)
            readers[key] = rawio.OcityTile(folder, raw.shape[1], raw.shape[2])
        else:
            lines = raw.view(np.uint16) | (rng.integers(0, 16, size=raw.shape).astype(np.uint16) << 12)
            rawfiles.write_srr(folder, lines, y_total=ss.n_y, b_total=1, apod=(0, 25), scan=(25, raw.shape[1]))
            readers[key] = rawio.SrrTile(folder)
    del gen
    sv = StreamedVolume(xp, ss, ring=2, gen_threads=8, estimator="given", readers=readers, io_threads=4)
    r = sv.run(validate=False)
    assert r["source"] == "files"
    assert np.array_equal(r["volume"]["db"].cpu().numpy(), ref, equal_nan=True)


def test_bounded_residency_same_bits(xp):
    """1 and 2 resident rows of a 4-row volume give the all-resident volume, bit for bit, with the
    plain validation run row by row."""
    from oct_recon_af_v1_amd.stream import StreamedVolume
    _say("case: streamed volume 2x4x2 tiles, 1 / 2 / all rows resident")
    ss = _small(tiles_y=4)
    full = StreamedVolume(xp, ss, ring=3, gen_threads=8, estimator="given")
    r = full.run(validate=True)
    ref = r["volume"]["db"].cpu().numpy()
    assert r["resident_rows"] == 4
    del full
    for rows in (1, 2):
        sv = StreamedVolume(xp, ss, ring=3, gen_threads=8, estimator="given", resident_rows=rows)
        got = sv.run(validate=True)
        _say(f"{rows} resident rows: {got['raw_resident_GB']:.3f} GB, within 1 LSB of plain "
             f"{got['within_1lsb_of_plain']:.6f}, {got['validation']}")
        assert got["resident_rows"] == rows and sv.slots.shape[0] == rows
        assert np.array_equal(got["volume"]["db"].cpu().numpy(), ref, equal_nan=True)
        assert got["within_1lsb_of_plain"] == r["within_1lsb_of_plain"]
    with pytest.raises(ValueError):
        StreamedVolume(xp, ss, estimator="given", resident_rows=2, rowwise=False)


def test_full_size_streamed_volume(xp):
    """Configuration 4 at full size: 864 tiles, host-generated and streamed through pinned async
    copies into a ring of 2 resident y-tile rows (about 32 GB instead of 145 GB), per-tile
    dispersion and focus from the learned estimator (search fallback), MFMA reconstruction
    row by row, the plain FP32 volume made beside it row by row as the referee."""
    from oct_recon_af_v1_amd import learned as LN
    from oct_recon_af_v1_amd.stream import StreamedVolume, StreamSpec
    _say("case: full configuration-4 streamed volume (12 x 9 x 8 tiles, 40 x 1000 A-scans), 2 rows resident")
    sv = StreamedVolume(xp, StreamSpec(), ring=8, estimator="learned", resident_rows=2)
    sv.warmup()
    r = sv.run(validate=True)
    _say({k: v for k, v in r.items() if k not in ("volume", "focus_per_depth")})
    _say(f"end to end {r['ascans_per_s'] / 1e6:.2f} M A-scans/s over {r['ascans'] / 1e6:.2f} M A-scans "
         f"({r['raw_GB']:.0f} GB raw, {r['raw_resident_GB']:.1f} GB resident); fallback rate "
         f"{r['fallback_rate_bscans']:.2%} of B-scans; within 1 LSB of plain {r['within_1lsb_of_plain']:.6%}")
    assert r["ascans"] == 12 * 9 * 8 * 40 * 1000
    assert r["within_1lsb_of_plain"] >= 0.999
    assert r["beta_err_idx_median"] <= LN.BETA_TOL_IDX
    assert r["focus_err_bins_median"] <= LN.FOCUS_TOL_BINS
    fpd = np.asarray(r["focus_per_depth"]) - 1                  # 0-based, as the generator's focus
    truth = [np.median(sv.truth["focus"][[i for i, k in enumerate(sv.keys) if k[2] == d]]) for d in range(8)]
    _say(f"focus per depth: {np.round(fpd, 1).tolist()} vs generator {np.round(truth, 1).tolist()}")
    assert np.max(np.abs(fpd - truth)) <= LN.FOCUS_TOL_BINS
