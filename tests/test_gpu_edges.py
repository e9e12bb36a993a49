"""Edge cases through the C ABI: empty batches, ragged batches across the kernels' block
boundaries, and arguments the library rejects.

  * empty batches (0 files / 0 rows / 0 planes) give empty outputs and launch nothing;
  * the operator (MFMA) form on ragged batches -- one A-scan, 257 A-scans (one past a 256-A-scan
    GEMM block), two files of 300 lines -- within 1e-5 of the largest float64 oracle value, the
    bar the full batches meet;
  * the plain path on the same ragged batches within 2e-6 of it (8-A-scan groups cut short);
  * rejected arguments raise ValueError carrying the library's message, and a valid call made
    afterwards still gives the right result.
"""
import numpy as np
import pytest

from oracle import ref_path as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def xp():
    from oct_recon_af_v1_amd.backend import get_xp
    return get_xp("hip")[0]


def _setup(xp, beta_idx, apod=25):
    from oct_recon_af_v1_amd import synthetic as SY
    from oct_recon_af_v1_amd.geometry import beta_from_index_units, build_spectral_geometry
    from oct_recon_af_v1_amd.spectral import HipSpectralProcessor
    lam = SY.SpectrumSpec().lambda_nm()
    bref = beta_from_index_units(beta_idx, lam)
    sg = build_spectral_geometry(lam, bref, 1.33)
    return lam, bref, HipSpectralProcessor(sg, apod, xp)


def test_empty_batches(xp):
    from oct_recon_af_v1_amd import _lib
    from oct_recon_af_v1_amd.tiffio import packbits_rows
    print("case: empty batches", flush=True)
    torch = xp.torch
    L = _lib.lib()
    _, _, sp = _setup(xp, 2e-5)
    empty = torch.zeros((0, 30, 2048), dtype=torch.int16, device=xp.device)
    assert tuple(sp.run_operator(empty).shape) == (0, 5, 1024)
    assert tuple(sp.run_operator(empty, 100, 40).shape) == (0, 5, 40)
    mag = torch.zeros((0, 6, 16), dtype=torch.float32, device=xp.device)
    assert tuple(sp.average_repeats(mag, 0, 2, 3, 2).shape) == (0, 3, 16)
    codes = torch.zeros((0, 10), dtype=torch.uint16, device=xp.device)
    enc, row_bytes = packbits_rows(xp, codes)
    assert enc.numel() == 0 and row_bytes.size == 0
    buf = torch.zeros(64, dtype=torch.int16, device=xp.device)
    _lib.check(L.oct_srr_unpack(buf.data_ptr(), 0, 4, 8, 0, 1, 1, 4, buf.data_ptr(), xp.stream()), "srr (0 files)")
    f = torch.zeros(8, dtype=torch.float32, device=xp.device)
    bits = torch.zeros(8, dtype=torch.uint16, device=xp.device)
    _lib.check(L.oct_quantize(f.data_ptr(), bits.data_ptr(), 0, 8, f.data_ptr(), f.data_ptr(), 1, xp.stream()),
               "quantize (0 planes)")
    _lib.check(L.oct_bin_lines(buf.data_ptr(), _lib.OCT_RAW_I16, 0, 4, 0, 0, 2, f.data_ptr(), xp.stream()),
               "bin (0 files)")
    xp.synchronize()


@pytest.mark.parametrize("files,lines", [(1, 1), (1, 257), (2, 300)])
def test_ragged_batches_both_forms(xp, files, lines):
    from oct_recon_af_v1_amd import synthetic as SY
    print(f"case: ragged batch {files} x {lines}", flush=True)
    lam, bref, sp = _setup(xp, 3e-5)
    raw = SY.make_bscans(files, lines, beta=3e-5, seed=40 + lines)
    geo = R.spectral_geometry(lam, bref, 1.33)
    o = R.magnitude_fp64(raw.numpy(), 25, geo)
    g_op = sp.run_operator(raw.to(xp.device)).cpu().numpy()
    g_pl = sp.magnitude(raw.to(xp.device)).cpu().numpy()
    assert g_op.shape == o.shape == g_pl.shape == (files, lines, 1024)
    scale = o.max()
    assert np.abs(g_op - o).max() <= 1e-5 * scale
    assert np.abs(g_pl - o).max() <= 2e-6 * scale
    part = sp.run_operator(raw.to(xp.device), 200, 33).cpu().numpy()      # a cropped operator, same batch
    assert np.abs(part - o[..., 200:233]).max() <= 1e-5 * scale


def test_rejected_arguments_leave_device_usable(xp):
    from oct_recon_af_v1_amd import _lib
    from oct_recon_af_v1_amd import synthetic as SY
    print("case: rejected arguments", flush=True)
    torch = xp.torch
    L = _lib.lib()
    x = torch.zeros(4 * 2 * 8 * 8, dtype=torch.bfloat16, device=xp.device)
    w = torch.zeros(16 * 2 * 9, dtype=torch.bfloat16, device=xp.device)
    b = torch.zeros(16, dtype=torch.float32, device=xp.device)
    y = torch.zeros(4 * 16 * 4 * 4, dtype=torch.bfloat16, device=xp.device)
    with pytest.raises(ValueError, match="bad conv shape"):
        _lib.check(L.oct_cnn_conv_fwd(x.data_ptr(), w.data_ptr(), b.data_ptr(), y.data_ptr(), 4, 2, 7, 7, 16,
                                      xp.stream()), "conv fwd (odd side)")
    with pytest.raises(ValueError, match="not a multiple"):
        _lib.check(L.oct_cnn_conv_fwd(x.data_ptr(), w.data_ptr(), b.data_ptr(), y.data_ptr(), 4, 2, 8, 8, 12,
                                      xp.stream()), "conv fwd (12 channels)")
    mag = torch.zeros((2, 10, 64), dtype=torch.float32, device=xp.device)
    out = torch.zeros((2, 1, 8, 8), dtype=torch.bfloat16, device=xp.device)
    with pytest.raises(ValueError, match="outside"):
        _lib.check(L.oct_logmag_crop(mag.data_ptr(), 2, 10, 64, 64, 5, 0, 4, 8, 8, 0.0, 1.0, out.data_ptr(), 1,
                                     xp.stream()), "crop past the last column")
    pay = torch.zeros((1, 4, 8), dtype=torch.int16, device=xp.device)
    dst = torch.zeros((1, 4, 8), dtype=torch.int16, device=xp.device)
    with pytest.raises(ValueError, match="outside"):
        _lib.check(L.oct_srr_unpack(pay.data_ptr(), 1, 4, 8, 0, 2, 2, 5, dst.data_ptr(), xp.stream()),
                   "srr scan region past the file")
    codes = torch.zeros((3, 7), dtype=torch.uint16, device=xp.device)
    with pytest.raises(ValueError, match="null"):
        _lib.check(L.oct_packbits_sizes(codes.data_ptr(), 3, 7, None, xp.stream()), "packbits without sizes")
    _, _, sp = _setup(xp, 0.0)
    with pytest.raises(ValueError):
        sp.run_operator(torch.zeros((1, 30, 1024), dtype=torch.int16, device=xp.device))   # 1024-pixel spectra
    # device windows are read with 16-byte loads: misaligned windows and odd strides are refused
    wins = torch.zeros((2, 2048, 2), dtype=torch.float32, device=xp.device)
    with pytest.raises(ValueError, match="16-byte aligned"):
        _lib.check(L.oct_plan_use_window(sp.plan.handle, wins.data_ptr() + 8), "window 8 bytes in")
    with pytest.raises(ValueError, match="stride"):
        _lib.check(L.oct_plan_use_windows(sp.plan.handle, wins.data_ptr(), 2 * 2048 + 2, 0), "odd window stride")
    with pytest.raises(ValueError, match="bad window shape"):
        _lib.check(L.oct_dispersion_windows(None, wins.data_ptr(), 4, 1.0, 0.0, 1.0, 1, wins.data_ptr(),
                                            wins.data_ptr(), 1024, wins.data_ptr(), xp.stream()), "1024-point windows")
    _lib.check(L.oct_plan_use_window(sp.plan.handle, None), "restore the plan's window")
    # the device still computes correctly after the rejected calls
    lam, bref, sp = _setup(xp, 1e-5)
    raw = SY.make_bscans(1, 64, beta=1e-5, seed=5)
    o = R.magnitude_fp64(raw.numpy(), 25, R.spectral_geometry(lam, bref, 1.33))
    g = sp.run_operator(raw.to(xp.device)).cpu().numpy()
    assert np.abs(g - o).max() <= 1e-5 * o.max()


@pytest.mark.parametrize("dtype", ["uint16", "int16", "float32"])
def test_background_rows_exact(xp, dtype):
    """Background rows ("lines" mode: mean of the first apod A-lines in float64, stored float32)
    equal numpy's: integer sums are exact in float64; float32 raw is summed in line order."""
    print(f"case: background rows ({dtype})", flush=True)
    torch = xp.torch
    _, _, sp = _setup(xp, 0.0)
    rng = np.random.default_rng(8)
    raw = rng.integers(0, 4096, (5, 40, 2048)).astype(dtype)
    if dtype == "float32":
        raw = raw + rng.random(raw.shape, dtype=np.float32)
    got = sp.background(xp.asarray(raw)).cpu().numpy()
    ref = (np.add.accumulate(raw[:, :25].astype(np.float64), axis=1)[:, -1] / 25).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))

