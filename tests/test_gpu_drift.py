"""Focus drift on the device (oct_recon_af_v1_amd.drift): the optical-path corrected x-median
depth profile against the oracle's restatement of the reference (bit-exact float64), and the
whole measurement on a synthetic deep scan whose focus drifts with the stage depth."""
import numpy as np
import pytest

from oracle import ref_focus as RF

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def xp():
    from oct_recon_af_v1_amd.backend import get_xp
    return get_xp("hip")[0]


def test_opc_depth_median_vs_oracle(xp):
    from oct_recon_af_v1_amd.drift import depth_profiles
    from oct_recon_af_v1_amd.stitching import OpticalPathCorrection
    print("case: optical-path corrected depth median", flush=True)
    torch = xp.torch
    rng = np.random.default_rng(5)
    F, nY, nX, nZ = 3, 10, 250, 1024
    mag = rng.gamma(2.0, 30.0, size=(F, nX, nZ)).astype(np.float32)
    mag[:, :, 100] = mag[:, :1, 100]                        # ties along x
    yy, xx = np.meshgrid(np.arange(nY), np.arange(nX), indexing="ij")
    pdz = 60.0 * np.sin(xx / 37.0) - 30.0 + 0.7 * yy + 0.25    # shifts both ways, rows leaving [0, nZ)
    pdz[4] = 50.25                                            # frame 4: rows >= 973 exist in no column
    oc = OpticalPathCorrection(np.floor(pdz + 0.5).astype(np.int64), pdz, nZ)
    lo, hi = oc.valid_bounds()
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.int32), device=xp.device)  # noqa: E731
    frames = [0, 4, 9]
    dev = torch.from_numpy(mag).to(xp.device)
    got = depth_profiles(xp, dev, frames, (t(oc.shift), t(lo), t(hi)))
    want = RF.opc_depth_median(mag, frames, oc.shift, pdz)
    assert got.shape == want.shape
    assert np.array_equal(got, want, equal_nan=True)
    assert np.isnan(want).any() and np.isfinite(want).any()
    plain = depth_profiles(xp, dev)
    assert np.array_equal(plain, np.median(mag.astype(float), axis=1))


def test_measure_focus_drift_synthetic(xp):
    """8 focus depths 50 um apart, the confocal focus drifting 0.04 px per um of stage depth:
    every measured depth accepted, the fitted slope within 40 % of the generator's and the focus
    per depth within 5 px of it (the detector's peak of the detrended log profile sits 3-4 px
    deeper than the gate centre of this sample: speckle medium over a noise floor)."""
    from oct_recon_af_v1_amd import synthetic as SY
    from oct_recon_af_v1_amd.geometry import ScanGrid, beta_from_index_units, build_spectral_geometry
    from oct_recon_af_v1_amd.pipeline import ReconConfig, Reconstructor
    print("case: focus drift on a synthetic deep scan", flush=True)
    spec = SY.SpectrumSpec()
    lam = spec.lambda_nm()
    bidx = 6e-5
    bref = beta_from_index_units(bidx, lam)
    sg = build_spectral_geometry(lam, bref, 1.33)
    dz = float(np.median(np.diff(sg.z_um)))
    ny, nx = 20, 300
    zd = np.arange(8) * 0.05
    grid = ScanGrid(np.array([0.0]), np.array([0.0]), zd, nx * 1e-3, ny * 1e-3, nx, ny, pixel_size_um=1.0)
    m, f0 = 0.04, 380.0
    truth = f0 + m * zd * 1e3                                  # 0-based focus bin per depth
    tiles = {(0, 0, zi): SY.make_focus_bscans(ny, nx, bidx, truth[zi], spec, seed=100 + zi, device=xp.device)
             for zi in range(len(zd))}
    cfg = ReconConfig(dispersion_quadratic_term=bref, focus_positions=400.0, focus_sigma=30.0 / dz,
                      crop_z_range_mm=[0.0, 0.1], output_pixel_size_um=1.0, n_medium=1.33, batch_frames=ny)
    rec = Reconstructor(cfg, grid, lam, tiles, spec.apod, xp=xp)
    focus, diag, per = rec.measure_focus_drift(prior_pix=truth[1] + 1)
    print("measured", [p["detected_pix"] for p in per], "fit", focus.tolist(), diag["driftRegime"],
          f"slope {diag['slopeUsed_pixPerUm']:.4f} px/um", flush=True)
    assert len(per) == 7 and all(p["accepted"] for p in per)
    assert diag["driftRegime"] == "normal"
    assert abs(diag["slopeUsed_pixPerUm"] - m) < 0.4 * m
    assert np.max(np.abs(focus - (truth + 1))) <= 5
