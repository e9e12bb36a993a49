"""Operator (MFMA) form of the spectral stage against the float64 oracle.

Tolerances:
  * the float64 operator built on the device equals the oracle's dense float64 operator to
    1e-12 of its largest entry;
  * magnitudes through the fp16 3-product GEMM: with one clim (the oracle's finite dB range),
    at least 99.9 % of voxels within 1 LSB of the float64 oracle (the precision bar); the
    measured rate and dB error are printed;
  * a cropped operator agrees with the same bins of the full one to 1e-6 of the largest value.
"""
import numpy as np
import pytest

from oracle import ref_path as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def xp():
    from oct_recon_af_v1_amd.backend import get_xp
    return get_xp("hip")[0]


def _setup(xp, beta_idx, apod=25, mode="lines"):
    from oct_recon_af_v1_amd import synthetic as SY
    from oct_recon_af_v1_amd.geometry import beta_from_index_units, build_spectral_geometry
    from oct_recon_af_v1_amd.spectral import HipSpectralProcessor
    lam = SY.SpectrumSpec().lambda_nm()
    bref = beta_from_index_units(beta_idx, lam)
    sg = build_spectral_geometry(lam, bref, 1.33)
    return lam, bref, HipSpectralProcessor(sg, apod, xp, apod_mode=mode)


def test_operator_fp64_equals_oracle(xp):
    print("case: operator fp64 build", flush=True)
    lam, bref, sp = _setup(xp, 4e-5)
    op = sp.operator_plan(300, 48, want_fp64=True)
    A = op.operator_fp64()
    geo = R.spectral_geometry(lam, bref, 1.33)
    Ao = R.operator_fp64(geo, 300, 48)
    err = np.abs(A - Ao).max() / np.abs(Ao).max()
    print(f"operator max rel diff {err:.2e}, scale 2^{op.scale_log2}", flush=True)
    assert err < 1e-12


def test_opform_sixty_db_lsb_parity(xp):
    """Config-2 style inputs (reflectivity over 60 dB), all 1024 bins."""
    from oct_recon_af_v1_amd import synthetic as SY
    print("case: opform 60 dB full crop", flush=True)
    spec = SY.SpectrumSpec(r_min=1e-7, r_max=1e-1)
    beta = 2e-5
    lam, bref, sp = _setup(xp, beta)
    raw = SY.make_bscans(6, 333, spec, beta=beta, seed=21).numpy()          # 1998 A-scans: ragged tile
    g = sp.run_operator(xp.asarray(raw)).cpu().numpy()
    geo = R.spectral_geometry(lam, bref, 1.33)
    o64 = R.magnitude_fp64(raw, 25, geo)
    frac, worst, clim = R.lsb_agreement(g, o64)
    with np.errstate(divide="ignore"):
        ddb = np.abs(20 * np.log10(g.astype(np.float64)) - 20 * np.log10(o64))
    ddb = ddb[np.isfinite(ddb)]
    rel = np.abs(g - o64).max() / o64.max()
    plain = sp.magnitude(xp.asarray(raw)).cpu().numpy()
    fp, _, _ = R.lsb_agreement(plain, o64, clim)
    print(f"opform vs fp64: within 1 LSB {frac:.5%} (max {worst}), dB err median {np.median(ddb):.2e} "
          f"p99.9 {np.quantile(ddb, 0.999):.2e} max {ddb.max():.2e}, max rel {rel:.2e}; plain path {fp:.5%}",
          flush=True)
    assert frac >= 0.999
    assert rel < 1e-5


def test_opform_crop_and_dtypes(xp):
    from oct_recon_af_v1_amd import synthetic as SY
    print("case: opform crop/int16", flush=True)
    lam, bref, sp = _setup(xp, 0.0)
    raw = SY.make_bscans(2, 300, seed=22).numpy().astype(np.int16)
    full = sp.run_operator(xp.asarray(raw)).cpu().numpy()
    part = sp.run_operator(xp.asarray(raw), 300, 77).cpu().numpy()
    assert np.abs(part - full[..., 300:377]).max() <= 1e-6 * full.max()
    geo = R.spectral_geometry(lam, bref, 1.33)
    o = R.magnitude_fp64(raw, 25, geo)
    assert np.abs(full - o).max() <= 1e-5 * o.max()
    with pytest.raises(ValueError):
        sp.run_operator(xp.asarray(raw.astype(np.float32)))


@pytest.mark.parametrize("dtype", ["uint16", "int16"])
def test_register_fragment_gemm_equals_lds_fragment_gemm(xp, dtype, monkeypatch):
    """The GEMM that forms the data fragments in registers (v_permlane32_swap) gives the bits of
    the one that stages them through LDS (OCT_OPGEMM=classic): same fragments, same products in
    the same order per accumulator.  Ragged batch, cropped and full operators."""
    from oct_recon_af_v1_amd import synthetic as SY
    print(f"case: register vs LDS fragments ({dtype})", flush=True)
    lam, bref, sp = _setup(xp, 6e-5)
    raw = SY.make_bscans(3, 301, seed=9).numpy().astype(dtype)
    d = xp.asarray(raw)
    out = {}
    for sel in ("classic", "default"):
        monkeypatch.setenv("OCT_OPGEMM", sel)
        out[sel] = (sp.run_operator(d).cpu().numpy(), sp.run_operator(d, 211, 120).cpu().numpy())
    monkeypatch.delenv("OCT_OPGEMM")
    for a, b in zip(out["classic"], out["default"]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("z0,nzc", [(400, 10), (180, 20), (300, 33), (512, 64), (211, 70), (90, 95), (600, 120)])
def test_narrow_operators_vs_oracle(xp, z0, nzc):
    """Operators of 32, 64, 96 and 128 rows are applied with 1-4 row tiles per workgroup (no
    padding to 128 rows); 70 and 95 bins take two workgroup row blocks of 3 tiles, 120 bins two
    of 4.  Each within 1e-5 of the largest float64 oracle value on its bins."""
    from oct_recon_af_v1_amd import synthetic as SY
    print(f"case: operator bins [{z0}, {z0 + nzc})", flush=True)
    lam, bref, sp = _setup(xp, -3e-5)
    raw = SY.make_bscans(2, 280, seed=12).numpy()
    geo = R.spectral_geometry(lam, bref, 1.33)
    o = R.magnitude_fp64(raw, 25, geo)
    g = sp.run_operator(xp.asarray(raw), z0, nzc).cpu().numpy()
    assert g.shape == (2, 280, nzc)
    assert np.abs(g - o[..., z0:z0 + nzc]).max() <= 1e-5 * o.max()

