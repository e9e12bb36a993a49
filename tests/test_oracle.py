"""The oracle pinned against what does not depend on it: the MATLAB plane stored with the
reference's tests, a brute-force DFT, closed forms and special cases.  CPU only."""
import numpy as np
import pytest

import legacy_case
from oracle import ref_path as R


def _legacy_plane(precision):
    c = legacy_case.load("telesto")
    s = c["scan"]
    geo = R.spectral_geometry(c["lambda_nm"], s["dispersion"], s["tissue_ri"])
    tg = R.tiled_geometry(s["x_centers_mm"], s["y_centers_mm"], s["z_depths_mm"], s["tile_range_x_mm"],
                          s["tile_range_y_mm"], s["n_x_px"], s["n_y_px"], geo.z_um, s["focus_pix"],
                          s["pixel_size_um"])
    raws = [c["raw"][k][None] for k in range(len(s["tiles"]))]
    p = R.reconstruct_row(raws, [tuple(t) for t in s["tiles"]], [s["frame"]], geo, tg, s["x_centers_mm"],
                          s["z_depths_mm"], s["focus_pix"], s["focus_sigma"], s["apod_size"], s["opc_poly"],
                          precision)[0]
    return p.astype(float), c["plane_db"].astype(float), c["tie"]


def test_oracle_float64_matches_matlab_plane():
    """Whole path (geometry, pre-FFT, FFT, optical path, focus weights, stitching, finalise)
    against the legacy MATLAB plane (reference tests/test_formats.py:164-180, 1e-4 dB)."""
    p, m, tie = _legacy_plane("float64")
    assert p.shape == m.shape
    fin = np.isfinite(p) & np.isfinite(m)
    assert np.abs(p - m)[fin].max() < 1e-4
    mism = np.isfinite(p) != np.isfinite(m)
    assert not np.any(mism & ~tie)
    assert mism.sum() <= max(1, 0.01 * tie.sum())


def test_oracle_float32_cpu_path_matches_matlab_plane():
    """The float32 reference CPU path (C pre-FFT with the Numba typing + pocketfft) is within
    the reference's float32 tolerance (tests/test_torch_backend.py:66-73: < 1e-2 dB)."""
    p, m, tie = _legacy_plane("float32")
    fin = np.isfinite(p) & np.isfinite(m)
    assert np.abs(p - m)[fin].max() < 1e-2
    assert not np.any((np.isfinite(p) != np.isfinite(m)) & ~tie)


def test_dispersion_window_identity():
    """Reference tests/test_estimation.py:98-106: only beta changes the window."""
    lam = np.linspace(800, 1000, 256)
    g0 = R.spectral_geometry(lam, 0.0, 1.33)
    g = R.spectral_geometry(lam, 8.949e7, 1.33)
    k = 2 * np.pi / g0.lambda_eq
    assert np.allclose(g0.window * np.exp(-1j * 8.949e7 * (k - k.mean()) ** 2), g.window, rtol=1e-9, atol=1e-9)
    # Hann normalised to unit rms
    assert abs(np.mean(np.abs(g0.window) ** 2) - 1) < 1e-12


def test_sinc_is_identity_on_equispaced_k():
    """lambda linear in 1/k: the equispaced grid coincides with the pixels, every sinc weight
    is sinc(integer) -> linearisation returns the (mean-removed, re-added) input."""
    N = 64
    k = np.linspace(2 * np.pi / 800, 2 * np.pi / 1000, N)
    lam = 2 * np.pi / k
    g = R.spectral_geometry(lam, 0.0, 1.33)
    W = np.zeros((N, N))
    for i in range(N):
        for t in range(g.idx.shape[1]):
            W[i, g.idx[i, t]] += g.w[i, t]
    assert np.allclose(W, np.eye(N), atol=1e-12)


def test_fft_against_brute_force_dft():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(3, 64)) + 1j * rng.normal(size=(3, 64))
    import scipy.fft
    assert np.allclose(scipy.fft.ifft(x.astype(np.complex64), axis=-1), R.dft_brute(x), atol=1e-5)


def test_c_prefft_matches_float64_restatement():
    """Two independent restatements of the pre-FFT stage (C with the Numba typing; numpy
    float64) agree to float32 rounding, for int16, uint16 and float32 raw input."""
    rng = np.random.default_rng(1)
    lam = 1300 + 0.0625 * (np.arange(256) - 127.5)
    g = R.spectral_geometry(lam, 3e7, 1.33)
    base = rng.integers(500, 3500, size=(2, 13, 256))
    for dt in (np.int16, np.uint16, np.float32):
        raw = base.astype(dt)
        spec = R.fused_pre_fft(raw, 5, g)
        x = raw.astype(np.float64)
        d = x[:, 5:, :] - x[:, :5, :].mean(axis=1, keepdims=True)
        mv = d.mean(-1, keepdims=True)
        dd = d - mv
        eq = np.zeros_like(dd)
        for t in range(g.idx.shape[1]):
            eq += g.w[:, t] * dd[..., g.idx[:, t]]
        ref = (eq + mv) * g.window
        assert np.abs(spec - ref).max() <= 2e-6 * np.abs(ref).max()


@pytest.mark.parametrize("n,z0,nzc,beta", [(256, 0, 128, 3e7), (256, 37, 20, -5e7), (2048, 480, 64, 8.9e7)])
def test_operator_fp64_equals_spectral_path(n, z0, nzc, beta):
    """The dense float64 operator (sinc band + the rank-1 mean term + window/N + cropped inverse
    DFT) applied to raw - background equals the float64 spectral path (pre-FFT restatement, then
    numpy's inverse FFT and the crop), complex values and magnitudes, to 1e-12 of the peak.  A
    dropped mean term, a conjugated DFT or a shifted crop fails it."""
    rng = np.random.default_rng(n + z0)
    lam = 1300 + (0.0625 * 2048 / n) * (np.arange(n) - (n - 1) / 2) - 4e-6 * (np.arange(n) - (n - 1) / 2) ** 2
    g = R.spectral_geometry(lam, beta, 1.33)
    apod = 3
    raw = rng.integers(200, 4000, size=(2, apod + 5, n)).astype(np.uint16)
    A = R.operator_fp64(g, z0, nzc)
    x = raw.astype(np.float64)
    d = x[:, apod:, :] - x[:, :apod, :].mean(axis=1, keepdims=True)
    X = np.einsum("zn,bxn->bxz", A, d)
    ref = R.complex_fp64(raw, apod, g)[..., z0:z0 + nzc]
    peak = np.abs(ref).max()
    assert np.abs(X - ref).max() <= 1e-12 * peak
    mag = R.magnitude_fp64(raw, apod, g)[..., z0:z0 + nzc]
    assert np.abs(np.abs(X) - mag).max() <= 1e-12 * peak


def test_c_prefft_background_only_is_zero():
    lam = 1300 + 0.0625 * (np.arange(256) - 127.5)
    g = R.spectral_geometry(lam, 0.0, 1.33)
    line = np.random.default_rng(2).integers(100, 4000, 256).astype(np.int16)
    raw = np.broadcast_to(line, (2, 9, 256)).copy()
    assert np.all(R.fused_pre_fft(raw, 4, g) == 0)


def test_interp_matrix_properties():
    src = np.array([0.0, 1.0, 2.5, 4.0])
    dst = np.array([-0.5, 0.0, 0.5, 1.0 + 1e-12, 2.0, 4.0, 4.5])
    W = R.interp_matrix(src, dst)
    f = 3 * src - 1
    inside = (dst >= 0) & (dst <= 4)
    assert np.allclose((W @ f)[inside], 3 * dst[inside] - 1)
    assert np.all(W[~inside] == 0)
    assert np.allclose(W[inside].sum(1), 1)
    assert W[3, 1] == 1.0 and W[3, 2] == 0.0          # node snapping within 1e-9


def test_quantisation_closed_forms():
    clim = (-10.0, 50.0)
    db = np.array([np.nan, -np.inf, np.inf, -10.0, 50.0, 20.0, -20.0, 80.0])
    b = R.db_to_bits(db, clim)
    assert b.tolist() == [0, 1, 65535, 1, 65535, 32768, 1, 65535]
    # round half away from zero (not to even) where the scaling is exact
    assert R.db_to_bits(np.array([0.5, 1.5, 2.5, 2.4999]), (0.0, 65534.0)).tolist() == [2, 3, 4, 3]
    lsb = 60.0 / 65534
    # dequantise(quantise) is within half an LSB
    x = np.random.default_rng(3).uniform(-10, 50, 1000)
    back = R.bits_to_db(R.db_to_bits(x, clim), clim)
    assert np.abs(back - x).max() <= 0.5 * lsb * (1 + 1e-9)


def test_factor_z_floor_and_peak():
    f = R.factor_z(100, 40.0, 10.0)
    assert f.argmax() == 39 and abs(f.max() - (1 + np.exp(-4.5))) < 1e-15
    assert f.min() > np.exp(-4.5)
    # effective Gaussian std is sqrt(2) sigma: e^-1/2 point at 40 +- sqrt(2)*10
    z = 40 + np.sqrt(2) * 10
    val = np.exp(-(z - 40) ** 2 / (2 * 10.0) ** 2)
    assert abs(val - np.exp(-0.5)) < 1e-12


@pytest.mark.parametrize("n", [1, 2, 7])
def test_colon_linspace_endpoints(n):
    y = R.linspace_m(0.3, 1.7, n + 1)
    assert y[0] == 0.3 and y[-1] == 1.7
    c = R.colon_m(0.0, 0.1, 0.3)
    assert len(c) == 4 and c[-1] == pytest.approx(0.3)


def test_binning_matches_matlab_filter2():
    """AScanBinning: filter2(ones(1,B)/B, 1:10) then (:, max(1,floor(B/2)):B:end), the MATLAB
    R2024a values held by the reference's tests (tests/test_formats.py:289-299)."""
    x = np.arange(1, 11, dtype=float)[None, :, None]
    full = {2: [1.5, 2.5, 3.5, 4.5, 5.5, 6.5, 7.5, 8.5, 9.5, 5.0],
            3: [1.0, 2, 3, 4, 5, 6, 7, 8, 9, 19 / 3],
            4: [1.5, 2.5, 3.5, 4.5, 5.5, 6.5, 7.5, 8.5, 6.75, 4.75]}
    for B, f in full.items():
        want = np.asarray(f)[max(1, B // 2) - 1::B]
        np.testing.assert_allclose(R.bin_ascans(x, B, np.float64)[0, :, 0], want, rtol=1e-15)
        np.testing.assert_allclose(R.bin_ascans(x, B, np.float32)[0, :, 0], want, rtol=2e-7)
        assert len(R.binned_columns(10, B)) == len(want)


@pytest.mark.parametrize("B", [1, 2, 3, 4, 5, 7])
def test_binning_brute_force_window(B):
    """Every kept line is the zero-padded mean of lines [c - ceil(B/2) + 1, c + floor(B/2)]
    (filter2 'same' with a length-B box), checked by a direct loop on ragged line counts."""
    rng = np.random.default_rng(B)
    for n in (1, B, 2 * B + 1, 23):
        if len(R.binned_columns(n, B)) == 0:          # fewer lines than the first kept one
            continue
        x = rng.normal(size=(2, n, 3))
        got = R.bin_ascans(x, B, np.float64)
        cols = R.binned_columns(n, B)
        assert got.shape == (2, len(cols), 3)
        for k, c in enumerate(cols):
            lo, hi = c - (B + 1) // 2 + 1, c + B // 2
            want = sum(x[:, j] for j in range(max(lo, 0), min(hi, n - 1) + 1)) / B
            np.testing.assert_allclose(got[:, k], want, rtol=1e-13, atol=1e-15)


def test_prepare_raw_keeps_unbinned_background():
    """Binning leaves the background lines (or the frame-mean line, taken from the unbinned
    lines) untouched and bins only the scan lines."""
    rng = np.random.default_rng(4)
    raw = rng.integers(0, 4000, size=(4, 25 + 11, 8)).astype(np.uint16)
    pre, a = R.prepare_raw(raw, 25, 3, "lines")
    assert a == 25 and pre.shape == (4, 25 + len(R.binned_columns(11, 3)), 8)
    assert np.array_equal(pre[:, :25], raw[:, :25].astype(np.float32))
    assert np.array_equal(pre[:, 25:], R.bin_ascans(raw[:, 25:], 3))
    pre, a = R.prepare_raw(raw, 25, 2, "frame_mean", 2)
    assert a == 1
    fm = raw.astype(np.float64).reshape(2, 2, 36, 8).mean(2).mean(1)
    np.testing.assert_allclose(pre[:, 0], np.repeat(fm, 2, 0), rtol=1e-6)
    assert np.array_equal(pre[:, 1:], R.bin_ascans(raw, 2))


def test_average_repeats_axes():
    """B-scan repeats are consecutive files, A-scan repeats consecutive A-lines: with values
    1000 f + 100 b + 10 x + a the means are exact, and a swapped axis changes them."""
    F, nb, X, na, Z = 3, 4, 5, 2, 3
    f, b, x, a, z = np.meshgrid(*(np.arange(n) for n in (F, nb, X, na, Z)), indexing="ij")
    mag = (1000 * f + 100 * b + 10 * x + a + 0 * z).astype(np.float32).reshape(F * nb, X * na, Z)
    got = R.average_repeats(mag, F, nb, X, na)
    want = (1000 * np.arange(F)[:, None, None] + 100 * (nb - 1) / 2 + 10 * np.arange(X)[None, :, None]
            + (na - 1) / 2 + 0 * np.arange(Z))
    assert got.shape == (F, X, Z) and np.array_equal(got, want.astype(np.float32))
    assert R.average_repeats(mag, F * nb, 1, X * na, 1) is mag
