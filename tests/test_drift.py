"""Focus-drift fit and depth selection (oct_recon_af_v1_amd.drift) against the known answers of
the reference's own tests (tests/test_estimation.py:32-84) and closed forms.  CPU only."""
import numpy as np
import pytest

from oct_recon_af_v1_amd import drift as D

DZ = 1.4339      # tests/test_estimation.py:32-33
NZ = 1024


def test_theil_sen_exact_line():
    z = np.array([50, 100, 150, 200.0])
    s, b, ok = D.theil_sen(z, 0.2 * z + 400)
    assert ok and np.isclose(s, 0.2) and np.isclose(b, 400)
    assert D.theil_sen([10, 10], [1, 2])[2] is False


def test_drift_synthetic_normal_with_outlier():
    rng = np.random.default_rng(0)
    na, ni, ns = 1.4, 1.33, 1.45
    m = (ns ** 2 - ni ** 2) / (ni * na) / DZ
    z_mm = np.arange(0.05, 0.401, 0.05)
    pix = 430 + m * z_mm * 1e3 + rng.normal(0, 1.0, z_mm.size)
    pix[3] -= 60
    zall = np.arange(-0.1, 0.4001, 0.01)
    focus, d = D.fit_drift(z_mm, pix, zall, DZ, NZ, tissue_ri=na)
    assert d["driftRegime"] == "normal"
    assert d["rejectedIdx"] == [3]
    assert abs(d["slopeUsed_pixPerUm"] - m) < 0.05 * m
    assert abs(d["tissueRI"] - ns) < 0.02
    assert np.all(focus[zall <= 0] == focus[0])
    assert np.max(np.abs(focus - np.round(430 + m * np.maximum(zall, 0) * 1e3))) <= 2
    assert np.all(np.isfinite(d["seAtDepth_pix"]))


def test_drift_no_drift_and_gel_clicks_ignored():
    z_mm = np.array([-0.05, 0.05, 0.1, 0.15, 0.2])
    pix = np.array([300, 433.4, 432.6, 433.2, 432.9])
    focus, d = D.fit_drift(z_mm, pix, [0, 0.1, 0.2], DZ, NZ, tissue_ri=1.33)
    assert d["driftRegime"] == "no-measurable-drift"
    assert 0 not in d["keptIdx"] and any("stage z < 0" in m for m in d["messages"])
    assert np.all(focus == 433)


def test_drift_single_click_clamps_and_structure():
    focus, d = D.fit_drift([0.0], [433], np.arange(-0.03, 0.041, 0.01), DZ, NZ)
    assert np.all(focus == 433) and d["driftSlope"] == 0
    z = np.array([0.05, 0.1, 0.15, 0.2])
    pix = 400 + 2.0 * z * 1e3 + np.array([0.3, -0.2, 0.1, -0.1])
    _, d = D.fit_drift(z, pix, [0.1], DZ, NZ)
    assert d["driftRegime"] == "clamped-to-max" and d["wasSlopeClamped"]
    pix = 500 - 0.8 * z * 1e3 + np.array([0.3, -0.2, 0.1, -0.1])
    _, d = D.fit_drift(z, pix, [0.1], DZ, NZ)
    assert d["driftRegime"] == "suspicious-structure-clicks" and np.isnan(d["tissueRI"])
    with pytest.raises(ValueError):
        D.fit_drift([np.nan], [1], [0], DZ, NZ)


def test_exact_line_is_reproduced_and_clipped():
    """A noiseless line inside the physical range comes back exactly (rounded), flat above z = 0,
    and clipped to [1, n_z]."""
    z = np.array([0.05, 0.1, 0.2, 0.3])
    m = 0.05
    f, d = D.fit_drift(z, 100 + m * z * 1e3, [-0.1, 0.0, 0.15, 0.3], DZ, NZ)
    assert d["driftRegime"] == "normal" and np.isclose(d["slopeUsed_pixPerUm"], m)
    assert f.tolist() == [100, 100, 108, 115]
    f, _ = D.fit_drift(z, 1020 + m * z * 1e3, [0.3], DZ, NZ)
    assert f.tolist() == [NZ]


def test_select_depths_to_measure():
    z = np.arange(8) * 0.05                                  # 0 .. 0.35 mm, 50 um apart
    assert D.select_depths_to_measure(z) == ([1, 2, 3, 4, 5, 6, 7], False)
    z = np.arange(-2, 30) * 0.01                             # 10 um apart: every 5th from 50 um
    idx, fb = D.select_depths_to_measure(z)
    assert not fb and np.allclose(z[idx], [0.05, 0.10, 0.15, 0.20, 0.25])
    idx, fb = D.select_depths_to_measure([0.0, 0.01, 0.02])   # shallow: from z = 0
    assert fb and idx == [0]                                # every 5th depth from z = 0
    with pytest.raises(ValueError):
        D.select_depths_to_measure([-0.02, -0.01])


def test_central_frames():
    # c = (n - 1) / 2, half = max(1, round(0.05 n)): 5 points on [c - half, c + half], numpy round
    assert D.central_frames(100) == [44, 47, 50, 52, 54]      # 44.5 47 49.5 52 54.5
    assert D.central_frames(40) == [18, 20, 22]               # 17.5 18.5 19.5 20.5 21.5
    assert D.central_frames(1) == [0]
