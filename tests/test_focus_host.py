"""Focus-detector helpers (host, 1-D) against the reference's known answers
(tests/test_estimation.py:159-164) and a synthetic confocal profile."""
import numpy as np

from oct_recon_af_v1_amd.focus import (depth_profile, detect_peak_near, findpeaks_matlab, imgaussfilt_1d,
                                       interface_pix)


def test_matlab_helpers_known_answers():
    v = np.array([0, 1, 3, 3, 2, 5, 1, 1, 4.0])
    assert findpeaks_matlab(v).tolist() == [2, 5]          # flat peak -> first sample; no endpoints
    s = imgaussfilt_1d(np.r_[np.zeros(20), 1, np.zeros(20)], 4)
    assert np.isclose(s.sum(), 1) and np.argmax(s) == 20
    assert np.allclose(imgaussfilt_1d(np.ones(10), 4), 1)  # replicate padding
    assert len(imgaussfilt_1d(np.zeros(50), 4)) == 50 and len(np.arange(-8, 9)) == 2 * int(np.ceil(8)) + 1


def test_detect_peak_near_synthetic_confocal():
    z = np.arange(1, 1025)
    prof = 100 * np.exp(-z / 300.0) * (1 + 3 * np.exp(-((z - 437) / 25.0) ** 2)) + 1.0
    pix, prom = detect_peak_near(prof, 420, search_px=60)
    assert abs(pix - 437) <= 2 and prom > 0.3
    assert np.isnan(detect_peak_near(prof, 100, search_px=60)[0])     # nothing within +-60 of 100


def test_interface_and_profile():
    rng = np.random.default_rng(0)
    m = rng.uniform(0.5, 1.0, (3, 40, 300))
    m[:, :, 150] = 50.0                                             # bright interface at row 151 (1-based)
    m[:, :, 20] = 99.0                                              # ignored: in the top third
    pix, vals = interface_pix(m)
    assert pix == 151 and vals == [151, 151, 151]
    p = depth_profile(m)
    assert p.shape == (300,) and np.argmax(p[100:]) + 100 == 150
