#!/usr/bin/env python
"""Write tests/golden/legacy_<case>.npz: raw tile inputs + the MATLAB output plane that the
reference stores for its legacy-parity tests (tests/data/legacy_formats_ref.npz, keys
``<case>__planeDb`` / ``<case>__tie``), so that the oracle and the device path can be checked
against MATLAB without the reference being present.

Runs in the build container only (it reads the reference checkout as data: the stored
MATLAB planes and the packaged chirp tables).  The raw inputs are regenerated here by a
restatement of the reference's seeded synthetic generator
(scripts/make_synthetic_volume.py:54-122, 268-356: numpy default_rng stream order, source
spectrum, reflector model, 12-bit rounding and clipping).  Only the y frame that the stored
plane needs is kept.  The generator's SHA-1 over the written tile files cannot be reused
(it also hashes the Thorlabs Header.xml text); the end-to-end match of the plane at 1e-4 dB
is the check that the regenerated inputs are the ones MATLAB saw.

    python tests/golden/make_legacy_fixture.py [/root/reference]
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = Path(__file__).resolve().parent

LAMBDA_RANGE_NM = {"gan632": (796.0, 1010.0), "ganymede": (796.23, 1010.02), "telesto": (1208.69, 1372.50)}
OPC_POLY = [0.0099168, 0.010817, -0.0001907, -0.00026705, 4.8047e-12]   # make_synthetic_volume.py:45
TILED = dict(nx=8, ny=2, tiles_x=2, tiles_y=1, depths=(0.0, 0.01), n_lambda=2048, seed=1,
             pixel_um=2.0, apod=25)                                       # tests/test_formats.py:47
PLANE_Y = 2                                                               # tests/test_formats.py:58
# Only the Telesto case is reproducible: for the Ganymede/GAN632 cases the reference's
# generator prefers a chirp file from the authors' instrument when it exists
# (scripts/make_synthetic_volume.py:42, 58-59), and the stored MATLAB planes for "base" and
# "ganymede" do not match the packaged ChirpGanymede.dat (about 1.8 dB median difference,
# same structure), so they were made from that unshipped chirp.
CASES = {"telesto": "telesto"}


def chirp_for(system: str, n: int) -> np.ndarray:
    name = "ChirpTelesto.dat" if system == "telesto" else "ChirpGanymede.dat"
    c = np.loadtxt(REF / "octrecon" / "data" / name).ravel()
    if len(c) != n:
        c = np.interp(np.linspace(0, len(c) - 1, n), np.arange(len(c)), c) * (n - 1) / (len(c) - 1)
    return c


def chirp_to_lambda(c, system):
    lmin, lmax = LAMBDA_RANGE_NM[system]
    return 1.0 / (c * (1.0 / (len(c) - 1) * (1.0 / lmax - 1.0 / lmin)) + 1.0 / lmin)


class Model:
    """Source spectrum x (1 + reflector fringes) + fixed pattern + read noise, 12-bit."""

    def __init__(self, lam, rng, n_medium=1.33):
        self.k = 2 * np.pi / lam
        n = len(lam)
        p = np.arange(n)
        self.dc = 1400.0 * np.exp(-((p - 0.5 * n) / (0.33 * n)) ** 2) + 250.0
        self.fixed = 4.0 * np.sin(p * 0.37) + 3.0 * rng.standard_normal(n)
        self.rng = rng
        self.n_medium = n_medium
        dk = (self.k.max() - self.k.min()) / (n - 1)
        self.z_nyq_um = np.pi / (2 * dk) / 1e3

    def reflectors(self, x_mm, y_mm, z_depth_mm):
        surf = 480.0 + 40.0 * np.sin(2 * np.pi * x_mm / 0.09) + 25.0 * np.cos(2 * np.pi * y_mm / 0.03)
        surf = surf - z_depth_mm * 1e3 * self.n_medium
        zs = np.array([surf, surf + 45.0, surf + 120.0, surf + 260.0, surf + 400.0])
        amp = np.array([0.20, 0.08, 0.05 * (1 + np.sin(9 * x_mm / 0.09)), 0.04, 0.03])
        return zs, amp

    def aline(self, zs, amps):
        s = np.ones_like(self.k)
        for z, a in zip(zs, amps):
            s = s + a * np.cos(2 * self.k * z * 1e3 + 0.3)
        return self.dc * s

    def counts(self, clean):
        v = clean + self.fixed + 6.0 * self.rng.standard_normal(clean.shape)
        return np.clip(np.round(v), 0, 4095)


def make_case(name: str, system: str) -> dict:
    a = TILED
    rng = np.random.default_rng(a["seed"])
    chirp = chirp_for(system, a["n_lambda"])
    model = Model(chirp_to_lambda(chirp, system), rng)
    tile_x_mm = a["nx"] * a["pixel_um"] * 1e-3
    tile_y_mm = a["ny"] * a["pixel_um"] * 1e-3
    xc = (np.arange(a["tiles_x"]) - (a["tiles_x"] - 1) / 2) * tile_x_mm
    yc = (np.arange(a["tiles_y"]) - (a["tiles_y"] - 1) / 2) * tile_y_mm
    zd = np.asarray(a["depths"], float)
    px = a["pixel_um"] * 1e-3
    frames = {}                       # (xi, zi) -> (ny, interf, N) int16
    for y in yc:                      # tile order y, x, z (scripts/make_synthetic_volume.py:301-306)
        for xi, x in enumerate(xc):
            for zi, z in enumerate(zd):
                files = []
                for yi in range(a["ny"]):
                    y_mm = y + (yi - a["ny"] / 2) * px
                    lines = [model.counts(np.broadcast_to(model.dc, (a["apod"], len(model.dc))))]
                    for xj in range(a["nx"]):
                        x_mm = x + (xj - a["nx"] / 2) * px
                        zs, amps = model.reflectors(x_mm, y_mm, z)
                        lines.append(model.counts(np.broadcast_to(model.aline(zs, amps), (1, len(model.dc)))))
                    files.append(np.concatenate(lines, 0).astype("<i2"))
                frames[(xi, zi)] = np.stack(files)
    z_air_px_um = model.z_nyq_um / (a["n_lambda"] / 2)
    focus_pix = int(round(480.0 / z_air_px_um)) + 1
    # the tile's Chirp.data is float32; the reconstruction derives lambda from that copy
    lam_recon = chirp_to_lambda(chirp.astype("<f4").astype(float), system)
    ref = np.load(REF / "tests" / "data" / "legacy_formats_ref.npz", allow_pickle=False)
    fy = PLANE_Y - 1
    raw = np.stack([frames[(xi, zi)][fy] for xi in range(len(xc)) for zi in range(len(zd))])
    scan = dict(x_centers_mm=xc.tolist(), y_centers_mm=yc.tolist(), z_depths_mm=zd.tolist(),
                tile_range_x_mm=tile_x_mm, tile_range_y_mm=tile_y_mm, n_x_px=a["nx"], n_y_px=a["ny"],
                pixel_size_um=a["pixel_um"], tissue_ri=1.33, apod_size=a["apod"], focus_pix=focus_pix,
                focus_sigma=10.0, dispersion=8.949e7, opc_poly=OPC_POLY, frame=fy,
                tiles=[[xi, zi] for xi in range(len(xc)) for zi in range(len(zd))], system=system)
    np.savez_compressed(OUT / f"legacy_{name}.npz", raw=raw, lambda_nm=lam_recon,
                        plane_db=ref[f"{name}__planeDb"], tie=ref[f"{name}__tie"],
                        scan=np.asarray(json.dumps(scan)))
    return scan


if __name__ == "__main__":
    for n, s in CASES.items():
        info = make_case(n, s)
        print(n, "focus_pix", info["focus_pix"], "->", OUT / f"legacy_{n}.npz")
