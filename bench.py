#!/usr/bin/env python
"""Benchmark: A-scans/s of the tiled OCT reconstruction on one MI355X.

Workload (default, the tiled GPU configuration): 3 x 3 tiles with 10 % lateral (x) overlap x 4
focus depths; each tile 100 B-scans x 1000 A-scans x 2048 px uint16 12-bit counts (+25
background lines per B-scan), one shared dispersion beta; focus planes 50 um apart, Gaussian
focus weight sigma = 30 um, n = 1.33, 1 um output pixels, output depth 0-0.15 mm.  Inputs are
generated on the device and resident in HBM before timing.  A step = the whole job: fused
spectral kernel for all 36 tiles (3.6 M A-scans), stitching into one output volume, finalise
(dB, plane/global clim) and legacy uint16 quantisation.

Prints one JSON line.  --gpus N > 1 is not supported (one process, one GPU).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "A-scans/s on 1 MI355X: plain FP32 vs MFMA form at ≤1 LSB; %HBM peak"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(xp, args):
    """Config: returns (Reconstructor, info)."""
    import torch

    from oct_recon_af_v1_amd import synthetic as SY
    from oct_recon_af_v1_amd.geometry import ScanGrid, beta_from_index_units, build_spectral_geometry
    from oct_recon_af_v1_amd.pipeline import ReconConfig, Reconstructor

    spec = SY.SpectrumSpec()
    lam = spec.lambda_nm()
    n_tx, n_ty, n_z = 3, 3, 4
    nx, ny = args.ascans, args.bscans
    px_mm = 1e-3
    tile_x = nx * px_mm
    tile_y = ny * px_mm
    xc = (np.arange(n_tx) - (n_tx - 1) / 2) * tile_x * 0.9          # 10 % overlap in x
    yc = (np.arange(n_ty) - (n_ty - 1) / 2) * tile_y                # no overlap in y
    zd = np.arange(n_z) * 0.05                                       # focus planes 50 um apart
    beta = float(np.random.default_rng(args.seed).normal(0, 5e-5))
    sg = build_spectral_geometry(lam, beta_from_index_units(beta, lam), 1.33)
    dz_um = float(sg.z_um[1] - sg.z_um[0])
    grid = ScanGrid(xc, yc, zd, tile_x, tile_y, nx, ny, pixel_size_um=1.0)
    cfg = ReconConfig(dispersion_quadratic_term=beta_from_index_units(beta, lam), focus_positions=512.0,
                      focus_sigma=30.0 / dz_um, crop_z_range_mm=[0.0, 0.15], output_pixel_size_um=1.0,
                      n_medium=1.33, batch_frames=ny, spectral_rows=args.spectral_rows)
    tiles = {}
    t0 = time.perf_counter()
    k = 0
    for yi in range(n_ty):
        for xi in range(n_tx):
            for zi in range(n_z):
                tiles[(xi, yi, zi)] = SY.make_bscans(ny, nx, spec, beta=beta, seed=args.seed * 1000 + k,
                                                     device=xp.device, chunk_files=10)
                k += 1
    torch.cuda.synchronize()
    log(f"generated {k} tiles ({k * ny * nx / 1e6:.2f} M A-scans, "
        f"{sum(t.numel() * 2 for t in tiles.values()) / 1e9:.1f} GB) in {time.perf_counter() - t0:.1f}s")
    rec = Reconstructor(cfg, grid, lam, tiles, spec.apod, xp=xp)
    info = {"tiles": k, "ascans": k * ny * nx, "beta_idx": beta, "lines": spec.apod + nx,
            "out_shape": rec.n_out, "dz_um": dz_um}
    return rec, info


def cpu_baseline(rec, info, seconds: float):
    """The reference CPU path (oracle restatement: C pre-FFT with the Numba typing + scipy
    pocketfft + numpy stitching) on a bounded sample of the same workload: the first 10
    B-scans of one tile through spectral + stitch, repeated for ~`seconds`."""
    import torch

    from oracle import ref_path as R
    workers = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cfg = rec.cfg
    tg = rec.tg
    geo = R.spectral_geometry(rec.sg.lambda_nm, cfg.dispersion_quadratic_term, cfg.n_medium)
    nb = 10
    raw = rec.tiles[(0, 0, 0)][:nb].cpu().numpy()
    st = R.Stitcher(tg.tile_x_mm, tg.tile_z_mm, tg.out_x_mm, tg.out_z_mm, None, rec.scan.x_centers_mm[0],
                    rec.scan.z_depths_mm[0], rec.focus[0], cfg.focus_sigma, np.float32)
    frames = np.arange(nb)
    num = np.zeros((nb, len(tg.out_z_mm), len(tg.out_x_mm)), np.float32)
    den = np.zeros_like(num)

    def once():
        mag = R.magnitude_cpu(raw, 25, geo, workers=workers)
        n_, d_ = st.contribute(mag, frames)
        num[:, :, st.c0:st.c1] += n_
        den[:, :, st.c0:st.c1] += d_
        return mag.shape[0] * mag.shape[1]

    once()                                      # warm-up pass, excluded
    t0 = time.perf_counter()
    done = 0
    while time.perf_counter() - t0 < seconds:
        done += once()
    el = time.perf_counter() - t0
    del torch
    return {"value": done / el, "unit": "A-scans/s", "cores": workers, "kind": "port",
            "sample": f"{done} A-scans ({nb} B-scans x {raw.shape[1] - 25} A-scans of one tile, repeated) "
                      f"through the oracle's C pre-FFT (Numba typing, OpenMP) + scipy pocketfft + numpy "
                      f"stitching in {el:.1f}s; os.cpu_count()={os.cpu_count()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--bscans", type=int, default=100)
    ap.add_argument("--ascans", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--spectral-rows", default="needed", choices=["needed", "all"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: multi-GPU runs are not part of this build "
              "(the path shards by tile / y-row with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch

    from oct_recon_af_v1_amd.backend import get_xp
    xp, _ = get_xp("hip")
    rec, info = make_workload(xp, args)
    vol = rec.allocate_volume()

    # per-launch HIP events around the fused spectral kernel, on the stream it runs on
    spectral_events = []
    rec.sp.timing_events = spectral_events

    for _ in range(args.warmup):
        rec.reconstruct_volume(vol)
    torch.cuda.synchronize()
    spectral_events.clear()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rec.reconstruct_volume(vol)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    rec.sp.timing_events = None
    ms = [a.elapsed_time(b) for a, b in spectral_events]
    launches = len(ms)
    avg_ms = float(np.mean(ms))
    ascans = info["ascans"]
    per_launch = ascans / (launches / args.steps)
    lo, hi = rec.stitchers[(0, 0)].needed_rows() if args.spectral_rows == "needed" else (0, 1024)
    rows = hi - lo
    bytes_per_ascan = 2 * 2048 + 4 * rows
    achieved = per_launch * bytes_per_ascan / (avg_ms * 1e-3) / 1e9
    value = ascans * args.steps / el
    clim = vol["clim"].cpu().numpy().tolist()
    log(f"step {el / args.steps * 1e3:.2f} ms; spectral launch avg {avg_ms:.3f} ms over {launches} launches "
        f"({per_launch:.0f} A-scans, {rows} rows); clim {clim}")
    res = {
        "metric": METRIC, "value": value, "unit": "A-scans/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "tiled plain FP32: 3x3 tiles (10% x overlap) x 4 focus depths, "
                               f"{args.bscans} B-scans x {args.ascans} A-scans x 2048 px uint16 per tile "
                               "(+25 background lines), shared beta, sigma 30 um, n 1.33, 1 um output, "
                               "depth 0-0.15 mm; spectral + stitch + finalise + quantise",
                   "ascans_per_step": ascans, "output_yzx": list(rec.n_out), "spectral_rows": [lo, hi]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "spectral_kernel<uint16_t,12,MAG>",
                     "bytes_per_ascan": bytes_per_ascan, "ascans_per_launch": per_launch,
                     "avg_launch_ms": avg_ms},
    }
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(rec, info, args.cpu_seconds)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
