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


def profiled_value(name: str, key: str = "traffic_bytes_per_ascan"):
    """A per-A-scan figure recorded in profiles/<name>.md as '<key>: value' (HBM bytes from
    rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated; VALU wave instructions from SQ_INSTS_VALU)
    for this kernel on this workload; None if absent."""
    p = ROOT / "profiles" / f"{name}.md"
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.strip().startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


def operator_row_tiles(nzc: int) -> int:
    """32-row tiles per GEMM workgroup for an operator of nzc bins (oct_opplan_rows)."""
    import ctypes
    from oct_recon_af_v1_amd import _lib
    rt = ctypes.c_int(0)
    _lib.lib().oct_opplan_rows(int(nzc), ctypes.byref(rt))
    return rt.value


def operator_rows(nzc: int) -> int:
    """Rows of the fp16 operator the GEMM applies for nzc bins (oct_opplan_rows)."""
    from oct_recon_af_v1_amd import _lib
    return int(_lib.lib().oct_opplan_rows(int(nzc), None))


# VALU issue peak without packed FP32: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 instruction per 4 cycles
VALU_PEAK_GIPS = 256 * 4 * 2.4 / 4


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


def _host_cores():
    """(threads used, os.cpu_count(), physical cores, CPUs this process may run on)."""
    workers = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    try:
        import psutil
        phys = psutil.cpu_count(logical=False)
    except Exception:                               # psutil missing: unknown
        phys = None
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = None
    return workers, os.cpu_count(), phys, allowed


def cpu_baseline(rec, info, seconds: float, seed: int = 1):
    """The reference CPU path (the oracle's restatement: C pre-FFT with the Numba kernel's
    float64-accumulate / float32-store typing on OpenMP threads, scipy pocketfft with as many
    workers, numpy stitching with one BLAS thread as octrecon/__init__.py:5 sets) timed on this
    host, warm-up pass excluded (as octrecon's scripts/benchmark.py:41-45), in two legs:
      config0  20 B-scans x 500 A-scans of one tile and one focus depth (10 000 A-scans):
               spectral path, 20 log10, clim and uint16 quantisation of the whole set per pass;
      config1  the tiled workload's process_row (octrecon/pipeline.py:223-302) over the first two
               frames of one y-tile row (all 12 tiles: 3 x-tiles x 4 focus depths, 24 000
               A-scans per pass): spectral path of every tile, stitching into num/den, finalise.
    Each leg repeats its pass for about half of `seconds`.  The reported value is config1's."""
    import torch
    from threadpoolctl import threadpool_limits

    from oct_recon_af_v1_amd import synthetic as SY
    from oct_recon_af_v1_amd.geometry import beta_from_index_units
    from oracle import ref_path as R
    workers, logical, phys, allowed = _host_cores()
    cfg, tg = rec.cfg, rec.tg

    def timed(pass_fn, budget):
        pass_fn()                                   # warm-up pass, excluded
        t0 = time.perf_counter()
        done = 0
        while time.perf_counter() - t0 < budget:
            done += pass_fn()
        return done, time.perf_counter() - t0

    # ---- config0: one tile, 20 x 500 A-scans, quantised
    spec = SY.SpectrumSpec()
    lam = spec.lambda_nm()
    beta0 = float(np.random.default_rng(seed + 3).normal(0, 5e-5))
    raw0 = SY.make_bscans(20, 500, spec, beta=beta0, seed=seed + 4).numpy()
    geo0 = R.spectral_geometry(lam, beta_from_index_units(beta0, lam), 1.33)

    def pass0():
        mag = R.magnitude_cpu(raw0, spec.apod, geo0, workers=workers)
        with np.errstate(divide="ignore"):
            db = (20 * np.log10(mag)).astype(np.float32)
        R.db_to_bits(db, R.plane_clim(db))
        return mag.shape[0] * mag.shape[1]

    n0, el0 = timed(pass0, 0.4 * seconds)

    # ---- config1: process_row over two frames of one y-tile row (all tiles of the row)
    geo = R.spectral_geometry(rec.sg.lambda_nm, cfg.dispersion_quadratic_term, cfg.n_medium)
    frames = np.arange(2)
    ids = [(xi, zi) for xi in range(len(rec.scan.x_centers_mm)) for zi in range(len(rec.scan.z_depths_mm))]
    raws = [rec.tiles[(xi, 0, zi)][:len(frames)].cpu().numpy() for xi, zi in ids]
    otg = R.OracleTiled(tg.tile_x_mm, tg.tile_y_mm, tg.tile_z_mm, tg.out_x_mm, tg.out_y_mm, tg.out_z_mm)

    def pass1():
        with threadpool_limits(1, "blas"):
            R.reconstruct_row(raws, ids, frames, geo, otg, rec.scan.x_centers_mm, rec.scan.z_depths_mm,
                              rec.focus, cfg.focus_sigma, rec.sp.raw_apod_size)
        return sum(r.shape[0] * (r.shape[1] - rec.sp.raw_apod_size) for r in raws)

    n1, el1 = timed(pass1, 0.6 * seconds)
    del torch
    host = (f"os.cpu_count()={logical}, physical cores={phys}, CPUs available to the process={allowed}, "
            f"threads used={workers}")
    return {"value": n1 / el1, "unit": "A-scans/s", "cores": workers, "kind": "port",
            "logical_cpus": logical, "physical_cores": phys, "cpus_available": allowed,
            "sample": f"config1 leg: {n1} A-scans (process_row over 2 frames x 12 tiles of one y-tile row, "
                      f"repeated) in {el1:.1f}s; {host}",
            "legs": {
                "config0": {"value": n0 / el0, "unit": "A-scans/s", "ascans": n0, "seconds": el0,
                            "sample": "20 B-scans x 500 A-scans (one tile, one depth): C pre-FFT + scipy pocketfft "
                                      "+ 20 log10 + clim + uint16 quantisation, repeated"},
                "config1": {"value": n1 / el1, "unit": "A-scans/s", "ascans": n1, "seconds": el1,
                            "sample": "tiled workload, one y-tile row (3 x-tiles x 4 depths), frames 0-1: C "
                                      "pre-FFT + scipy pocketfft + numpy stitching (1 BLAS thread) + finalise, "
                                      "repeated"}}}


def time_steps(rec, vol, steps, warmup, busy=None):
    """Warm-up, then `steps` timed whole-job steps; returns (seconds, per-launch kernel ms).
    busy (a list) receives the milliseconds during which at least one of the timed kernel's
    launches ran (their union: the operator form overlaps launches of different focus depths
    on separate streams, so per-launch durations do not add up)."""
    import torch
    events = []
    rec.sp.timing_events = events
    for _ in range(warmup):
        rec.reconstruct_volume(vol)
    torch.cuda.synchronize()
    events.clear()
    ref = torch.cuda.Event(enable_timing=True)
    ref.record()
    t0 = time.perf_counter()
    for _ in range(steps):
        rec.reconstruct_volume(vol)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    rec.sp.timing_events = None
    if busy is not None:
        iv = sorted((ref.elapsed_time(a), ref.elapsed_time(b)) for a, b in events)
        tot, cur0, cur1 = 0.0, None, None
        for a0, a1 in iv:
            if cur1 is None or a0 > cur1:
                if cur1 is not None:
                    tot += cur1 - cur0
                cur0, cur1 = a0, a1
            else:
                cur1 = max(cur1, a1)
        if cur1 is not None:
            tot += cur1 - cur0
        busy.append(tot)
    return el, [a.elapsed_time(b) for a, b in events]


def copy_calibration(gib: float = 1.0):
    """A plain device copy of known size, for scaling FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = int(gib * (1 << 30)) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    log(f"copy calibration: {n * 4} bytes read and written")
    del a, b


def all_bins_form(xp, args):
    """Operator-form precision workload: 200 000 A-scans (200 B-scans x 1000 + 25 background
    lines), reflectivity over 60 dB, all 1024 depth bins; MFMA form vs plain FP32 timing and
    agreement with the float64 oracle on a sample of 2 B-scans."""
    import torch

    from oct_recon_af_v1_amd import synthetic as SY
    from oct_recon_af_v1_amd.geometry import beta_from_index_units, build_spectral_geometry
    from oct_recon_af_v1_amd.spectral import HipSpectralProcessor
    from oracle import ref_path as R
    spec = SY.SpectrumSpec(r_min=1e-7, r_max=1e-1)
    lam = spec.lambda_nm()
    beta = 3e-5
    bref = beta_from_index_units(beta, lam)
    sp = HipSpectralProcessor(build_spectral_geometry(lam, bref, 1.33), spec.apod, xp)
    raw = SY.make_bscans(200, 1000, spec, beta=beta, seed=args.seed + 77, device=xp.device, chunk_files=10)
    out = torch.empty((200, 1000, 1024), dtype=torch.float32, device=xp.device)
    res = {}
    # the matrix-core launches settle after about six (the clock ramps down to its sustained
    # level under this load): at least six untimed launches per form
    warm = max(args.warmup, 6)
    for form in ("mfma", "plain"):
        ev = []
        sp.timing_events = ev
        for it in range(warm + args.steps):
            if it == warm:
                torch.cuda.synchronize()
                ev.clear()
            if form == "mfma":
                sp.run_operator(raw, 0, 1024, out=out)
            else:
                sp.run(raw, 0, 0, 1024, out=out)
        torch.cuda.synchronize()
        sp.timing_events = None
        ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
        res[form] = {"ms": ms, "ascans_per_s": 200000 / (ms * 1e-3)}
        if form == "mfma":
            g = out[:2].cpu().numpy()
    flops = 3 * 2 * 2048 * 2048 * 200000
    res["mfma"]["tflops"] = flops / (res["mfma"]["ms"] * 1e-3) / 1e12
    res["mfma"]["frac_of_fp16_dense_peak"] = res["mfma"]["tflops"] / 2500.0
    h = raw[:2].cpu().numpy()
    o64 = R.magnitude_fp64(h, spec.apod, R.spectral_geometry(lam, bref, 1.33))
    frac, worst, _ = R.lsb_agreement(g, o64)
    with np.errstate(divide="ignore"):
        ddb = np.abs(20 * np.log10(g.astype(np.float64)) - 20 * np.log10(o64))
    ddb = ddb[np.isfinite(ddb)]
    res["mfma"].update({"lsb_within1_vs_fp64": frac, "lsb_max_vs_fp64": worst,
                        "db_err_median": float(np.median(ddb)), "db_err_p999": float(np.quantile(ddb, 0.999)),
                        "sample_ascans": int(g.shape[0] * g.shape[1])})
    res["mfma_vs_plain"] = res["mfma"]["ascans_per_s"] / res["plain"]["ascans_per_s"]
    del raw, out
    torch.cuda.empty_cache()
    return res


def estimator_bench(xp, args):
    """Learned beta / focus estimator on 64 held-out B-scans (512 A-scans each, beta uniform over
    +-2e-4, focus uniform over 0..1023): accuracy against the sharpness search and the confocal
    detector (the referee), and estimation time per tile of 8 B-scans, network vs the restated
    reference search (grid + bounded refinement of the mean and of each B-scan) plus detector."""
    import time

    from oct_recon_af_v1_amd import learned as LN
    from oct_recon_af_v1_amd import synthetic as SY
    from oct_recon_af_v1_amd.estimation import HipDispersionScorer, depth_median, estimate_dispersion
    from oct_recon_af_v1_amd.geometry import beta_from_index_units
    spec = SY.SpectrumSpec()
    lam = spec.lambda_nm()
    rng = np.random.default_rng(args.seed + 500)
    n = 64
    raw = SY.make_focus_bscans(n, 512, rng.uniform(-2e-4, 2e-4, n), rng.uniform(0, 1024, n), spec,
                               seed=args.seed + 501, device=xp.device)
    est = LN.LearnedEstimator(xp, lam, spec.apod)
    p = est.predict(raw)
    lab = LN.make_labels(est.scorer)
    drop = lab["score"] - est.scorer.scores_per_file(p["beta_ref"])
    det = np.isfinite(lab["focus"])
    b_ok = drop <= LN.SCORE_TOL
    f_ok = det & (np.abs(p["focus"] - lab["focus"]) <= LN.FOCUS_TOL_BINS)
    conf = p["confident"]
    tiles = [raw[i:i + 8] for i in range(0, n, 8)]
    sc = HipDispersionScorer(lam, tiles[0], spec.apod, xp)
    lo, hi = beta_from_index_units(-2.5e-4, lam), beta_from_index_units(2.5e-4, lam)

    def search(t):
        sc.set_raw(t)
        sc.n_evals = 0
        r = estimate_dispersion(sc, 0.0, bounds=(lo, hi))
        prof = depth_median(sc.magnitude_at(r["value"]), xp).cpu().numpy().mean(0)
        LN.focus_from_profiles(prof[None])
        return r["n_evaluations"]

    search(tiles[0])
    est.predict(tiles[0], graph=True)
    xp.synchronize()
    t0 = time.perf_counter()
    evals = [search(t) for t in tiles]
    ts = (time.perf_counter() - t0) / len(tiles)
    t0 = time.perf_counter()
    for t in tiles:
        est.predict(t, graph=True)
    xp.synchronize()
    tc = (time.perf_counter() - t0) / len(tiles)
    t0 = time.perf_counter()
    for t in tiles:
        est.predict(t)
    xp.synchronize()
    te = (time.perf_counter() - t0) / len(tiles)
    est.predict(raw, graph=True)
    xp.synchronize()
    t0 = time.perf_counter()
    for _ in range(4):
        est.predict(raw, graph=True)                  # the 8 tiles in one batch
    xp.synchronize()
    tv = (time.perf_counter() - t0) / (4 * len(tiles))
    return {"bscans": n, "beta_within_tol": float(b_ok.mean()), "focus_within_tol": float(f_ok.mean()),
            "beta_abs_err_median_idx": float(np.median(np.abs(p["beta_idx"] - lab["beta_idx"]))),
            "focus_abs_err_median_bins": float(np.nanmedian(np.abs(p["focus"] - lab["focus"]))),
            "accepted": float(conf.mean()), "both_within_tol_accepted": float((b_ok & f_ok)[conf].mean()),
            "tolerance": {"sharpness_drop": LN.SCORE_TOL, "focus_bins": LN.FOCUS_TOL_BINS},
            "tile_ms_search": ts * 1e3, "search_evaluations": float(np.mean(evals)),
            "tile_ms_network_batched": tv * 1e3, "tile_ms_network_single": tc * 1e3,
            "tile_ms_network_single_eager": te * 1e3, "speedup_batched": ts / tv, "speedup_single": ts / tc}


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
    ap.add_argument("--no-mfma", action="store_true", help="skip the operator-form measurements")
    ap.add_argument("--copy-calibration", action="store_true", help="run a 1 GiB device copy first")
    ap.add_argument("--no-all-bins", action="store_true", help="skip the 200 000-A-scan all-bins measurement")
    ap.add_argument("--no-estimator", action="store_true", help="skip the learned-estimator measurement")
    ap.add_argument("--plain-serial", action="store_true",
                    help="plain form: every tile's launch on the caller's stream (no per-depth streams)")
    ap.add_argument("--side-streams", type=int, default=None,
                    help="streams for the focus depths' operator-form launches (default: the pipeline's)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: multi-GPU runs are not part of this build "
              "(the path shards by tile / y-row with no exchange)", file=sys.stderr)
        sys.exit(2)

    import torch

    from oct_recon_af_v1_amd.backend import get_xp
    xp, _ = get_xp("hip")
    if args.copy_calibration:
        copy_calibration()
    rec, info = make_workload(xp, args)
    if args.plain_serial:
        rec.concurrent_forms = ("mfma",)
    vol = rec.allocate_volume()
    ascans = info["ascans"]

    # ---- plain FP32 path (the headline value)
    busy_p = []
    el, ms = time_steps(rec, vol, args.steps, args.warmup, busy_p)
    launches = len(ms)
    avg_ms = float(np.mean(ms))
    per_launch = ascans / (launches / args.steps)
    # the focus depths' launches overlap on their own streams (each launch's tail shares the chip
    # with the next launch), so the kernel's time is the union of the launch intervals
    busy_ms = busy_p[0] / args.steps
    eff_ms = busy_ms / (launches / args.steps)
    # every tile computes its stitcher's row window (wider where the optical-path shift spreads it)
    wins = [rec._tile_rows(st) for st in rec.stitchers.values() if not st.empty]
    lo, hi = min(w[0] for w in wins), max(w[1] for w in wins)
    rows = float(np.mean([w[1] - w[0] for w in wins]))
    bytes_per_ascan = 2 * 2048 + 4 * rows
    achieved = per_launch * bytes_per_ascan / (eff_ms * 1e-3) / 1e9
    value = ascans * args.steps / el
    clim = vol["clim"].cpu().numpy().tolist()
    log(f"plain: step {el / args.steps * 1e3:.2f} ms; spectral launch avg {avg_ms:.3f} ms over {launches} launches "
        f"({per_launch:.0f} A-scans, {rows:.2f} rows on average), busy {busy_ms:.2f} ms per step = {eff_ms:.3f} ms per launch; "
        f"clim {clim}")
    tpa = profiled_value("spectral_kernel_tiled_fp32") if args.spectral_rows == "needed" else None
    ipa = profiled_value("spectral_kernel_tiled_fp32", "valu_instructions_per_ascan") \
        if args.spectral_rows == "needed" else None
    valu = None
    if ipa:
        gips = ipa * per_launch / (eff_ms * 1e-3) / 1e9
        valu = {"achieved": gips, "peak": VALU_PEAK_GIPS, "unit": "G wave-instructions/s", "frac": gips / VALU_PEAK_GIPS,
                "instructions_per_ascan": ipa,
                "source": "profiles/spectral_kernel_tiled_fp32.md (rocprofv3 SQ_INSTS_VALU); peak = non-packed FP32 issue"}
    res = {
        "metric": METRIC, "value": value, "unit": "A-scans/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "tiled plain FP32: 3x3 tiles (10% x overlap) x 4 focus depths, "
                               f"{args.bscans} B-scans x {args.ascans} A-scans x 2048 px uint16 per tile "
                               "(+25 background lines), shared beta, sigma 30 um, n 1.33, 1 um output, "
                               "depth 0-0.15 mm; spectral + stitch + finalise + quantise",
                   "ascans_per_step": ascans, "output_yzx": list(rec.n_out), "spectral_rows": [lo, hi],
                   "spectral_rows_per_tile": rows},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": tpa * per_launch if tpa else None,
                     "traffic_source": "profiles/spectral_kernel_tiled_fp32.md (rocprofv3 FETCH_SIZE x2.0 + "
                                       "WRITE_SIZE x1.0, copy-calibrated)" if tpa else None,
                     "kernel": "spectral16_kernel<uint16_t,12,MAG,STAGE2,4>",
                     "bound_note": "bound by the vector-memory path (TD 72 % busy) and latency, VALU issue next (non-packed FP32); the HBM fraction is reported as asked",
                     "bytes_per_ascan": bytes_per_ascan, "ascans_per_launch": per_launch,
                     "avg_launch_ms": avg_ms, "busy_ms_per_step": busy_ms, "busy_ms_per_launch": eff_ms,
                     "time_note": "achieved = a launch's bytes / (union of the step's launch intervals / launches)",
                     "valu_issue": valu},
    }

    if not args.no_mfma:
        # ---- the same workload through the operator form (MFMA, fp16 3-product)
        bits_plain = vol["bits"].clone()
        clim_plain = vol["clim"].clone()
        rec.cfg.form = "mfma"
        if args.side_streams:
            rec.max_side_streams = args.side_streams
        busy_m = []
        el_m, ms_m = time_steps(rec, vol, args.steps, args.warmup, busy_m)
        rec.cfg.form = "plain"
        ulo, uhi = rec.union_rows()
        bits_m = rec.fin.quantize(vol["db"], clim_plain)
        d = (bits_m.to(torch.int32) - bits_plain.to(torch.int32)).abs()
        within = float((d <= 1).float().mean().item())
        # one operator per focus depth's row window; flops and output rows averaged over tiles
        wins = [rec._tile_rows(st) for st in rec.stitchers.values() if not st.empty]
        rpads = [operator_rows(hi - lo) for lo, hi in wins]
        rpad = float(np.mean(rpads))
        out_rows = float(np.mean([hi - lo for lo, hi in wins]))
        op = rec.sp.operator_plan(*((wins[0][0], wins[0][1] - wins[0][0])))
        per_m = ascans / (len(ms_m) / args.steps)
        avg_m = float(np.mean(ms_m))
        # launches overlap (one stream per focus depth): flops over the time any of them ran
        busy_ms = busy_m[0] / args.steps
        tfl = ascans * 3 * 2 * rpad * 2048 / (busy_ms * 1e-3) / 1e12
        res["mfma_form"] = {
            "value": ascans * args.steps / el_m, "unit": "A-scans/s", "ms_per_step": el_m / args.steps * 1e3,
            "ratio_vs_plain": (ascans * args.steps / el_m) / value,
            "voxels_within_1lsb_of_plain": within, "max_lsb_diff_vs_plain": int(d.max().item()),
            "rows": [ulo, uhi], "rows_per_tile": out_rows, "operator_rows_padded": rpad,
            "operator_scale_log2": op.scale_log2,
            "roofline": {"bound": "mfma", "achieved": tfl, "peak": 2500.0, "unit": "TFLOP/s", "frac": tfl / 2500.0,
                         "kernel": "opgemm_lt_kernel<uint16_t,true,RT>, RT " +
                                   "/".join(str(t) for t in sorted({operator_row_tiles(hi - lo) for lo, hi in wins})),
                         "flops_per_ascan": 3 * 2 * rpad * 2048,
                         "avg_launch_ms": avg_m, "ascans_per_launch": per_m, "busy_ms_per_step": busy_ms,
                         "launches_overlap": True,
                         "hbm_GBps": ascans * (2 * 2048 + 4 * out_rows) / (busy_ms * 1e-3) / 1e9},
        }
        log(f"mfma: step {el_m / args.steps * 1e3:.2f} ms; opgemm avg {avg_m:.3f} ms per launch, busy {busy_ms:.2f} ms per step; "
            f"within 1 LSB of plain {within:.6f}")
        del bits_plain
        if not args.no_all_bins:
            res["all_bins_precision_form"] = all_bins_form(xp, args)
            log(f"all bins: {res['all_bins_precision_form']}")
    if not args.no_estimator:
        res["learned_estimator"] = estimator_bench(xp, args)
        log(f"estimator: {res['learned_estimator']}")
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(rec, info, args.cpu_seconds)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
