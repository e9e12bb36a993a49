"""Train the learned beta / focus estimator on the GPU.

Workload: synthetic B-scans of 512 A-scans x 2048 px uint16 (the config-0 spectrometer and
reflectors, plus a scattering medium under a confocal gate), beta uniform over +-2e-4 (index
units) and focus depth uniform over depth bins 0..1023, each reconstructed by this package's
spectral kernel at the two probe dispersions and cropped to the network input.  Labels: each
B-scan's own sharpness-search maximum (beta) and the confocal detector on the depth profile
reconstructed at that beta (focus), both on the device.

    python -m oct_recon_af_v1_amd.train_estimator --n-bscans 20000 --out estimator.npz
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def build_dataset(xp, n_bscans: int, seed: int, n_x: int = 512, chunk: int = 250, keep_raw: bool = False,
                  log=print, host_every: int = 0):
    """Stage-1 and stage-2 inputs X1, X2 (M, 2, 256, 256) bf16 device, labels and truth for
    M = n_bscans B-scans; stage-2 probes are centred at the label plus a uniform +-3e-5 offset
    (the spread of stage-1 errors it corrects).  keep_raw: also return the raw B-scans.
    host_every = k > 0: every k-th chunk comes from the host tile generator (the streamed
    volume's source: bank-built speckle, one generator call per B-scan) instead of the device
    generator, so the network sees both."""
    from . import learned as LN
    from . import synthetic as SY
    from .estimation import HipDispersionScorer
    torch = xp.torch
    rng = np.random.default_rng(seed)
    beta_true = rng.uniform(-2e-4, 2e-4, n_bscans)
    focus_true = rng.uniform(0, 1024, n_bscans)
    offset2 = rng.uniform(-1, 1, n_bscans) * LN.STAGE2.beta_scale
    spec = SY.SpectrumSpec()
    lam = spec.lambda_nm()
    X1 = torch.empty((n_bscans, 2, LN.IMG, LN.IMG), dtype=torch.bfloat16, device=xp.device)
    X2 = torch.empty_like(X1)
    keys = ("beta_idx", "score", "focus", "prominence")
    lab = {k: np.zeros(n_bscans) for k in keys}
    if keep_raw:
        raw_all = torch.empty((n_bscans, spec.apod + n_x, spec.n), dtype=torch.uint16, device=xp.device)
    else:
        raw = torch.empty((chunk, spec.apod + n_x, spec.n), dtype=torch.uint16, device=xp.device)
    scorer = b1 = b2 = None
    hostgen = hbuf = None
    t0 = time.perf_counter()
    tg = tl = 0.0
    for c0 in range(0, n_bscans, chunk):
        nc = min(chunk, n_bscans - c0)
        sl = slice(c0, c0 + nc)
        ta = time.perf_counter()
        dst = raw_all[sl] if keep_raw else raw[:nc]
        if host_every and (c0 // chunk) % host_every == host_every - 1:
            if hostgen is None:
                hostgen = SY.HostTileGenerator(spec)
                hbuf = torch.empty((chunk, spec.apod + n_x, spec.n), dtype=torch.uint16, pin_memory=True)
            for k in range(nc):
                hostgen.generate(hbuf[k:k + 1], n_x, beta_true[c0 + k], focus_true[c0 + k], seed * 100003 + c0 + k)
            dst.copy_(hbuf[:nc])
            r = dst
        else:
            r = SY.make_focus_bscans(nc, n_x, beta_true[sl], focus_true[sl], spec, seed=seed * 100003 + c0,
                                     device=xp.device, out=dst)
        xp.synchronize()
        tb = time.perf_counter()
        if scorer is None:
            scorer = HipDispersionScorer(lam, r, spec.apod, xp)
            b1, b2 = LN.InputBuilder(scorer, LN.STAGE1), LN.InputBuilder(scorer, LN.STAGE2)
        else:
            scorer.set_raw(r)
        L = LN.make_labels(scorer)
        for k in keys:
            lab[k][sl] = L[k]
        b1.build(out=X1[sl])
        b2.build(out=X2[sl], centers_idx=L["beta_idx"] + offset2[sl])
        xp.synchronize()
        tg += tb - ta
        tl += time.perf_counter() - tb
        if log and (c0 // chunk) % 10 == 0:
            log(f"dataset {c0 + nc}/{n_bscans}: generate {tg:.1f} s, label + inputs {tl:.1f} s")
    lab.update(beta_true=beta_true, focus_true=focus_true, center2=lab["beta_idx"] + offset2, generate_s=tg,
               label_s=tl, elapsed_s=time.perf_counter() - t0)
    return X1, X2, lab, (raw_all if keep_raw else None)


def calibrate(xp, params, vlab, vraw, chunk=250):
    """The two-stage cascade on validation B-scans (stage 2 centred at stage 1's own estimate):
    errors against the search labels and the gates (largest predicted sigma keeping >= 99 % of
    the accepted B-scans within tolerance)."""
    from . import learned as LN
    from . import synthetic as SY
    spec = SY.SpectrumSpec()
    est = LN.LearnedEstimator(xp, spec.lambda_nm(), spec.apod, params=params, gate_beta=np.inf, gate_focus=np.inf)
    n = vraw.shape[0]
    preds = [est.predict(vraw[i:i + chunk]) for i in range(0, n, chunk)]
    cat = {k: np.concatenate([p_[k] for p_ in preds]) for k in ("beta_idx", "beta_stage1", "focus", "sigma_beta",
                                                                 "sigma_focus")}
    b, b1, f, sb, sf = (cat[k] for k in ("beta_idx", "beta_stage1", "focus", "sigma_beta", "sigma_focus"))
    vb = vlab["beta_idx"]
    vf = np.where(np.isfinite(vlab["focus"]), vlab["focus"], vlab["focus_true"])
    gb, cb, pb = LN.calibrate_gate(b - vb, sb, LN.BETA_TOL_IDX)
    gf, cf, pf = LN.calibrate_gate(f - vf, sf, LN.FOCUS_TOL_BINS)
    both = (sb <= gb) & (sf <= gf)
    within = (np.abs(b - vb) <= LN.BETA_TOL_IDX) & (np.abs(f - vf) <= LN.FOCUS_TOL_BINS)
    return dict(val_stage1_beta_abs_err_median=float(np.median(np.abs(b1 - vb))),
                val_beta_abs_err_median=float(np.median(np.abs(b - vb))),
                val_beta_abs_err_p90=float(np.percentile(np.abs(b - vb), 90)),
                val_focus_abs_err_median=float(np.median(np.abs(f - vf))),
                val_focus_abs_err_p90=float(np.percentile(np.abs(f - vf), 90)),
                calibration=dict(gate_beta=gb, gate_focus=gf, beta_tol_idx=LN.BETA_TOL_IDX,
                                 focus_tol_bins=LN.FOCUS_TOL_BINS, accepted_beta=cb, precision_beta=pb,
                                 accepted_focus=cf, precision_focus=pf, accepted_both=float(both.mean()),
                                 within_tol_all=float(within.mean()),
                                 within_tol_both_accepted=float(within[both].mean()) if both.any() else float("nan")))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n-bscans", type=int, default=20000)
    ap.add_argument("--n-val", type=int, default=1000)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--host-every", type=int, default=0,
                    help="every k-th chunk of B-scans from the host tile generator (0: none)")
    ap.add_argument("--out", required=True)
    ap.add_argument("--metrics", default=None)
    ap.add_argument("--calibrate-only", default=None, metavar="WEIGHTS",
                    help="recalibrate the gates of trained weights on a fresh validation set and write --out")
    a = ap.parse_args(argv)
    from . import learned as LN
    from . import synthetic as SY
    from .backend import get_xp

    def log(m):
        print(m, flush=True)

    xp = get_xp("hip")[0]
    torch = xp.torch
    if a.calibrate_only:
        params, meta = LN.load_params(a.calibrate_only)
        _, _, vlab, vraw = build_dataset(xp, a.n_val, a.seed + 1, keep_raw=True, log=log, host_every=a.host_every)
        meta.update(calibrate(xp, params, vlab, vraw))
        log(json.dumps(meta))
        LN.save_params(a.out, params, meta=meta)
        return 0
    X1, X2, lab, _ = build_dataset(xp, a.n_bscans, a.seed, log=log, host_every=a.host_every)
    V1, V2, vlab, vraw = build_dataset(xp, a.n_val, a.seed + 1, keep_raw=True, log=log, host_every=a.host_every)
    ok = np.isfinite(lab["focus"])
    log(f"labels: focus found on {ok.mean():.2%}; |beta_search - beta_true| median "
        f"{np.median(np.abs(lab['beta_idx'] - lab['beta_true'])):.3g}, |focus_det - focus_true| median "
        f"{np.nanmedian(np.abs(lab['focus'] - lab['focus_true'])):.2f} bins")
    # B-scans without a detected focus keep the truth as their target (the detector is the referee
    # of the held-out set, the generator's gate position its stand-in for training)
    focus_t = np.where(ok, lab["focus"], lab["focus_true"])
    vfocus_t = np.where(np.isfinite(vlab["focus"]), vlab["focus"], vlab["focus_true"])
    res = dict(n_train=a.n_bscans, n_val=a.n_val, epochs=a.epochs, batch=a.batch, lr=a.lr, host_every=a.host_every,
               dataset_s=lab["elapsed_s"], generate_s=lab["generate_s"], label_s=lab["label_s"], stages=[])
    nets = []
    for st, (X, V, spec, centre, vcentre) in enumerate(((X1, V1, LN.STAGE1, 0.0, 0.0),
                                                        (X2, V2, LN.STAGE2, lab["center2"], vlab["center2"]))):
        Y = torch.tensor(LN.normalise_targets(lab["beta_idx"], focus_t, spec, centre), device=xp.device)
        VY = torch.tensor(LN.normalise_targets(vlab["beta_idx"], vfocus_t, spec, vcentre), device=xp.device)
        net = LN.EstimatorNet(xp, seed=a.seed + st)
        t0 = time.perf_counter()
        hist = LN.train(net, X, Y, epochs=a.epochs, batch=a.batch, lr=a.lr, log=log, val=(V, VY))
        res["stages"].append(dict(train_s=time.perf_counter() - t0, history=hist))
        nets.append(net)
    res["train_s"] = sum(s_["train_s"] for s_ in res["stages"])
    del X1, X2
    res.update(calibrate(xp, [n.params_numpy() for n in nets], vlab, vraw))
    res["n_params"] = nets[0].layout.n_params
    log(json.dumps({k: v for k, v in res.items() if k != "stages"}))
    LN.save_params(a.out, nets, meta={k: v for k, v in res.items() if k != "stages"})
    if a.metrics:
        with open(a.metrics, "w") as fh:
            json.dump(res, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
