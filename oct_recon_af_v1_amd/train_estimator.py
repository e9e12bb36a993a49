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


def build_dataset(xp, n_bscans: int, seed: int, n_x: int = 512, chunk: int = 250, log=print):
    """Inputs X (M, 2, 256, 256) bf16 device, targets and truth for M = n_bscans B-scans."""
    from . import learned as LN
    from . import synthetic as SY
    from .estimation import HipDispersionScorer
    torch = xp.torch
    rng = np.random.default_rng(seed)
    beta_true = rng.uniform(-2e-4, 2e-4, n_bscans)
    focus_true = rng.uniform(0, 1024, n_bscans)
    spec = SY.SpectrumSpec()
    lam = spec.lambda_nm()
    X = torch.empty((n_bscans, 2, LN.IMG, LN.IMG), dtype=torch.bfloat16, device=xp.device)
    keys = ("beta_idx", "score", "focus", "prominence")
    lab = {k: np.zeros(n_bscans) for k in keys}
    raw = torch.empty((chunk, spec.apod + n_x, spec.n), dtype=torch.uint16, device=xp.device)
    scorer = builder = None
    t0 = time.perf_counter()
    tg = tl = 0.0
    for c0 in range(0, n_bscans, chunk):
        nc = min(chunk, n_bscans - c0)
        ta = time.perf_counter()
        r = SY.make_focus_bscans(nc, n_x, beta_true[c0:c0 + nc], focus_true[c0:c0 + nc], spec,
                                 seed=seed * 100003 + c0, device=xp.device, out=raw[:nc])
        xp.synchronize()
        tb = time.perf_counter()
        if scorer is None:
            scorer = HipDispersionScorer(lam, r, spec.apod, xp)
            builder = LN.InputBuilder(scorer)
        else:
            scorer.set_raw(r)
        L = LN.make_labels(scorer)
        for k in keys:
            lab[k][c0:c0 + nc] = L[k]
        builder.build(out=X[c0:c0 + nc])
        xp.synchronize()
        tg += tb - ta
        tl += time.perf_counter() - tb
        if log and (c0 // chunk) % 10 == 0:
            log(f"dataset {c0 + nc}/{n_bscans}: generate {tg:.1f} s, label + inputs {tl:.1f} s")
    lab.update(beta_true=beta_true, focus_true=focus_true, generate_s=tg, label_s=tl,
               elapsed_s=time.perf_counter() - t0)
    return X, lab


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n-bscans", type=int, default=20000)
    ap.add_argument("--n-val", type=int, default=1000)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", required=True)
    ap.add_argument("--metrics", default=None)
    a = ap.parse_args(argv)
    from . import learned as LN
    from .backend import get_xp

    def log(m):
        print(m, flush=True)

    xp = get_xp("hip")[0]
    torch = xp.torch
    X, lab = build_dataset(xp, a.n_bscans + a.n_val, a.seed, log=log)
    ok = np.isfinite(lab["focus"])
    log(f"labels: focus found on {ok.mean():.2%}; |beta_search - beta_true| median "
        f"{np.median(np.abs(lab['beta_idx'] - lab['beta_true'])):.3g}, |focus_det - focus_true| median "
        f"{np.nanmedian(np.abs(lab['focus'] - lab['focus_true'])):.2f} bins")
    # B-scans without a detected focus keep the truth as their target (the detector is the judge
    # of the held-out set, the generator's gate position its stand-in for training)
    focus_t = np.where(ok, lab["focus"], lab["focus_true"])
    Y = torch.tensor(LN.normalise_targets(lab["beta_idx"], focus_t), device=xp.device)
    ntr = a.n_bscans
    net = LN.EstimatorNet(xp, seed=a.seed)
    t0 = time.perf_counter()
    hist = LN.train(net, X[:ntr], Y[:ntr], epochs=a.epochs, batch=a.batch, lr=a.lr, log=log,
                    val=(X[ntr:], Y[ntr:]))
    train_s = time.perf_counter() - t0
    head = LN.predict(net, X[ntr:])
    b, f, sb, sf = LN.denormalise(head)
    vb = lab["beta_idx"][ntr:]
    vf = focus_t[ntr:]
    res = dict(n_train=ntr, n_val=a.n_val, epochs=a.epochs, batch=a.batch, lr=a.lr, train_s=train_s,
               dataset_s=lab["elapsed_s"], generate_s=lab["generate_s"], label_s=lab["label_s"],
               val_beta_abs_err_median=float(np.median(np.abs(b - vb))),
               val_beta_abs_err_p90=float(np.percentile(np.abs(b - vb), 90)),
               val_focus_abs_err_median=float(np.median(np.abs(f - vf))),
               val_focus_abs_err_p90=float(np.percentile(np.abs(f - vf), 90)),
               history=hist, n_params=net.layout.n_params)
    log(json.dumps({k: v for k, v in res.items() if k != "history"}))
    LN.save_params(a.out, net, meta={k: v for k, v in res.items() if k != "history"})
    if a.metrics:
        with open(a.metrics, "w") as fh:
            json.dump(res, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
