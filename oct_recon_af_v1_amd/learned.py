"""Learned dispersion (beta) and focus-depth estimator: a small BF16 CNN trained and run by this
package's own HIP kernels (oct_cnn_* in liboct_hip).

It stands in for the dispersion sharpness search (octrecon/estimation/dispersion.py:154-267) and
the confocal focus detector (octrecon/estimation/focus.py:191-218, 221-279).  Both stay: they make
the training labels, referee the network's error on held-out data, and are the fallback for
B-scans where the network's own predicted uncertainty is above a threshold.

Input: two 256 x 256 log-magnitude images of the central 256 A-scans of a B-scan, reconstructed
with two probe dispersions (beta = -1e-4 and +1e-4 in index units; the residual blur of the two
differs in size, which fixes the sign of beta that a single image leaves open), depth rows
0..1023 max-pooled by 4.  Output: beta, focus depth and the log variance of each (a
heteroscedastic Gaussian head), the variance being the confidence.

Network: 6 x (conv 3x3 stride 2 + ReLU) 2-16-32-32-64-64-64 channels (256 -> 4 pixels), then
fully connected 1024 -> 128 (ReLU) -> 4; 238 k parameters, bf16 weights and activations, float32
master weights, gradients and accumulation, Adam.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np

from . import _lib
from .geometry import beta_from_index_units

ARCH_CONV = ((2, 16), (16, 32), (32, 32), (32, 64), (64, 64), (64, 64))
IMG = 256
FC_HIDDEN = 128
N_TARGETS = 2                   # beta, focus
HEAD = 2 * N_TARGETS            # means, log variances
FLAT = ARCH_CONV[-1][1] * (IMG >> len(ARCH_CONV)) ** 2


@dataclass
class InputSpec:
    n_x_crop: int = IMG
    zpool: int = 4
    n_z: int = 1024
    probes_idx: tuple = (-1e-4, 1e-4)     # probe dispersions, index units (rad per squared k sample)
    relative: bool = False                # probes are offsets from a per-B-scan centre (second stage)
    offset_db: float = 10.0
    scale_db: float = 20.0
    beta_scale: float = 2e-4              # target t_beta = (beta_idx - centre) / beta_scale
    focus_center: float = 512.0           # target t_focus = (z_bin - center) / scale
    focus_scale: float = 256.0


STAGE1 = InputSpec()
# second stage: probes at the first stage's beta +-2e-5, residual within +-3e-5
STAGE2 = InputSpec(probes_idx=(-2e-5, 2e-5), relative=True, beta_scale=3e-5)


class ParamLayout:
    """Flat parameter vector: per layer weight then bias, each slice aligned to 64 elements."""

    def __init__(self):
        self.entries = {}
        off = 0

        def add(name, shape):
            nonlocal off
            self.entries[name] = (tuple(shape), off)
            off += (int(np.prod(shape)) + 63) // 64 * 64

        for i, (ci, co) in enumerate(ARCH_CONV):
            add(f"conv{i}.w", (co, ci, 3, 3))
            add(f"conv{i}.b", (co,))
        add("fc0.w", (FC_HIDDEN, FLAT))
        add("fc0.b", (FC_HIDDEN,))
        add("fc1.w", (HEAD, FC_HIDDEN))
        add("fc1.b", (HEAD,))
        self.total = off
        self.n_params = sum(int(np.prod(s)) for s, _ in self.entries.values())

    def init(self, seed: int = 0) -> np.ndarray:
        """He-uniform weights, zero biases, a small head."""
        rng = np.random.default_rng(seed)
        p = np.zeros(self.total, np.float32)
        for name, (shape, off) in self.entries.items():
            if name.endswith(".w"):
                fan_in = int(np.prod(shape[1:]))
                lim = math.sqrt(6.0 / fan_in) * (0.1 if name == "fc1.w" else 1.0)
                p[off:off + int(np.prod(shape))] = rng.uniform(-lim, lim, int(np.prod(shape)))
        return p

    def unflatten(self, flat) -> dict:
        return {n: np.asarray(flat[o:o + int(np.prod(s))]).reshape(s) for n, (s, o) in self.entries.items()}


def _need(t, numel, what):
    """Host-side extent check before a launch: the kernel addresses exactly `numel` elements."""
    if t.numel() != numel or not t.is_contiguous():
        raise ValueError(f"{what}: tensor {tuple(t.shape)} ({t.numel()} elements), kernel needs {numel}")


class EstimatorNet:
    """Device network state (master float32 weights, bf16 copy, Adam moments) and the forward /
    backward sequences over the oct_cnn_* kernels."""

    def __init__(self, xp, params: np.ndarray | None = None, seed: int = 0):
        self.xp = xp
        torch = self.torch = xp.torch
        self.layout = ParamLayout()
        p = self.layout.init(seed) if params is None else np.asarray(params, np.float32)
        if p.shape != (self.layout.total,):
            raise ValueError(f"parameter vector of {p.shape}, expected ({self.layout.total},)")
        dev = xp.device
        self.p32 = torch.tensor(p, dtype=torch.float32, device=dev)
        self.pbf = self.p32.to(torch.bfloat16)
        self.grad = torch.zeros_like(self.p32)
        self.m = torch.zeros_like(self.p32)
        self.v = torch.zeros_like(self.p32)
        self.steps = 0
        self._bufs = {}
        # training backward convolutions on the matrix cores (oct_cnn_conv_bwd_*_mfma) where the
        # shape allows; False: the FP32-VALU kernels
        self.matrix_cores_backward = True

    # -- slices
    def _slice(self, flat, name):
        shape, off = self.layout.entries[name]
        return flat[off:off + int(np.prod(shape))]

    def _ptr(self, flat, name):
        return flat.data_ptr() + self.layout.entries[name][1] * flat.element_size()

    def params_numpy(self) -> np.ndarray:
        return self.p32.cpu().numpy()

    # -- buffers per batch size
    def _buffers(self, n):
        if n in self._bufs:
            return self._bufs[n]
        torch = self.torch
        dev = self.xp.device
        acts, dpre, ws, mw, md = [], [], [], [], []
        h = IMG
        L = _lib.lib()
        for ci, co in ARCH_CONV:
            acts.append(torch.empty((n, co, h // 2, h // 2), dtype=torch.bfloat16, device=dev))
            dpre.append(torch.empty((n, co, h // 2, h // 2), dtype=torch.float32, device=dev))
            nb = L.oct_cnn_conv_bwd_weight_workspace_bytes(n, ci, h, h, co)
            ws.append(torch.empty((max(1, nb // 4),), dtype=torch.float32, device=dev))
            mw.append(int(L.oct_cnn_conv_bwd_weight_mfma_workspace_bytes(n, ci, h, h, co)))
            md.append(int(L.oct_cnn_conv_bwd_data_mfma_workspace_bytes(n, ci, h, h, co)))
            h //= 2
        # one workspace for the matrix-core backward of every layer (the layers run in stream order)
        mbytes = max(mw + md)
        b = dict(acts=acts, dpre=dpre, ws=ws, mw=mw, md=md,
                 mws=torch.empty((max(1, (mbytes + 3) // 4),), dtype=torch.float32, device=dev),
                 hid=torch.empty((n, FC_HIDDEN), dtype=torch.bfloat16, device=dev),
                 dhid=torch.empty((n, FC_HIDDEN), dtype=torch.float32, device=dev),
                 head=torch.empty((n, HEAD), dtype=torch.float32, device=dev),
                 dhead=torch.empty((n, HEAD), dtype=torch.float32, device=dev),
                 loss=torch.empty((1 + N_TARGETS,), dtype=torch.float32, device=dev))
        self._bufs[n] = b
        return b

    # -- passes
    def forward(self, x, matrix_cores: bool = True):
        """x (n, 2, 256, 256) bf16 device tensor -> head (n, 4) float32 (kept for backward).
        matrix_cores: convolutions through oct_cnn_conv_fwd_mfma (inference); False: the FP32-VALU
        oct_cnn_conv_fwd, which training uses (with the matrix-core forward one of four stage
        trainings of the 20 000-B-scan recipe ended without learning beta; with this one none did)."""
        torch = self.torch
        if x.dtype != torch.bfloat16 or tuple(x.shape[1:]) != (ARCH_CONV[0][0], IMG, IMG) or not x.is_contiguous():
            raise ValueError(f"input must be contiguous bf16 (n, {ARCH_CONV[0][0]}, {IMG}, {IMG}), got "
                             f"{x.dtype} {tuple(x.shape)}")
        n = x.shape[0]
        b = self._buffers(n)
        L = _lib.lib()
        s = self.xp.stream()
        cur, h = x, IMG
        for i, (ci, co) in enumerate(ARCH_CONV):
            _need(cur, n * ci * h * h, f"conv{i} input")
            _need(b["acts"][i], n * co * (h // 2) ** 2, f"conv{i} output")
            conv = L.oct_cnn_conv_fwd_mfma if matrix_cores else L.oct_cnn_conv_fwd
            _lib.check(conv(cur.data_ptr(), self._ptr(self.pbf, f"conv{i}.w"),
                                          self._ptr(self.p32, f"conv{i}.b"), b["acts"][i].data_ptr(), n, ci, h, h, co,
                                          s), f"conv{i} forward")
            cur, h = b["acts"][i], h // 2
        _lib.check(L.oct_cnn_fc_fwd(cur.data_ptr(), self._ptr(self.pbf, "fc0.w"), self._ptr(self.p32, "fc0.b"),
                                    b["hid"].data_ptr(), None, n, FLAT, FC_HIDDEN, s), "fc0 forward")
        _lib.check(L.oct_cnn_fc_fwd(b["hid"].data_ptr(), self._ptr(self.pbf, "fc1.w"), self._ptr(self.p32, "fc1.b"),
                                    None, b["head"].data_ptr(), n, FC_HIDDEN, HEAD, s), "fc1 forward")
        b["x"] = x
        return b["head"]

    def loss(self, n, target, weights):
        """Gaussian NLL of the last forward's head against target (n, 2) float32 device; fills
        dhead; returns the (3,) device tensor (loss, mse_beta, mse_focus)."""
        b = self._buffers(n)
        _lib.check(_lib.lib().oct_cnn_gauss_nll(b["head"].data_ptr(), target.data_ptr(), weights.data_ptr(),
                                                b["loss"].data_ptr(), b["dhead"].data_ptr(), n, N_TARGETS,
                                                self.xp.stream()), "loss")
        return b["loss"]

    def backward(self, n, dhead=None):
        """Gradients of every parameter into self.grad from dhead (default: the loss's)."""
        b = self._buffers(n)
        L = _lib.lib()
        s = self.xp.stream()
        dh = b["dhead"] if dhead is None else dhead
        g = self.grad
        _lib.check(L.oct_cnn_fc_bwd_weight(dh.data_ptr(), b["hid"].data_ptr(), self._ptr(g, "fc1.w"),
                                           self._ptr(g, "fc1.b"), n, FC_HIDDEN, HEAD, s), "fc1 backward (weights)")
        _lib.check(L.oct_cnn_fc_bwd_data(dh.data_ptr(), self._ptr(self.pbf, "fc1.w"), b["hid"].data_ptr(),
                                         b["dhid"].data_ptr(), n, FC_HIDDEN, HEAD, s), "fc1 backward (data)")
        last = len(ARCH_CONV) - 1
        _lib.check(L.oct_cnn_fc_bwd_weight(b["dhid"].data_ptr(), b["acts"][last].data_ptr(), self._ptr(g, "fc0.w"),
                                           self._ptr(g, "fc0.b"), n, FLAT, FC_HIDDEN, s), "fc0 backward (weights)")
        _lib.check(L.oct_cnn_fc_bwd_data(b["dhid"].data_ptr(), self._ptr(self.pbf, "fc0.w"),
                                         b["acts"][last].data_ptr(), b["dpre"][last].data_ptr(), n, FLAT, FC_HIDDEN,
                                         s), "fc0 backward (data)")
        for i in range(last, -1, -1):
            ci, co = ARCH_CONV[i]
            hin = IMG >> i                          # input side of conv i
            xin = b["x"] if i == 0 else b["acts"][i - 1]
            ws = b["ws"][i]
            _need(xin, n * ci * hin * hin, f"conv{i} input")
            _need(b["dpre"][i], n * co * (hin // 2) ** 2, f"conv{i} output gradient")
            if i > 0:
                _need(b["dpre"][i - 1], n * ci * hin * hin, f"conv{i} input gradient")
            mws, mcores = b["mws"], self.matrix_cores_backward
            if mcores and b["mw"][i]:
                _lib.check(L.oct_cnn_conv_bwd_weight_mfma(b["dpre"][i].data_ptr(), xin.data_ptr(),
                                                          self._ptr(g, f"conv{i}.w"), self._ptr(g, f"conv{i}.b"), n, ci,
                                                          hin, hin, co, mws.data_ptr(), mws.numel() * 4, s),
                           f"conv{i} backward (weights, matrix cores)")
            else:
                _lib.check(L.oct_cnn_conv_bwd_weight(b["dpre"][i].data_ptr(), xin.data_ptr(),
                                                     self._ptr(g, f"conv{i}.w"), self._ptr(g, f"conv{i}.b"), n, ci,
                                                     hin, hin, co, ws.data_ptr(), ws.numel() * 4, s),
                           f"conv{i} backward (weights)")
            if i > 0:
                if mcores and b["md"][i]:
                    _lib.check(L.oct_cnn_conv_bwd_data_mfma(b["dpre"][i].data_ptr(), self._ptr(self.pbf, f"conv{i}.w"),
                                                            xin.data_ptr(), b["dpre"][i - 1].data_ptr(), n, ci, hin,
                                                            hin, co, mws.data_ptr(), mws.numel() * 4, s),
                               f"conv{i} backward (data, matrix cores)")
                else:
                    _lib.check(L.oct_cnn_conv_bwd_data(b["dpre"][i].data_ptr(), self._ptr(self.pbf, f"conv{i}.w"),
                                                       xin.data_ptr(), b["dpre"][i - 1].data_ptr(), n, ci, hin, hin,
                                                       co, s), f"conv{i} backward (data)")

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0):
        self.steps += 1
        _lib.check(_lib.lib().oct_cnn_adam(self.p32.data_ptr(), self.grad.data_ptr(), self.m.data_ptr(),
                                           self.v.data_ptr(), self.pbf.data_ptr(), self.layout.total, lr, beta1,
                                           beta2, eps, weight_decay, self.steps, self.xp.stream()), "adam")


# --------------------------------------------------------------------------- inputs and labels
class InputBuilder:
    """Network inputs of raw B-scans: the two probe reconstructions, cropped and log-scaled on the
    device (oct_logmag_crop).  Probes are fixed dispersions, or offsets from a per-B-scan centre
    (spec.relative), reconstructed with per-B-scan windows in one launch per probe."""

    def __init__(self, scorer, spec: InputSpec = STAGE1):
        self.scorer = scorer
        self.spec = spec
        self.lam = scorer.sg0.lambda_nm
        self.to_ref = beta_from_index_units(1.0, self.lam)
        if not spec.relative:
            self.probe_win = scorer.windows([b * self.to_ref for b in spec.probes_idx])

    def build(self, raw=None, out=None, centers_idx=None, centers_head=None, bg=None):
        """(B, 2, 256, 256) bf16 for the scorer's B-scans (or `raw`, which becomes its set): both
        probes in one spectral launch (oct_spectral_repeat) and one crop launch.  A relative spec
        takes the centres (index units) as a host array `centers_idx`, or as `centers_head` =
        (device head (B, 4) float32, beta scale): centre = head[b, 0] * scale, turned into
        windows on the device without a host round trip.  bg: the batch's background rows
        (computed here when not given)."""
        sc, sp = self.scorer, self.spec
        torch = sc.torch
        if raw is not None:
            sc.set_raw(raw)
        if sc.nX < sp.n_x_crop or sp.n_z > sc.nz or sp.n_z // sp.zpool != IMG:
            raise ValueError(f"B-scans of {sc.nX} A-scans x {sc.nz} rows cannot give the {IMG} x {IMG} crop")
        B, P = sc.B, len(sp.probes_idx)
        dev = sc.xp.device
        if bg is None:
            bg = sc.sp.background(sc.raw)
        if sp.relative:
            win = torch.empty((P * B, sc.sg0.n, 2), dtype=torch.float32, device=dev)
            if centers_head is not None:
                if centers_head[0].shape[0] != B:
                    raise ValueError(f"{centers_head[0].shape[0]} heads for {B} B-scans")
                for k, probe in enumerate(sp.probes_idx):
                    sc.windows(head=centers_head[0], head_scale=centers_head[1], offset=probe, to_ref=self.to_ref,
                               out=win[k * B:(k + 1) * B])
            else:
                c = np.asarray(centers_idx, float).reshape(-1)
                if c.size != B:
                    raise ValueError(f"{c.size} centres for {B} B-scans")
                sc.windows(np.concatenate([(c + probe) * self.to_ref for probe in sp.probes_idx]), out=win)
        else:
            win = self.probe_win
        mag = self._mag(P * B)                       # only the cropped scan lines are reconstructed
        x0 = (sc.nX - sp.n_x_crop) // 2
        sc.magnitude_repeat(win, P, bg, mag, x0=x0, nx=sp.n_x_crop)
        out = torch.empty((B, P, IMG, IMG), dtype=torch.bfloat16, device=dev) if out is None else out
        _lib.check(_lib.lib().oct_logmag_crop(mag.data_ptr(), B, sp.n_x_crop, mag.stride(1), sc.nz, 0, 0, sp.zpool,
                                              IMG, sp.n_x_crop, sp.offset_db, sp.scale_db, out.data_ptr(), P,
                                              sc.xp.stream()), "oct_logmag_crop")
        return out

    def _mag(self, n):
        """Probe magnitudes of n images, one buffer per shape kept for the builder's lifetime: a
        HIP graph captured for one batch shape holds this buffer's address, so it must survive
        calls with other shapes (it is never freed or reallocated)."""
        sc = self.scorer
        shape = (n, self.spec.n_x_crop, sc.nz)
        bufs = self.__dict__.setdefault("_magbufs", {})
        buf = bufs.get(shape)
        if buf is None:
            buf = bufs[shape] = sc.torch.empty(shape, dtype=sc.torch.float32, device=sc.xp.device)
        return buf


def focus_from_profiles(profiles, search_px=None, guess=None):
    """Focus (0-based depth bin) and prominence (dB) of each depth profile by the reference's
    confocal detector (octrecon/estimation/focus.py:191-206); nan where no peak is found.
    Without a guess the whole depth range is searched."""
    from .focus import detect_peak_near
    P = np.asarray(profiles, float)
    n = P.shape[-1]
    out = np.full(P.shape[0], np.nan)
    prom = np.zeros(P.shape[0])
    for i, prof in enumerate(P):
        g = n / 2 if guess is None else float(np.atleast_1d(guess)[min(i, np.size(guess) - 1)]) + 1
        pix, pr = detect_peak_near(prof, g, search_px=n if search_px is None else search_px)
        if np.isfinite(pix):
            out[i] = pix - 1
            prom[i] = pr
    return out, prom


def make_labels(scorer, lo_idx=-2.5e-4, hi_idx=2.5e-4, n_grid=41, iters=24):
    """Search labels for the scorer's B-scans: per-B-scan sharpness maximum (beta, index units),
    its score, and the focus detector on the depth profile reconstructed at that beta."""
    from .estimation import depth_median, search_per_bscan
    lam = scorer.sg0.lambda_nm
    to_ref = beta_from_index_units(1.0, lam)
    beta_ref, score, _ = search_per_bscan(scorer, lo_idx * to_ref, hi_idx * to_ref, n_grid, iters)
    mag = scorer.magnitude_at(beta_ref)
    prof = depth_median(mag, scorer.xp).cpu().numpy()
    focus, prom = focus_from_profiles(prof)
    return dict(beta_idx=beta_ref / to_ref, beta_ref=beta_ref, score=score, focus=focus, prominence=prom,
                profiles=prof)


def normalise_targets(beta_idx, focus, spec: InputSpec = STAGE1, center_idx=0.0):
    return np.stack([(np.asarray(beta_idx) - np.asarray(center_idx)) / spec.beta_scale,
                     (np.asarray(focus) - spec.focus_center) / spec.focus_scale], -1).astype(np.float32)


def denormalise(head, spec: InputSpec = STAGE1, center_idx=0.0):
    """head (n, 4) numpy -> beta_idx, focus bin, sigma_beta_idx, sigma_focus (bins)."""
    h = np.asarray(head, float)
    beta = np.asarray(center_idx, float) + h[:, 0] * spec.beta_scale
    focus = h[:, 1] * spec.focus_scale + spec.focus_center
    sb = np.exp(0.5 * h[:, 2]) * spec.beta_scale
    sf = np.exp(0.5 * h[:, 3]) * spec.focus_scale
    return beta, focus, sb, sf


# --------------------------------------------------------------------------- training
def train(net: EstimatorNet, X, Y, epochs: int = 20, batch: int = 64, lr: float = 2e-3, weight_decay: float = 1e-4,
          weights=(1.0, 1.0), warmup: int = 200, seed: int = 0, log=None, val=None):
    """Adam with linear warm-up and cosine decay over `epochs` passes of (X, Y) device tensors
    (X (M, 2, 256, 256) bf16, Y (M, 2) float32 normalised targets); drops the last partial
    batch.  Returns the per-epoch history (train loss, validation MSEs)."""
    torch = net.torch
    dev = net.xp.device
    M = X.shape[0]
    steps_per_epoch = M // batch
    if steps_per_epoch < 1:
        raise ValueError(f"{M} samples for batches of {batch}")
    total = epochs * steps_per_epoch
    w = torch.tensor(weights, dtype=torch.float32, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))
    xb = torch.empty((batch,) + tuple(X.shape[1:]), dtype=X.dtype, device=dev)
    yb = torch.empty((batch, Y.shape[1]), dtype=Y.dtype, device=dev)
    hist = []
    t0 = time.perf_counter()
    step = 0
    for ep in range(epochs):
        perm = torch.randperm(M, generator=gen, device=dev)
        acc = torch.zeros((1 + N_TARGETS,), dtype=torch.float32, device=dev)
        for k in range(steps_per_epoch):
            idx = perm[k * batch:(k + 1) * batch]
            torch.index_select(X, 0, idx, out=xb)
            torch.index_select(Y, 0, idx, out=yb)
            net.forward(xb, matrix_cores=False)
            acc += net.loss(batch, yb, w)
            net.backward(batch)
            f = min(1.0, (step + 1) / warmup) * 0.5 * (1 + math.cos(math.pi * step / total))
            net.adam(lr * f, weight_decay=weight_decay)
            step += 1
        rec = dict(epoch=ep, train=(acc / steps_per_epoch).cpu().numpy().tolist(),
                   elapsed_s=time.perf_counter() - t0)
        if val is not None:
            rec["val_mse"] = evaluate(net, val[0], val[1], batch)
        hist.append(rec)
        if log:
            log(f"epoch {ep}: {rec}")
    return hist


def predict(net: EstimatorNet, X, batch: int = 256):
    """head (M, 4) numpy for X (M, 2, 256, 256) bf16 device."""
    torch = net.torch
    out = []
    for i in range(0, X.shape[0], batch):
        xb = X[i:i + batch].contiguous()
        out.append(net.forward(xb).clone())
    return torch.cat(out).cpu().numpy()


def evaluate(net, X, Y, batch=256):
    h = predict(net, X, batch)
    y = Y.cpu().numpy() if hasattr(Y, "cpu") else np.asarray(Y)
    return np.mean((h[:, :N_TARGETS] - y) ** 2, 0).tolist()


def save_params(path, stages, meta=None):
    """Trained weights of the stages (EstimatorNet or flat float32 vectors) as an .npz of plain
    float32 arrays plus JSON metadata."""
    import json
    stages = stages if isinstance(stages, (list, tuple)) else [stages]
    arrays = {f"params{i}": (st.params_numpy() if isinstance(st, EstimatorNet) else np.asarray(st, np.float32))
              for i, st in enumerate(stages)}
    np.savez_compressed(path, meta=np.array(json.dumps(dict(meta or {}, n_params=ParamLayout().n_params,
                                                            stages=len(stages)))), **arrays)


def load_params(path):
    """([stage parameter vectors], meta)."""
    import json
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(str(z["meta"]))
        return [z[f"params{i}"].astype(np.float32) for i in range(meta.get("stages", 1))], meta


def calibrate_gate(err, sigma, tol, target=0.99):
    """Largest predicted-sigma gate such that, of the samples it accepts, at least `target` have
    |err| <= tol.  Returns (gate, accepted fraction, precision)."""
    err, sigma = np.abs(np.asarray(err, float)), np.asarray(sigma, float)
    order = np.argsort(sigma, kind="stable")
    ok = err[order] <= tol
    prec = np.cumsum(ok) / np.arange(1, len(ok) + 1)
    good = np.nonzero(prec >= target)[0]
    if good.size == 0:
        return float(sigma.min() * 0.5), 0.0, float("nan")
    k = int(good.max())
    return float(sigma[order][k]), (k + 1) / len(ok), float(prec[k])


# --------------------------------------------------------------------------- estimator with fallback
DEFAULT_PARAMS = __import__("pathlib").Path(__file__).resolve().parent / "data" / "estimator_bf16.npz"
SCORE_TOL = 0.05          # referee tolerance: sharpness (log top-1 % intensity) within 0.05 of the optimum
BETA_TOL_IDX = 5.5e-6     # |beta error| (index units) at which that drop is reached on this spectrometer
FOCUS_TOL_BINS = 8.8      # referee tolerance: the stitching focus weight's standard deviation (sqrt(2) x 30 um / dz)


def take_bscans(raw, idx):
    """raw[idx] along the first axis for a device uint16 tensor (gathered through an int16 view)."""
    torch = __import__("torch")
    i = torch.as_tensor(np.asarray(idx, np.int64), device=raw.device)
    if raw.dtype == torch.uint16:
        return torch.index_select(raw.view(torch.int16), 0, i).view(torch.uint16)
    return torch.index_select(raw, 0, i)


class LearnedEstimator:
    """beta (index and reference units) and focus depth of B-scans from the two-stage network, with
    the sharpness search / confocal detector as fallback where the predicted sigma exceeds the gate.

    Stage 1 sees probes at beta = -+1e-4 and predicts beta over +-2e-4; stage 2 sees probes at the
    stage-1 beta -+2e-5 and predicts the residual (and the focus, from the sharper images).
    gate_beta / gate_focus: largest accepted stage-2 sigma (index units / depth bins); defaults
    come from the calibration stored with the weights."""

    def __init__(self, xp, lambda_nm, apod_size: int, params=None, gate_beta=None, gate_focus=None,
                 n_medium: float = 1.33):
        from .estimation import HipDispersionScorer
        meta = {}
        if params is None or isinstance(params, (str, bytes, __import__("os").PathLike)):
            params, meta = load_params(params or DEFAULT_PARAMS)
        if len(params) != 2:
            raise ValueError(f"{len(params)} parameter sets, the estimator has 2 stages")
        self.xp = xp
        self.nets = [EstimatorNet(xp, p) for p in params]
        self.specs = (STAGE1, STAGE2)
        self.lambda_nm = np.asarray(lambda_nm, float)
        self.apod = apod_size
        self.n_medium = n_medium
        cal = meta.get("calibration", {})
        self.gate_beta = gate_beta if gate_beta is not None else cal.get("gate_beta", 2e-6)
        self.gate_focus = gate_focus if gate_focus is not None else cal.get("gate_focus", FOCUS_TOL_BINS)
        self._scorer_cls = HipDispersionScorer
        self.scorer = None
        self.builders = None
        self._graphs = {}

    def _bind(self, raw):
        if self.scorer is None:
            self.scorer = self._scorer_cls(self.lambda_nm, raw, self.apod, self.xp, n_medium=self.n_medium)
            self.builders = [InputBuilder(self.scorer, sp) for sp in self.specs]
        else:
            self.scorer.set_raw(raw)

    def _heads(self):
        """Both stages on the bound batch, entirely on the device (no host round trip)."""
        bg = self.scorer.sp.background(self.scorer.raw)
        h1 = self.nets[0].forward(self.builders[0].build(bg=bg))
        h2 = self.nets[1].forward(self.builders[1].build(centers_head=(h1, self.specs[0].beta_scale), bg=bg))
        return self.xp.torch.stack([h1, h2])

    def predict(self, raw, graph: bool = False):
        """Network only: per B-scan beta_idx, beta_ref, focus (0-based bin), sigmas, confident
        (plus the stage-1 beta).  graph=True replays the whole two-stage sequence as one HIP
        graph captured per batch shape (the batch is copied into the graph's input buffer)."""
        torch = self.xp.torch
        if graph:
            key = (tuple(raw.shape), raw.dtype)
            if key not in self._graphs:
                static = raw.clone()
                self._bind(static)
                self._heads()                                  # allocates every buffer
                self.xp.synchronize()
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    heads = self._heads()
                self._graphs[key] = (g, static, heads)
            g, static, heads = self._graphs[key]
            static.copy_(raw)
            self._bind(static)
            g.replay()
            H = heads.cpu().numpy()
        else:
            self._bind(raw)
            H = self._heads().cpu().numpy()                    # the one host synchronisation
        b1 = H[0][:, 0].astype(np.float64) * self.specs[0].beta_scale   # as the window kernel forms it
        b, f, sb, sf = denormalise(H[1], self.specs[1], center_idx=b1)
        return dict(beta_idx=b, beta_ref=beta_from_index_units(b, self.lambda_nm), focus=f, sigma_beta=sb,
                    sigma_focus=sf, beta_stage1=b1,
                    confident=(sb <= self.gate_beta) & (sf <= self.gate_focus) & (np.abs(b - b1) < self.specs[1].beta_scale))

    def estimate(self, raw):
        """Per B-scan estimates with fallback: B-scans the network is not confident about get the
        per-B-scan sharpness search and the confocal detector (the labels' procedure)."""
        p = self.predict(raw)
        p["source"] = np.where(p["confident"], "cnn", "search")
        bad = np.nonzero(~p["confident"])[0]
        if bad.size:
            raw_bad = take_bscans(raw, bad) if len(bad) < raw.shape[0] else raw
            self.scorer.set_raw(raw_bad)
            lab = make_labels(self.scorer)
            p["beta_idx"][bad] = lab["beta_idx"]
            p["beta_ref"][bad] = lab["beta_ref"]
            p["focus"][bad] = np.where(np.isfinite(lab["focus"]), lab["focus"], p["focus"][bad])
        p["fallback_rate"] = float(bad.size / len(p["confident"]))
        return p


def _time_main(argv=None):
    """Time the estimator on synthetic B-scans (one tile at a time and all tiles in one batch):
    python -m oct_recon_af_v1_amd.learned --tiles 8 --bscans 8 --reps 3"""
    import argparse
    from . import synthetic as SY
    from .backend import get_xp
    ap = argparse.ArgumentParser()
    ap.add_argument("--tiles", type=int, default=8)
    ap.add_argument("--bscans", type=int, default=8)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--seed", type=int, default=5)
    a = ap.parse_args(argv)
    xp = get_xp("hip")[0]
    spec = SY.SpectrumSpec()
    rng = np.random.default_rng(a.seed)
    n = a.tiles * a.bscans
    raw = SY.make_focus_bscans(n, 512, rng.uniform(-2e-4, 2e-4, n), rng.uniform(0, 1024, n), spec, seed=a.seed,
                               device=xp.device)
    est = LearnedEstimator(xp, spec.lambda_nm(), spec.apod)
    est.predict(raw[:a.bscans], graph=True)
    est.predict(raw, graph=True)
    xp.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.reps):
        for i in range(0, n, a.bscans):
            est.predict(raw[i:i + a.bscans], graph=True)
    xp.synchronize()
    t1 = time.perf_counter()
    for _ in range(a.reps):
        est.predict(raw, graph=True)
    xp.synchronize()
    t2 = time.perf_counter()
    print(f"per tile of {a.bscans} B-scans: {(t1 - t0) / (a.reps * a.tiles) * 1e3:.3f} ms one tile at a time, "
          f"{(t2 - t1) / (a.reps * a.tiles) * 1e3:.3f} ms with {a.tiles} tiles in one batch", flush=True)


if __name__ == "__main__":
    _time_main()
