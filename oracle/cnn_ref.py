"""CPU restatement of the learned estimator's network, its gradients and its optimiser.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nothing in the product path).  The network has
no counterpart in the reference; what the reference fixes here is only what the network is
compared with (octrecon/estimation/dispersion.py:46-63, 154-267; focus.py:44-72, 191-218).

Everything is float64 numpy.  With `bf16=True` the forward pass rounds what the device stores in
bfloat16 (weights, the input, every ReLU output) exactly as the device does (round to nearest
even), so the device and this restatement differ only by float32 accumulation order.
Pinned in tests/test_cnn_oracle.py against torch (float64 conv2d / linear autograd, AdamW) and
finite differences.
"""
from __future__ import annotations

import numpy as np


def bf16_round(a):
    """float32 -> nearest bfloat16 (ties to even), returned as float32 values; NaN stays NaN."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16) & 0xFFFFFFFF
    r = np.where(nan, 0x7FC00000, r).astype(np.uint32)
    return r.view(np.float32).reshape(a32.shape)


def bf16_bits(a):
    return (bf16_round(a).view(np.uint32) >> 16).astype(np.uint16)


def from_bf16_bits(h):
    return (np.asarray(h, np.uint16).astype(np.uint32) << 16).view(np.float32)


def _taps(x):
    """x (N, C, H, W) -> the 9 stride-2 tap views (N, C, H/2, W/2) of the zero-padded input."""
    N, C, H, W = x.shape
    xp = np.zeros((N, C, H + 2, W + 2), x.dtype)
    xp[:, :, 1:H + 1, 1:W + 1] = x
    return [xp[:, :, ky:ky + H:2, kx:kx + W:2] for ky in range(3) for kx in range(3)]


def conv_fwd(x, w, b):
    """pre (N, Co, H/2, W/2) = b + sum_{ci, ky, kx} w[co, ci, ky, kx] x[n, ci, 2oy-1+ky, 2ox-1+kx]."""
    t = _taps(np.asarray(x, np.float64))
    w = np.asarray(w, np.float64)
    out = np.zeros((x.shape[0], w.shape[0]) + t[0].shape[2:])
    for k in range(9):
        out += np.einsum("nchw,oc->nohw", t[k], w[:, :, k // 3, k % 3])
    return out + np.asarray(b, np.float64)[None, :, None, None]


def conv_bwd(x, w, dpre):
    """(dx, dw, db) of conv_fwd for the pre-activation gradient dpre."""
    x = np.asarray(x, np.float64)
    w = np.asarray(w, np.float64)
    N, C, H, W = x.shape
    t = _taps(x)
    dw = np.zeros(w.shape)
    dxp = np.zeros((N, C, H + 2, W + 2))
    for k in range(9):
        ky, kx = k // 3, k % 3
        dw[:, :, ky, kx] = np.einsum("nohw,nchw->oc", dpre, t[k])
        dxp[:, :, ky:ky + H:2, kx:kx + W:2] += np.einsum("nohw,oc->nchw", dpre, w[:, :, ky, kx])
    return dxp[:, :, 1:H + 1, 1:W + 1], dw, dpre.sum((0, 2, 3))


def fc_fwd(x, w, b):
    return np.asarray(x, np.float64) @ np.asarray(w, np.float64).T + np.asarray(b, np.float64)


def fc_bwd(x, w, dpre):
    return dpre @ np.asarray(w, np.float64), dpre.T @ np.asarray(x, np.float64), dpre.sum(0)


def gauss_nll(head, target, weights):
    """Mean over n of sum_j wt_j 0.5 ((t - mu)^2 exp(-s) + s); returns (loss, mse per target, dhead)."""
    head = np.asarray(head, np.float64)
    T = target.shape[1]
    mu, s = head[:, :T], head[:, T:]
    wt = np.asarray(weights, np.float64)[None, :]
    r = target - mu
    e = np.exp(-s)
    n = head.shape[0]
    loss = float(np.sum(wt * 0.5 * (r * r * e + s)) / n)
    d = np.concatenate([-wt * r * e, wt * 0.5 * (1 - r * r * e)], 1) / n
    return loss, np.mean(r * r, 0), d


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """Adam with bias correction and decoupled weight decay (returns new p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * wd * p
    p = p - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v


class NetRef:
    """Forward / backward of the estimator network for a parameter dict {conv{i}.w, conv{i}.b,
    fc0.w, fc0.b, fc1.w, fc1.b} (layer order given by `arch_conv`)."""

    def __init__(self, arch_conv, bf16=True):
        self.arch = arch_conv
        self.bf16 = bf16

    def _q(self, a):
        return bf16_round(a).astype(np.float64) if self.bf16 else np.asarray(a, np.float64)

    def forward(self, P, x):
        q = self._q
        acts = [q(x)]
        for i in range(len(self.arch)):
            pre = conv_fwd(acts[-1], q(P[f"conv{i}.w"]), P[f"conv{i}.b"])
            acts.append(q(np.maximum(pre, 0.0)))
        flat = acts[-1].reshape(acts[-1].shape[0], -1)
        hid = q(np.maximum(fc_fwd(flat, q(P["fc0.w"]), P["fc0.b"]), 0.0))
        head = fc_fwd(hid, q(P["fc1.w"]), P["fc1.b"])
        self.cache = (acts, flat, hid)
        return head

    def backward(self, P, dhead):
        q = self._q
        acts, flat, hid = self.cache
        G = {}
        dh, G["fc1.w"], G["fc1.b"] = fc_bwd(hid, q(P["fc1.w"]), dhead)
        dh = dh * (hid > 0)
        dflat, G["fc0.w"], G["fc0.b"] = fc_bwd(flat, q(P["fc0.w"]), dh)
        d = (dflat * (flat > 0)).reshape(acts[-1].shape)
        for i in range(len(self.arch) - 1, -1, -1):
            dx, G[f"conv{i}.w"], G[f"conv{i}.b"] = conv_bwd(acts[i], q(P[f"conv{i}.w"]), d)
            d = dx * (acts[i] > 0)
        return G


def logmag_crop(mag, x0, z0, zpool, height, width, offset_db, scale_db):
    """(B, nX, rows) float32 -> (B, height, width) float32 values of the bf16 network input:
    clamp((20 log10(max over zpool rows) - offset) / scale, -8, 8), rounded to bf16."""
    m = np.asarray(mag, np.float32)[:, x0:x0 + width, z0:z0 + height * zpool]
    m = m.reshape(m.shape[0], width, height, zpool).max(-1)              # (B, W, H)
    db = (np.float32(20.0) * np.log10(np.maximum(m, np.float32(1e-30)))).astype(np.float32)
    y = np.clip((db - np.float32(offset_db)) * np.float32(1.0 / scale_db), -8, 8).astype(np.float32)
    return bf16_round(np.transpose(y, (0, 2, 1)))
