"""Host-side checks of the learned estimator (no GPU): every launch of the network's forward,
loss, backward and optimiser sequence is replayed against a stand-in library that checks each
pointer argument's extent (from the call's own shape arguments) against the tensors the host
allocated, so a wrong shape passed to a kernel is caught on the CPU."""
import numpy as np
import pytest
import torch

from oct_recon_af_v1_amd import _lib
from oct_recon_af_v1_amd import learned as LN


def _tensors(obj, out):
    if isinstance(obj, torch.Tensor):
        out.append(obj)
    elif isinstance(obj, dict):
        for v in obj.values():
            _tensors(v, out)
    elif isinstance(obj, (list, tuple)):
        for v in obj:
            _tensors(v, out)
    return out


class _ExtentLib:
    """Each entry point checks [ptr, ptr + bytes) against the registered host tensors."""

    def __init__(self, roots):
        self.roots = roots
        self.calls = []

    def _chk(self, ptr, nbytes, what):
        if ptr is None:
            return
        for t in _tensors(self.roots, []):
            base = t.data_ptr()
            if base <= ptr and ptr + nbytes <= base + t.numel() * t.element_size():
                return
        raise AssertionError(f"{what}: {nbytes} bytes at {ptr:#x} outside every allocation")

    def oct_cnn_conv_fwd(self, x, w, b, y, n, ci, h, wd, co, s):
        self.calls.append(("conv_fwd", n, ci, h, co))
        self._chk(x, 2 * n * ci * h * wd, "conv x")
        self._chk(w, 2 * co * ci * 9, "conv w")
        self._chk(b, 4 * co, "conv b")
        self._chk(y, 2 * n * co * (h // 2) * (wd // 2), "conv y")
        return 0

    oct_cnn_conv_fwd_mfma = oct_cnn_conv_fwd

    def oct_cnn_conv_bwd_data(self, dpre, w, xmask, dx, n, ci, h, wd, co, s):
        self.calls.append(("conv_bwd_data", n, ci, h, co))
        self._chk(dpre, 4 * n * co * (h // 2) * (wd // 2), "dpre")
        self._chk(w, 2 * co * ci * 9, "w")
        self._chk(xmask, 2 * n * ci * h * wd, "xmask")
        self._chk(dx, 4 * n * ci * h * wd, "dx")
        return 0

    def oct_cnn_conv_bwd_weight_workspace_bytes(self, n, ci, h, wd, co):
        return 40 * co * ci

    def oct_cnn_conv_bwd_weight(self, dpre, x, dw, db, n, ci, h, wd, co, ws, wsb, s):
        self.calls.append(("conv_bwd_weight", n, ci, h, co))
        self._chk(dpre, 4 * n * co * (h // 2) * (wd // 2), "dpre")
        self._chk(x, 2 * n * ci * h * wd, "x")
        self._chk(dw, 4 * co * ci * 9, "dw")
        self._chk(db, 4 * co, "db")
        self._chk(ws, wsb, "workspace")
        assert wsb >= 40 * co * ci
        return 0

    # the matrix-core backward: the same extents, plus a workspace of the size its query gives
    # (the data gradient's query is 0 where the input channels are not a multiple of 16)
    def oct_cnn_conv_bwd_weight_mfma_workspace_bytes(self, n, ci, h, wd, co):
        return 2 * (9 * ci + 16) * (n * (h // 2) * (wd // 2) + 32) + 64 * co * (9 * ci + 16)

    def oct_cnn_conv_bwd_weight_mfma(self, dpre, x, dw, db, n, ci, h, wd, co, ws, wsb, s):
        self.calls.append(("conv_bwd_weight", n, ci, h, co))
        self._chk(dpre, 4 * n * co * (h // 2) * (wd // 2), "dpre")
        self._chk(x, 2 * n * ci * h * wd, "x")
        self._chk(dw, 4 * co * ci * 9, "dw")
        self._chk(db, 4 * co, "db")
        self._chk(ws, wsb, "workspace")
        assert wsb >= self.oct_cnn_conv_bwd_weight_mfma_workspace_bytes(n, ci, h, wd, co)
        return 0

    def oct_cnn_conv_bwd_data_mfma_workspace_bytes(self, n, ci, h, wd, co):
        return 0 if ci % 16 else 4 * n * co * (h // 2) * (wd // 2) + 18 * ci * co

    def oct_cnn_conv_bwd_data_mfma(self, dpre, w, xmask, dx, n, ci, h, wd, co, ws, wsb, s):
        assert self.oct_cnn_conv_bwd_data_mfma_workspace_bytes(n, ci, h, wd, co) > 0
        self.calls.append(("conv_bwd_data", n, ci, h, co))
        self._chk(dpre, 4 * n * co * (h // 2) * (wd // 2), "dpre")
        self._chk(w, 2 * co * ci * 9, "w")
        self._chk(xmask, 2 * n * ci * h * wd, "xmask")
        self._chk(dx, 4 * n * ci * h * wd, "dx")
        self._chk(ws, wsb, "workspace")
        assert wsb >= self.oct_cnn_conv_bwd_data_mfma_workspace_bytes(n, ci, h, wd, co)
        return 0

    def oct_cnn_fc_fwd(self, x, w, b, ybf, yf, n, k, o, s):
        self.calls.append(("fc_fwd", n, k, o))
        assert (ybf is None) != (yf is None)
        self._chk(x, 2 * n * k, "fc x")
        self._chk(w, 2 * o * k, "fc w")
        self._chk(b, 4 * o, "fc b")
        self._chk(ybf, 2 * n * o, "fc y bf16")
        self._chk(yf, 4 * n * o, "fc y f32")
        return 0

    def oct_cnn_fc_bwd_data(self, dpre, w, xmask, dx, n, k, o, s):
        self.calls.append(("fc_bwd_data", n, k, o))
        self._chk(dpre, 4 * n * o, "dpre")
        self._chk(w, 2 * o * k, "w")
        self._chk(xmask, 2 * n * k, "xmask")
        self._chk(dx, 4 * n * k, "dx")
        return 0

    def oct_cnn_fc_bwd_weight(self, dpre, x, dw, db, n, k, o, s):
        self.calls.append(("fc_bwd_weight", n, k, o))
        self._chk(dpre, 4 * n * o, "dpre")
        self._chk(x, 2 * n * k, "x")
        self._chk(dw, 4 * o * k, "dw")
        self._chk(db, 4 * o, "db")
        return 0

    def oct_cnn_gauss_nll(self, head, tgt, wt, loss, dhead, n, t, s):
        self.calls.append(("loss", n, t))
        self._chk(head, 8 * n * t, "head")
        self._chk(tgt, 4 * n * t, "target")
        self._chk(wt, 4 * t, "weights")
        self._chk(loss, 4 * (1 + t), "loss")
        self._chk(dhead, 8 * n * t, "dhead")
        return 0

    def oct_cnn_adam(self, p, g, m, v, pbf, count, lr, b1, b2, eps, wd, step, s):
        self.calls.append(("adam", count))
        for ptr, nb, w in ((p, 4, "p"), (g, 4, "g"), (m, 4, "m"), (v, 4, "v"), (pbf, 2, "pbf")):
            self._chk(ptr, nb * count, w)
        return 0

    def oct_last_error(self):
        return b""


class _CpuXP:
    torch = torch
    device = torch.device("cpu")

    def stream(self):
        return None


@pytest.mark.parametrize("matrix_cores", [True, False])
def test_network_launch_extents(monkeypatch, matrix_cores):
    xp = _CpuXP()
    net = LN.EstimatorNet(xp, seed=0)
    net.matrix_cores_backward = matrix_cores
    n = 3
    x = torch.zeros((n, 2, LN.IMG, LN.IMG), dtype=torch.bfloat16)
    tgt = torch.zeros((n, 2), dtype=torch.float32)
    wt = torch.ones(2, dtype=torch.float32)
    fake = _ExtentLib([net.p32, net.pbf, net.grad, net.m, net.v, net._bufs, x, tgt, wt])
    monkeypatch.setattr(_lib, "lib", lambda: fake)
    net.forward(x)
    net.loss(n, tgt, wt)
    net.backward(n)
    net.adam(1e-3)
    kinds = [c[0] for c in fake.calls]
    assert kinds.count("conv_fwd") == len(LN.ARCH_CONV) and kinds.count("conv_bwd_weight") == len(LN.ARCH_CONV)
    assert kinds.count("conv_bwd_data") == len(LN.ARCH_CONV) - 1 and kinds.count("fc_fwd") == 2
    # the conv backward calls see the input side of their own layer
    for c in fake.calls:
        if c[0].startswith("conv"):
            i = [ci for ci, _ in LN.ARCH_CONV].index(c[2]) if c[0] == "conv_fwd" else None
            assert c[3] in [LN.IMG >> j for j in range(len(LN.ARCH_CONV))]
    bw = [c for c in fake.calls if c[0] == "conv_bwd_weight"]
    assert [c[3] for c in bw] == [LN.IMG >> i for i in range(len(LN.ARCH_CONV) - 1, -1, -1)]


def test_network_rejects_wrong_input():
    xp = _CpuXP()
    net = LN.EstimatorNet(xp, seed=0)
    with pytest.raises(ValueError):
        net.forward(torch.zeros((2, 2, 128, 128), dtype=torch.bfloat16))
    with pytest.raises(ValueError):
        net.forward(torch.zeros((2, 2, LN.IMG, LN.IMG), dtype=torch.float32))
    with pytest.raises(ValueError):
        LN.EstimatorNet(xp, params=np.zeros(10, np.float32))
