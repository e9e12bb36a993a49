"""Device selection with the MI355X ``hip`` device.

Mirrors octrecon/backend.py (get_xp :61-83, is_torch :86-87, to_numpy :90-95,
synchronize :98-104) and the TorchXP namespace of octrecon/core/torch_backend.py:27-63.
Device memory and streams come from PyTorch-ROCm; the arithmetic of the path runs in
liboct_hip.so.

Devices: ``hip`` (an AMD GPU through liboct_hip), ``auto`` (hip if usable, else an error
naming why: this package has no CPU path).
"""
from __future__ import annotations

import numpy as np

_NP2T = {np.dtype(np.float32): "float32", np.dtype(np.float64): "float64", np.dtype(np.int16): "int16",
         np.dtype(np.int32): "int32", np.dtype(np.int64): "int64", np.dtype(np.bool_): "bool",
         np.dtype(np.uint8): "uint8", np.dtype(np.uint16): "uint16", np.dtype(np.complex64): "complex64"}


class HipXP:
    """The numpy-like namespace the pipeline's accumulation code uses (same members as the
    reference's TorchXP), on a ROCm torch device."""
    nan = float("nan")

    def __init__(self, device: str = "cuda"):
        import torch
        self.torch = torch
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("HipXP needs a ROCm GPU device ('cuda' in PyTorch-ROCm)")
        self.name = "hip"

    def tdtype(self, dtype):
        return getattr(self.torch, _NP2T[np.dtype(dtype)])

    def zeros(self, shape, dtype=np.float32):
        return self.torch.zeros(shape, dtype=self.tdtype(dtype), device=self.device)

    def empty(self, shape, dtype=np.float32):
        return self.torch.empty(shape, dtype=self.tdtype(dtype), device=self.device)

    def asarray(self, a, dtype=None):
        torch = self.torch
        if isinstance(a, torch.Tensor):
            t = a.to(self.device)
        else:
            arr = np.ascontiguousarray(np.asarray(a))
            if arr.dtype.kind in "iu" and arr.dtype not in _NP2T:
                arr = arr.astype(np.int64)
            t = torch.from_numpy(arr).to(self.device)
        return t if dtype is None else t.to(self.tdtype(dtype))

    def log10(self, x):
        return self.torch.log10(x)

    def stream(self):
        """Raw hipStream_t of the current torch stream on this device."""
        return self.torch.cuda.current_stream(self.device).cuda_stream

    def synchronize(self):
        self.torch.cuda.synchronize(self.device)


def hip_status() -> tuple[bool, str]:
    """(usable, message) for the hip device: a visible ROCm GPU and a loadable liboct_hip."""
    try:
        import torch
    except ImportError:
        return False, "PyTorch is not installed"
    if not torch.cuda.is_available() or getattr(torch.version, "hip", None) is None:
        return False, "no ROCm GPU visible to PyTorch"
    try:
        from ._lib import lib
        lib()
    except Exception as e:  # noqa: BLE001
        return False, f"liboct_hip unavailable: {e}"
    return True, f"{torch.cuda.get_device_name(0)} via liboct_hip"


def get_xp(device: str = "auto"):
    """Return (xp, device_name); device_name is 'hip'."""
    if device not in ("auto", "hip", "gpu"):
        raise ValueError(f"device must be auto|hip|gpu, got {device!r}")
    ok, msg = hip_status()
    if not ok:
        raise RuntimeError(f"hip device requested but not usable: {msg}")
    return HipXP("cuda"), "hip"


def is_torch(xp) -> bool:
    return hasattr(xp, "torch") and hasattr(xp, "device")


def to_numpy(a):
    if isinstance(a, np.ndarray):
        return a
    if hasattr(a, "detach"):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def synchronize(xp):
    if xp is np:
        return
    if is_torch(xp):
        xp.synchronize()
