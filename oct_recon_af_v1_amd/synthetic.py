"""Seeded synthetic spectral-domain OCT raw data (the benchmark and test workloads).

Generator specification (all sizes per B-scan file, uint16 12-bit counts):
  * wavelength map: cubic lambda(p) = 1300 + 0.0625 u - 4e-6 u^2 + 1.5e-9 u^3 nm,
    u = p - 1023.5 (1230-1361 nm over 2048 pixels);
  * source: S(lambda) = peak * exp(-4 ln2 (lambda - 1300)^2 / 100^2), peak 3000 counts;
  * each A-scan: 5 reflectors at depths uniform in [lo, hi) equispaced-k depth bins,
    reflectivity R log-uniform in [r_min, r_max], fringe amplitude sqrt(R), random phase;
    fringe phase 2 pi z (k - k_min) / (N dk) + phi, z in bins, plus the quadratic
    dispersion +beta ((k - k0)/dk)^2 with beta in rad per squared equispaced-k sample
    (k0 the mean wavenumber, dk the equispaced step); reconstructing with the window's
    dispersion term beta / dk^2 (geometry.beta_from_index_units) compensates it (the
    equispaced k grid runs from k_max down, so the kept half of the inverse FFT carries
    exp(-i beta dq), which the window's exp(-i beta_ref (k - k0)^2) cancels for beta_ref > 0);
  * the first `apod` A-lines of every file are background lines (source only);
  * counts = clip(Poisson(S (1 + sum_r sqrt(R_r) cos(phase_r))), 0, 4095).
Arithmetic runs on any torch device (the GPU for the large workloads); identical seeds give
identical data on one device type.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .geometry import cubic_lambda_map


@dataclass
class SpectrumSpec:
    n: int = 2048
    peak: float = 3000.0
    center_nm: float = 1300.0
    fwhm_nm: float = 100.0
    n_reflectors: int = 5
    r_min: float = 1e-4
    r_max: float = 1e-1
    z_lo: float = 16.0
    z_hi: float = 1000.0
    apod: int = 25

    def lambda_nm(self):
        return cubic_lambda_map(self.n)


def _k_axes(lam):
    k = 2 * np.pi / np.asarray(lam, float)
    dk = (k.max() - k.min()) / (len(k) - 1)
    kn = (k - k.min()) / dk                   # position in equispaced-k samples, 0..N-1
    k0n = (k.mean() - k.min()) / dk
    return kn, k0n


def make_bscans(n_files: int, n_x: int, spec: SpectrumSpec = SpectrumSpec(), beta: float = 0.0, seed: int = 0,
                device="cpu", chunk_files: int = 8, out=None):
    """(n_files, apod + n_x, N) uint16 torch tensor on `device`.  `beta` may be a float or a
    per-file sequence."""
    import torch
    dev = torch.device(device)
    lam = spec.lambda_nm()
    N = spec.n
    kn, k0n = _k_axes(lam)
    S = spec.peak * np.exp(-4 * np.log(2) * (lam - spec.center_nm) ** 2 / spec.fwhm_nm ** 2)
    kn_t = torch.tensor(kn, dtype=torch.float32, device=dev)
    dq_t = torch.tensor((kn - k0n) ** 2, dtype=torch.float32, device=dev)
    S_t = torch.tensor(S, dtype=torch.float32, device=dev)
    lines = spec.apod + n_x
    if out is None:
        out = torch.empty((n_files, lines, N), dtype=torch.uint16, device=dev)
    betas = np.broadcast_to(np.asarray(beta, float), (n_files,))
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    R = spec.n_reflectors
    two_pi_over_n = 2 * np.pi / N
    for f0 in range(0, n_files, chunk_files):
        nf = min(chunk_files, n_files - f0)
        z = spec.z_lo + (spec.z_hi - spec.z_lo) * torch.rand((nf, n_x, R, 1), generator=g, device=dev)
        lr = np.log(spec.r_min) + (np.log(spec.r_max) - np.log(spec.r_min)) * torch.rand(
            (nf, n_x, R, 1), generator=g, device=dev)
        amp = torch.exp(0.5 * lr)
        phi = 2 * np.pi * torch.rand((nf, n_x, R, 1), generator=g, device=dev)
        bt = torch.tensor(betas[f0:f0 + nf], dtype=torch.float32, device=dev).view(nf, 1, 1, 1)
        acc = torch.ones((nf, n_x, N), dtype=torch.float32, device=dev)
        for r in range(R):
            ph = (two_pi_over_n * z[:, :, r]) * kn_t + phi[:, :, r] + bt[:, :, 0] * dq_t
            acc += amp[:, :, r] * torch.cos(ph)
        rate = torch.clamp(S_t * acc, min=0.0)
        sig = torch.poisson(rate, generator=g)
        bg = torch.poisson(S_t.expand(nf, spec.apod, N).contiguous(), generator=g)
        blk = torch.cat([bg, sig], dim=1).clamp_(0, 4095)
        out[f0:f0 + nf] = blk.to(torch.int32).to(torch.uint16)
    return out
