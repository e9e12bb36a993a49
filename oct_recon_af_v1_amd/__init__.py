"""MI355X-native (gfx950) OCT reconstruction hot path: a ``hip`` device for the tiled
spectral-domain OCT reconstruction of atesfet/oct_recon_AF_v1.

Public entry points mirror the reference's names:
  backend.get_xp / HipXP                    device selection, torch-ROCm namespace
  geometry.build_spectral_geometry, ...     float64 host operator builders
  spectral.HipSpectralProcessor             magnitude / linearise / complex_scan
  stitching.HipTileStitcher                 contribute (optical path, focus weight, interp)
  quantize.DeviceFinalizer                  finalise, clim, legacy uint16 quantisation
  pipeline.Reconstructor / ReconConfig      process_row / process_planes / run
The arithmetic runs in liboct_hip.so (HIP kernels, C ABI in include/oct_hip.h); there is no
CPU fallback.
"""
from .geometry import (ScanGrid, SpectralGeometry, TiledGeometry, beta_from_index_units,  # noqa: F401
                       build_spectral_geometry, build_tiled_geometry, chirp_to_lambda, cubic_lambda_map,
                       matlab_colon, matlab_linspace)

__all__ = ["ScanGrid", "SpectralGeometry", "TiledGeometry", "beta_from_index_units", "build_spectral_geometry",
           "build_tiled_geometry", "chirp_to_lambda", "cubic_lambda_map", "matlab_colon", "matlab_linspace"]
