"""ctypes binding of liboct_hip.so (C ABI declared in include/oct_hip.h).

The library is the only compute path: if it is missing or fails to load, every call raises.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "liboct_hip.so"

OCT_RAW_I16, OCT_RAW_U16, OCT_RAW_F32 = 0, 1, 2
OCT_APOD_LINES, OCT_APOD_FRAME_MEAN = 0, 1
OCT_OUT_MAG, OCT_OUT_COMPLEX, OCT_OUT_LINEARISED = 0, 1, 2
OCT_METRICS = {"top_percentile": 0, "kurtosis": 1, "entropy": 2, "legacy_peak": 3}

_vp = ctypes.c_void_p
_i = ctypes.c_int
_ll = ctypes.c_longlong
_sz = ctypes.c_size_t
_f = ctypes.c_float


class StitchDesc(ctypes.Structure):
    """Mirror of oct_stitch_desc."""
    _fields_ = [
        ("mag", _vp), ("mag_stride", _ll), ("mag_row0", _i), ("mag_rows", _i),
        ("n_x", _i), ("n_z", _i),
        ("frame_idx", _vp), ("shift", _vp), ("valid_lo", _vp), ("valid_hi", _vp),
        ("fz", _vp), ("z_rows", _vp), ("z_w", _vp), ("x_cols", _vp), ("x_w", _vp),
        ("num", _vp), ("den", _vp),
        ("B", _i), ("n_out_z", _i), ("C", _i), ("ldx", _ll), ("out_b0", _i), ("out_c0", _i),
        ("accumulate", _i),
    ]


class RowTile(ctypes.Structure):
    """Mirror of oct_row_tile."""
    _fields_ = [("mag", _vp), ("mag_stride", _ll), ("mag_row0", _i), ("c0", _i), ("C", _i),
                ("z_rows", _vp), ("z_w", _vp), ("x_cols", _vp), ("x_w", _vp), ("fz", _vp)]


# name -> (restype, argtypes); every symbol include/oct_hip.h declares
SIGNATURES = {
    "oct_plan_create": (_i, [ctypes.POINTER(_vp), _i, _i, _i, _vp, _vp, _vp]),
    "oct_plan_destroy": (_i, [_vp]),
    "oct_spectral_workspace_bytes": (_sz, [_vp, _i]),
    "oct_background": (_i, [_vp, _vp, _i, _i, _i, _i, _i, _i, _vp, _vp]),
    "oct_spectral": (_i, [_vp, _vp, _i, _i, _i, _i, _i, _i, _i, _i, _vp, _ll, _vp, _vp]),
    "oct_stitch": (_i, [ctypes.POINTER(StitchDesc), _vp]),
    "oct_stitch_row_blocks": (_i, [_i, _i]),
    "oct_stitch_row": (_i, [ctypes.POINTER(RowTile), _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _i, _f, _vp, _i, _vp,
                            _vp]),
    "oct_stitch_row_table": (_i, [ctypes.POINTER(RowTile), _vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _i, _f, _vp,
                                  _i, _vp, _vp]),
    "oct_volume_clim": (_i, [_vp, _i, _i, _vp, _vp, _vp]),
    "oct_finalize_workspace_bytes": (_sz, [_i, _ll]),
    "oct_finalize": (_i, [_vp, _vp, _vp, _i, _ll, _f, _vp, _vp, _vp, _sz, _vp]),
    "oct_quantize": (_i, [_vp, _vp, _i, _ll, _vp, _vp, _i, _vp]),
    "oct_opplan_create": (_i, [ctypes.POINTER(_vp), _i, _i, _i, _vp, _vp, _vp, _i, _i, ctypes.POINTER(_i), _vp]),
    "oct_opplan_destroy": (_i, [_vp]),
    "oct_opplan_set_window": (_i, [_vp, _vp, _vp]),
    "oct_srr_unpack": (_i, [_vp, _i, _i, _i, _i, _i, _i, _i, _vp, _vp]),
    "oct_packbits_chunks": (_i, [_i]),
    "oct_packbits_sizes": (_i, [_vp, _ll, _i, _vp, _vp]),
    "oct_packbits_write": (_i, [_vp, _ll, _i, _vp, _vp, _vp]),
    "oct_opplan_scale_bound": (_i, [_vp, ctypes.POINTER(ctypes.c_double)]),
    "oct_opplan_rows": (_i, [_i, ctypes.POINTER(ctypes.c_int)]),
    "oct_opspectral": (_i, [_vp, _vp, _i, _i, _i, _i, _i, _vp, _ll, _vp, _vp]),
    "oct_plan_use_window": (_i, [_vp, _vp]),
    "oct_topfrac_metric_workspace_bytes": (_sz, [_i, _i, _i, _i]),
    "oct_topfrac_metric": (_i, [_vp, _i, _i, _i, _i, ctypes.c_double, _vp, _vp, _sz, _vp]),
    "oct_topfrac_metric_noalloc": (_i, [_vp, _i, _i, _i, _i, ctypes.c_double, _vp, _vp]),
    "oct_sharpness_metric_workspace_bytes": (_sz, [_i, _i, _i, _i, _i]),
    "oct_sharpness_metric": (_i, [_i, _vp, _i, _i, _i, _i, _vp, _vp, _sz, _vp]),
    "oct_order_stats_workspace_bytes": (_sz, [_i, _i, _i, _i, _i]),
    "oct_order_stats": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _i, _vp, _vp, _sz, _vp]),
    "oct_plan_use_windows": (_i, [_vp, _vp, _ll, _ll]),
    "oct_spectral_repeat": (_i, [_vp, _vp, _i, _i, _i, _i, _i, _i, _i, _i, _i, _i, _i, _vp, _ll, _vp, _vp]),
    "oct_depth_median": (_i, [_vp, _i, _i, _ll, _i, _vp, _vp]),
    "oct_depth_median_opc": (_i, [_vp, _i, _i, _ll, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "oct_logmag_crop": (_i, [_vp, _i, _i, _ll, _i, _i, _i, _i, _i, _i, _f, _f, _vp, _i, _vp]),
    "oct_cnn_conv_fwd": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp]),
    "oct_cnn_conv_fwd_mfma": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp]),
    "oct_cnn_conv_bwd_data": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp]),
    "oct_cnn_conv_bwd_weight_workspace_bytes": (_sz, [_i, _i, _i, _i, _i]),
    "oct_cnn_conv_bwd_weight": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp, _sz, _vp]),
    "oct_cnn_conv_bwd_weight_mfma_workspace_bytes": (_sz, [_i, _i, _i, _i, _i]),
    "oct_cnn_conv_bwd_weight_mfma": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp, _sz, _vp]),
    "oct_cnn_conv_bwd_data_mfma_workspace_bytes": (_sz, [_i, _i, _i, _i, _i]),
    "oct_cnn_conv_bwd_data_mfma": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp, _sz, _vp]),
    "oct_cnn_fc_fwd": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _vp]),
    "oct_cnn_fc_bwd_data": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _vp]),
    "oct_cnn_fc_bwd_weight": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _vp]),
    "oct_cnn_gauss_nll": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _vp]),
    "oct_cnn_adam": (_i, [_vp, _vp, _vp, _vp, _vp, _ll, _f, _f, _f, _f, _f, _i, _vp]),
    "oct_dispersion_windows": (_i, [_vp, _vp, _i, ctypes.c_double, ctypes.c_double, ctypes.c_double, _i, _vp, _vp,
                                    _i, _vp, _vp]),
    "oct_binned_lines": (_i, [_i, _i]),
    "oct_bin_lines": (_i, [_vp, _i, _i, _i, _i, _i, _i, _vp, _vp]),
    "oct_average_repeats": (_i, [_vp, _vp, _i, _i, _i, _i, _i, _ll, _ll, _vp]),
    "oct_last_error": (ctypes.c_char_p, []),
}

_LIB = None


class HipLibraryError(RuntimeError):
    pass


def lib():
    """Load liboct_hip.so once; raise if it is not built."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise HipLibraryError(f"{LIB_PATH} is not built: run `python -m oct_recon_af_v1_amd.build` "
                                  "(there is no CPU fallback)")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc: int, what: str):
    if rc != 0:
        msg = lib().oct_last_error().decode(errors="replace")
        if rc == -1:
            raise ValueError(f"{what}: {msg}")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")
