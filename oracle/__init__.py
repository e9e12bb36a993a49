"""ORACLE - test infrastructure only.

CPU restatements of the reference's reconstruction path (octrecon, atesfet/oct_recon_AF_v1),
used as the checker of the device implementation.  Only ``tests/``, ``__graft_entry__.smoke``
and the ``cpu_baseline`` leg of ``bench.py`` may import this package; the product package
never does.

Modules
  ref_path     float32 "reference CPU path" (C pre-FFT loop + scipy pocketfft, as the
               reference's Numba backend) and the float64 oracle of the whole path,
               stitching, finalise, clim and quantisation.
  legacy_volume  bit-compatible restatement of the reference's synthetic-volume generator,
               used to reproduce the MATLAB planes stored with the reference's tests.

Parity status of every function is stated in its docstring; functions without an
independent pin say "parity unpinned".
"""
