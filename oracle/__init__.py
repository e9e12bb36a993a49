"""ORACLE - test infrastructure only.

CPU restatements of the reference's reconstruction path (octrecon, atesfet/oct_recon_AF_v1),
used as the checker of the device implementation.  Only ``tests/``, ``__graft_entry__.smoke``
and the ``cpu_baseline`` leg of ``bench.py`` may import this package; the product package
never does.

Modules
  ref_path     float32 "reference CPU path" (C pre-FFT loop + scipy pocketfft, as the
               reference's Numba backend) and the float64 oracle of the whole path,
               stitching, finalise, clim and quantisation.
  ref_focus    the optical-path corrected x-median depth profile of the focus detector.
  ref_io       raw-file readers (Thorlabs OCITY, SRR 12-bit payloads).
  ref_tiff     a BigTIFF page reader for the legacy writer's output.
  cnn_ref      the learned estimator's network, gradients, loss and Adam in float64.

The inputs of the MATLAB plane stored with the reference's tests are regenerated by
tests/golden/make_legacy_fixture.py (a restatement of the reference's synthetic-volume
generator) and kept in tests/golden/legacy_telesto.npz.

Parity status of every function is stated in its docstring; functions without an
independent pin say "parity unpinned".
"""
