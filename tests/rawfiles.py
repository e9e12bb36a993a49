"""Writers of raw tile folders in the reference's layouts, for the reader tests."""
from pathlib import Path

import numpy as np


def write_ocity(folder, raw):
    """raw (files, interf_size, n_lambda) -> <folder>/data/Spectral{i}.data (<i2)."""
    d = Path(folder) / "data"
    d.mkdir(parents=True, exist_ok=True)
    for i, a in enumerate(np.asarray(raw)):
        a.astype("<i2").tofile(d / f"Spectral{i}.data")


def write_srr(folder, payload, y_total, b_total, apod, scan, system="Ganymede_SRR", pad_tokens=True):
    """payload (y_total b_total, file_lines, n_lambda) uint16 -> SRR files; apod / scan are
    (start, end) A-line regions."""
    from oct_recon_af_v1_amd.rawio import srr_file_name
    folder = Path(folder)
    folder.mkdir(parents=True, exist_ok=True)
    n = payload.shape[2]
    toks = [f"size1={n}", f"scanregions=1,{scan[0]},{scan[1]}", f"aporegions=1,{apod[0]},{apod[1]}"]
    if pad_tokens:
        toks += ["version=2", "exposure=12.5e-6", "name='probe'"]
    body = (" ".join(toks) + "\n").encode()
    head = f"headersize={len(body)}\n".encode() + body
    for f in range(payload.shape[0]):
        y0, b0 = divmod(f, b_total)
        with open(folder / srr_file_name(y0 + 1, y_total, b0 + 1, b_total, system), "wb") as fh:
            fh.write(head)
            fh.write(payload[f].astype("<u2").tobytes())
