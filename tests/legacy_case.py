"""Loader for tests/golden/legacy_<case>.npz (inputs + MATLAB plane of the reference's
legacy-parity test, written by tests/golden/make_legacy_fixture.py)."""
import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(case="telesto"):
    d = np.load(GOLDEN / f"legacy_{case}.npz", allow_pickle=False)
    scan = json.loads(str(d["scan"]))
    return {"raw": d["raw"], "lambda_nm": d["lambda_nm"], "plane_db": d["plane_db"], "tie": d["tie"],
            "scan": scan}
