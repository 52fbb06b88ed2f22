"""Run the reference's interval certifier (built from its own source into
oracle/_ref) on a packing: the acceptance checker of the tests."""
from __future__ import annotations

import json
import subprocess
import tempfile
from pathlib import Path

import numpy as np

DILATIONS = (1.000002, 1.00001, 1.0001)  # cuda/demo.py:31


def write_csv(path: Path, puzzles):
    """puzzles: list of (n, rows) with rows = [(x_str, y_str, angle_str)]."""
    lines = ["id,n,square_id,x,y,angle"]
    k = 0
    for n, rows in puzzles:
        for i, (x, y, a) in enumerate(rows):
            lines.append(f"{k},{n},{i},{x},{y},{a}")
            k += 1
    path.write_text("\n".join(lines) + "\n")


def certify_rows(exe: Path, n: int, rows, roundtrip: bool = False):
    """Returns (ok, L string or None, reason)."""
    with tempfile.TemporaryDirectory() as d:
        csv = Path(d) / "p.csv"
        det = Path(d) / "d.json"
        write_csv(csv, [(n, rows)])
        cmd = [str(exe), str(csv), "--max-n", str(n), "--threads", "2", "--details", str(det)]
        if roundtrip:
            cmd.append("--roundtrip")
        subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300)
        res = json.loads(det.read_text())["results"][n - 1]
        return res["L"] is not None, res["L"], res["reason"]


def certify_q(exe: Path, q: np.ndarray, dilations=DILATIONS):
    """The export gate of cuda/demo.py:25-40: dilate centres, %.17g, certify."""
    q = np.asarray(q, dtype=np.float64)
    n = q.shape[0]
    for dil in dilations:
        rows = [(format(x * dil, ".17g"), format(y * dil, ".17g"), format(a, ".17g")) for x, y, a in q]
        ok, L, _ = certify_rows(exe, n, rows)
        if ok:
            return dil, L
    return None, None
