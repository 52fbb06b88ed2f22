"""`bin/certifier` keeps the process contract of the reference's certifier
(cuda/certifier.cpp:933-1077): on the same multi-puzzle CSV files the two
executables give the same counts, the same per-n decisions and 50-decimal L
strings and the same scores, literally and with --roundtrip, and both refuse
malformed files with exit code 2.  The reference executable is built from its
own source into oracle/_ref."""
from __future__ import annotations

import json
import math
import subprocess
import sys

import numpy as np
import pytest

from conftest import PKG, ROOT

sys.path.insert(0, str(ROOT / "tests" / "golden"))
import analytic  # noqa: E402
from refcert import write_csv  # noqa: E402

OURS = PKG / "bin" / "certifier"


def run(exe, csv, tmp_path, *extra):
    det = tmp_path / f"{exe.name}-det.json"
    r = subprocess.run([str(exe), str(csv), "--threads", "4", "--details", str(det), *extra],
                       capture_output=True, text=True, timeout=600)
    return r, (json.loads(det.read_text()) if r.returncode == 0 else None)


def stream():
    """Grids, Goebel and strip packings, recorded GPU outputs, overlapping and
    malformed puzzles and gaps (missing n)."""
    puzzles = {}
    for n in (1, 2, 3, 4, 7, 9, 12, 16, 25):
        puzzles[n] = analytic.grid(n)
    for a in (1, 2):
        rows = analytic.goebel(a)
        puzzles[len(rows)] = rows
    for n in (11, 17):
        raw = json.loads((ROOT / "tests/golden/raw" / f"raw{n}.json").read_text())
        q = np.asarray(raw["q"], dtype=np.float64)
        puzzles[n] = [(format(x * 1.0001, ".17g"), format(y * 1.0001, ".17g"), format(a, ".17g")) for x, y, a in q]
    rows = analytic.grid(6)
    rows[1] = (".9", rows[1][1], rows[1][2])  # overlap
    puzzles[6] = rows
    rows = analytic.grid(8)
    puzzles[8] = rows[:-1] + [rows[0]]  # duplicate square id handled below
    return puzzles


def write_stream(path, puzzles, bad_ids=(8,)):
    lines = ["id,n,square_id,x,y,angle"]
    k = 0
    for n in sorted(puzzles):
        for i, (x, y, a) in enumerate(puzzles[n]):
            sid = 0 if (n in bad_ids and i == n - 1) else i
            lines.append(f"{k},{n},{sid},{x},{y},{a}")
            k += 1
    path.write_text("\n".join(lines) + "\n")


@pytest.mark.parametrize("roundtrip", [False, True])
def test_same_decisions_L_and_scores_as_reference(ref_certifier, tmp_path, roundtrip):
    assert OURS.exists()
    csv = tmp_path / "stream.csv"
    write_stream(csv, stream())
    extra = ("--max-n", "30") + (("--roundtrip",) if roundtrip else ())
    r1, d1 = run(OURS, csv, tmp_path, *extra)
    r2, d2 = run(ref_certifier, csv, tmp_path, *extra)
    assert r1.returncode == 0 and r2.returncode == 0, (r1.stderr, r2.stderr)
    s1, s2 = json.loads(r1.stdout), json.loads(r2.stdout)
    for k in ("max_n", "accepted", "missing", "rejected", "roundtrip"):
        assert s1[k] == s2[k], k
    assert abs(s1["score"] - s2["score"]) <= 1e-12
    assert s1["accepted"] >= 10 and s1["rejected"] >= 2 and s1["missing"] > 0
    for a, b in zip(d1["results"], d2["results"]):
        assert a["n"] == b["n"]
        assert a["L"] == b["L"], a["n"]
        assert (a["reason"] == "ok") == (b["reason"] == "ok")
        assert (a["reason"] == "missing from submission") == (b["reason"] == "missing from submission")
        assert abs(a["score"] - b["score"]) <= 1e-15 * max(1.0, abs(b["score"]))
    # the scores are the documented formula
    for a in d1["results"]:
        n = a["n"]
        if a["L"] is None:
            assert a["score"] == math.ceil(math.sqrt(n)) / math.sqrt(n) - 1
        else:
            assert abs(a["score"] - max(0.0, float(a["L"]) / math.sqrt(n) - 1)) <= 1e-15


@pytest.mark.parametrize("text", ["", "id,n,x,y\n", "id,n,square_id,x,y,angle\n5,1,0,.5,.5,0\n",
                                  "id,n,square_id,x,y,angle\n0,1,0,.5,.5\n",
                                  "id,n,square_id,x,y,angle\n0,1001,0,.5,.5,0\n"])
def test_malformed_files_exit_2(ref_certifier, tmp_path, text):
    csv = tmp_path / "bad.csv"
    csv.write_text(text)
    for exe in (OURS, ref_certifier):
        r = subprocess.run([str(exe), str(csv), "--max-n", "5"], capture_output=True, text=True, timeout=60)
        assert r.returncode == 2, exe


def test_usage_and_missing_file_exit_2(tmp_path):
    assert subprocess.run([str(OURS)], capture_output=True).returncode == 2
    assert subprocess.run([str(OURS), str(tmp_path / "none.csv")], capture_output=True).returncode == 2


def test_write_csv_helper_roundtrip(tmp_path):
    """The CSV writer used by the tests emits the header and consecutive ids."""
    p = tmp_path / "a.csv"
    write_csv(p, [(1, [(".5", ".5", "0")]), (2, [(".5", ".5", "0"), ("1.5", ".5", "0")])])
    lines = p.read_text().splitlines()
    assert lines[0] == "id,n,square_id,x,y,angle" and [ln.split(",")[0] for ln in lines[1:]] == ["0", "1", "2"]
