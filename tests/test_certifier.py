"""The product's exact certifier (libsisamd: sis_certify_strings /
sis_export_certify) against the reference's interval certifier built from its
own source (oracle/_ref/certifier), on the reference's own fixtures: contact
strings (tests/test_core.py:27-35), grids, Goebel and strip constructions
(tests/test_cpp_certifier.py:27-56, tests/test_cuda_quality.py:17-21,41-47),
the recorded GPU outputs (tests/test_cuda_demo.py:28-35), input limits
(tests/test_core.py:38-46) and perturbed packings.  No GPU."""
from __future__ import annotations

import ctypes
import json
import math
import sys
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, str(ROOT / "tests" / "golden"))
import analytic  # noqa: E402
from refcert import certify_rows  # noqa: E402


def ours(sis, rows):
    n = len(rows)
    arr = lambda k: (ctypes.c_char_p * n)(*[r[k].encode() for r in rows])  # noqa: E731
    buf = ctypes.create_string_buffer(64)
    rc = sis.lib().sis_certify_strings(n, arr(0), arr(1), arr(2), buf, 64)
    return rc == 0, (buf.value.decode() if rc == 0 else None)


def both(sis, ref_certifier, rows):
    ok1, L1 = ours(sis, rows)
    ok2, L2, _ = certify_rows(ref_certifier, len(rows), rows)
    return (ok1, L1), (ok2, L2)


@pytest.mark.parametrize("x,accepted", [("1.5", True), ("1.499999999999999999999999999999", False),
                                        ("1.500000000000000000000000000001", True), (".6", False)])
def test_contact_no_tolerance(sis, ref_certifier, x, accepted):
    rows = [(".5", ".5", "0"), (x, ".5", "0")]
    a, b = both(sis, ref_certifier, rows)
    assert a == b
    assert a[0] is accepted


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 9, 10, 16, 31])
def test_grid_L_is_ceil_sqrt(sis, ref_certifier, n):
    rows = analytic.grid(n)
    a, b = both(sis, ref_certifier, rows)
    assert a == b and a[0]
    assert float(a[1]) == math.ceil(math.sqrt(n))


def test_goebel_and_strip_strings_match(sis, ref_certifier):
    for a in (1, 2, 4):
        rows = analytic.goebel(a)
        x, y = both(sis, ref_certifier, rows)
        assert x == y and x[0], a
        assert abs(float(x[1]) - (a + 1 + 1 / 2 ** 0.5)) < 1e-14
    ok, base_L, _ = certify_rows(ref_certifier, 5, analytic.tilted_five())
    assert ok
    for k in (2, 3, 4):
        rows = analytic.strip(k, base_L)
        x, y = both(sis, ref_certifier, rows)
        assert x == y and x[0], k
        assert k + 0.70710678 < float(x[1]) < k + 0.7071069


def test_roundtrip_exposes_micro_dilation(sis, ref_certifier):
    """goebel(2)'s 1e-40 dilation vanishes under the float round trip: the
    literal strings certify, the round-tripped ones do not
    (tests/test_cpp_certifier.py:59-102)."""
    rows = analytic.goebel(2)
    assert ours(sis, rows)[0]
    rt = [(repr(float(x)), repr(float(y)), repr(float(a))) for x, y, a in rows]
    ok_ref, _, _ = certify_rows(ref_certifier, len(rows), rows, roundtrip=True)
    assert not ok_ref
    assert not ours(sis, rt)[0]


@pytest.mark.parametrize("n", [5, 10, 11, 17, 50])
def test_recorded_gpu_outputs_certify_identically(sis, ref_certifier, n):
    """cuda/results/raw{n}.json through the dilation ladder of
    cuda/demo.py:25-40: same accepted dilation and the same 50-decimal L."""
    from refcert import certify_q
    raw = json.loads((ROOT / "tests/golden/raw" / f"raw{n}.json").read_text())
    q = np.asarray(raw["q"], dtype=np.float64)
    d_ref, L_ref = certify_q(ref_certifier, q)
    d = ctypes.c_double()
    buf = ctypes.create_string_buffer(64)
    rc = sis.lib().sis_export_certify(n, q.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(d), buf, 64)
    assert rc == 0
    assert d.value == d_ref and buf.value.decode() == L_ref


@pytest.mark.parametrize("field,value", [("x", "sqrt(2)"), ("y", "nan"), ("x", "1e400"), ("x", "10001"),
                                         ("angle", "101"), ("x", "1." + "0" * 101), ("x", ""), ("x", "1e"),
                                         ("x", "."), ("y", "+"), ("x", "0x10")])
def test_input_limits(sis, ref_certifier, field, value):
    rows = [[".5", ".5", "0"]]
    rows[0][{"x": 0, "y": 1, "angle": 2}[field]] = value
    rows = [tuple(r) for r in rows]
    assert ours(sis, rows)[0] is False
    if value != "":
        assert certify_rows(ref_certifier, 1, rows)[0] is False


def test_exponent_and_sign_forms(sis, ref_certifier):
    rows = [("5e-1", "+0.5", "0.0"), ("1.5E0", "0.50", "-0")]
    a, b = both(sis, ref_certifier, rows)
    assert a == b and a[0]


def test_perturbed_packings_agree(sis, ref_certifier):
    """Random tilts and shifts around tight grids: accept/reject and L agree."""
    rng = np.random.default_rng(17)
    agree = accepted = 0
    for trial in range(40):
        n = int(rng.integers(2, 13))
        w = math.isqrt(n - 1) + 1
        gap = float(rng.choice([0.0, 1e-9, 1e-7, 1e-4]))
        rows = []
        for i in range(n):
            x = (i % w + 0.5) * (1 + gap) + float(rng.uniform(-1, 1)) * float(rng.choice([0, 1e-12, 1e-6]))
            y = (i // w + 0.5) * (1 + gap)
            a = float(rng.choice([0.0, float(rng.uniform(-1e-6, 1e-6)), float(rng.uniform(-0.3, 0.3))]))
            rows.append((format(x, ".17g"), format(y, ".17g"), format(a, ".17g")))
        x, y = both(sis, ref_certifier, rows)
        assert x == y, (trial, rows, x, y)
        agree += 1
        accepted += x[0]
    assert agree == 40 and 0 < accepted < 40
