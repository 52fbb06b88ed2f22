"""The allocation-free plain solve (sis_solve_noalloc) through the C ABI: the
same results as sis_solve, bit for bit, on two calls in a row of which the
second is smaller (the shared device array keeps the first call's contents),
in both improvement modes; a request too large for the array is rejected
before any device work."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("target_L", "L_cert", "lambda", "e32", "e64_before", "e64_after", "proposals", "accepted", "sweeps",
          "exchanges_tried", "exchanges_accepted", "candidates", "near_pairs", "shrinks", "feasible", "replicas",
          "fp64_rejects", "kicks", "feasible_checkpoints")


def say(*a):
    print("[gpu-case]", *a, flush=True)


def _same(a, b):
    assert np.array_equal(a.q, b.q)
    for f in FIELDS:
        assert getattr(a, f) == getattr(b, f), (f, getattr(a, f), getattr(b, f))


@pytest.mark.parametrize("improve,cool", [(0, 0), (1, 200)])
def test_noalloc_solve_equals_plain_solve(sis, improve, cool):
    calls = [(60, 9.0, 64, 400), (12, 5.0, 16, 300)]  # the second smaller than the first
    for n, L, R, sweeps in calls:
        p = sis.default_params(replicas=R, ladder=8, sweeps=sweeps, sweeps_per_exchange=10, tmin=1e-4, tmax=1e-2,
                               shrink=1e-3, improve=improve, cool_period=cool, init_mode=1, rng_seed=11,
                               polish_iters=50)
        say("noalloc", improve, cool, n, R, sweeps)
        a = sis.solve_noalloc(n, L, None, params=p)
        b = sis.solve(n, L, None, params=p)
        _same(a, b)
        assert a.proposals == n * R * sweeps


def test_noalloc_rejects_requests_beyond_its_buffer(sis):
    p = sis.default_params(replicas=4096, sweeps=10)
    with pytest.raises(sis.SisError) as e:
        sis.solve_noalloc(1000, 40.0, None, params=p)
    assert e.value.code == sis.SIS_EINVAL and "too large" in str(e.value)
