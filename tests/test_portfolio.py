"""Portfolio starts (init_mode 2) on the host: the families of
cuda/packing.cu:304-452 by ladder position and the structured jitter of
cuda/packing.cu:454-475.  The reference writes no golden vectors for them, so
they are pinned by geometry: the FP64 SAT oracle, the known optimal layout of
five squares and Goebel's family a + 1 + 1/sqrt(2) (the formulas of
cuda/scheduler.py:25-74, restated in tests/golden/analytic.py)."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle_lib import f64, ptr

H = 1 / math.sqrt(2)


def seed_grid(n):
    w = math.isqrt(n - 1) + 1
    i = np.arange(n)
    return float(w), np.stack([i % w + 0.5, i // w + 0.5, np.zeros(n)], axis=1)


def candidates(sis, n, L, q):
    """Replica 8k + 2 holds topology candidate k (k < 64) when the request has
    64 ladders of 8."""
    P = sis.portfolio(n, L, q, replicas=512, ladder=8, rng_seed=3, jitter=False)
    seen, out = set(), []
    for k in range(64):
        c = P[8 * k + 2]
        key = c.round(9).tobytes()
        if key not in seen:
            seen.add(key)
            out.append(c)
    return P, out


def overlap_free(oracle, q, touch=1e-12):
    """No pair penetrates deeper than `touch`: the corner stacks sit at unit
    pitch, so their neighbours touch and FP64 rounding of the shared eps can
    leave a penetration of an ulp."""
    n = len(q)
    for i in range(n):
        for j in range(i + 1, n):
            if oracle.lib.o_sat_oracle64(ptr(q[i].copy(), f64), ptr(q[j].copy(), f64)) > touch:
                return False
    return True


def side_of(q):
    """Bounding side of the squares (centre extent + half the rotated width)."""
    e = 0.5 * (np.abs(np.cos(q[:, 2])) + np.abs(np.sin(q[:, 2])))
    lo = min((q[:, 0] - e).min(), (q[:, 1] - e).min())
    hi = max((q[:, 0] + e).max(), (q[:, 1] + e).max())
    return hi - lo, lo


@pytest.mark.parametrize("a", [1, 2, 4])
def test_goebel_family_is_among_the_topologies(sis, oracle, a):
    """Goebel's packing of n = a(a+1) + 3 + floor((a-1) sqrt2) squares in
    a + 1 + 1/sqrt2 (two corner stacks of a rows, two single corners, one
    tilted band) is one of the candidates, overlap-free in FP64."""
    n = a * (a + 1) + 3 + math.floor((a - 1) * math.sqrt(2))
    L, q = seed_grid(n)
    _, cands = candidates(sis, n, L, q)
    target = a + 1 + H
    best = None
    for c in cands:
        side, lo = side_of(c)
        if lo >= -1e-9 and side <= target + 1e-5 and overlap_free(oracle, c):
            best = side
            break
    assert best is not None, f"no overlap-free candidate of side {target} for n={n}"
    assert abs(best - target) < 1e-5


def test_tilted_bands_are_overlap_free_and_inside(sis, oracle):
    """In every candidate the band of 45-degree squares is overlap-free and
    inside the box (side at most seed L + 0.15): a wrong sign or index of the
    diagonal frame, the column pitch or the v range fails this.  (The corner
    stacks of the smallest boxes may overflow or meet each other: the
    reference keeps such layouts as starting points, filtered by capacity
    only.)"""
    for n in (5, 11, 17, 30, 50):
        L, q = seed_grid(n)
        _, cands = candidates(sis, n, L, q)
        assert cands
        bands = 0
        for c in cands:
            tilt = c[np.isclose(c[:, 2], math.pi / 4)]
            if len(tilt) == 0:
                continue
            bands += 1
            assert overlap_free(oracle, tilt, touch=0.0)
            e = 0.5 * math.sqrt(2)
            assert (tilt[:, :2] - e).min() >= -1e-9
            assert (tilt[:, :2] + e).max() <= L + 0.15 + 1e-9
        assert bands > 0


def test_classes_by_ladder_position(sis):
    """Classes 0-1 keep the seed on the first two ladders, 2-4 share ladder k's
    topology, 5-6 are wall grids (6 with its far corner cut), 7 has the four
    corner squares."""
    n = 30
    L, q = seed_grid(n)
    P = sis.portfolio(n, L, q, replicas=64, ladder=8, rng_seed=3, jitter=False)
    for lad in range(2):
        for g in (0, 1):
            assert np.array_equal(P[8 * lad + g], q)
    for lad in range(8):
        assert np.array_equal(P[8 * lad + 2], P[8 * lad + 3])
        assert np.array_equal(P[8 * lad + 2], P[8 * lad + 4])
        wall = P[8 * lad + 5]
        k = math.isqrt(n)
        cx = max(1, k - 1 - lad % 2)
        cy = max(1, k - 1 - (lad // 2) % 2)
        grid = {(i + 0.5, j + 0.5) for i in range(cx) for j in range(cy)}
        got = {(x, y) for x, y, a in wall[: cx * cy]}
        assert got == grid
        cut = P[8 * lad + 6]
        assert (cx - 0.5, cy - 0.5) not in {(x, y) for x, y, a in cut}
        c7 = P[8 * lad + 7]
        Lb = max(math.sqrt(n) + 0.05, L - 0.02)
        assert {(x, y) for x, y, a in c7[:4]} == {(0.5, 0.5), (Lb - 0.5, 0.5), (0.5, Lb - 0.5), (Lb - 0.5, Lb - 0.5)}
    for lad in range(2, 8):
        assert np.array_equal(P[8 * lad], P[8 * lad + 1])  # topology (lad - 2)


def test_structured_jitter_amplitudes(sis):
    n = 30
    L, q = seed_grid(n)
    base = sis.portfolio(n, L, q, replicas=64, ladder=8, rng_seed=5, jitter=False)
    jit = sis.portfolio(n, L, q, replicas=64, ladder=8, rng_seed=5, jitter=True)
    again = sis.portfolio(n, L, q, replicas=64, ladder=8, rng_seed=5, jitter=True)
    assert np.array_equal(jit, again)
    for r in range(64):
        g = r % 8
        if g == 0:
            assert np.array_equal(jit[r], base[r])
            continue
        # centres move by at most the amplitude, then are clamped to
        # [1/2, L - 1/2] (the clamp is 1-Lipschitz); angles just move
        amp = 1e-2 if g == 1 else 0.03
        d_xy = np.abs(jit[r][:, :2] - np.clip(base[r][:, :2], 0.5, L - 0.5))
        d_a = np.abs(jit[r][:, 2] - base[r][:, 2])
        assert d_xy.max() <= amp + 1e-12 and 0 < d_a.max() <= amp + 1e-15
        assert np.all(jit[r][:, :2] >= 0.5) and np.all(jit[r][:, :2] <= L - 0.5)
        if g >= 2:
            assert d_a.max() > 0.02  # 30 draws of U(-0.03, 0.03)


def test_portfolio_scales_to_other_ladder_sizes(sis):
    """A ladder of 4 maps temperature positions 0..3 to classes 0, 2, 4, 6."""
    n = 12
    L, q = seed_grid(n)
    P4 = sis.portfolio(n, L, q, replicas=16, ladder=4, rng_seed=1, jitter=False)
    assert np.array_equal(P4[0], q)   # class 0 on ladder 0
    assert not np.array_equal(P4[3], q)  # class 6: wall grid
