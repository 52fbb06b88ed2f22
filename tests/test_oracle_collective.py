"""The oracle's model of the collective moves (cuda/packing.cu:564-616) and the
kick (cuda/packing.cu:696-745), pinned by geometry rather than by its own
formulas: a translation mode moves every member of the group by one common
vector, a rotation mode moves the group rigidly (pairwise distances kept, all
angles turned by the same amount), the split mode moves alternate members in
opposite directions, non-members never move, energies agree with an FP64
recompute, and the kick keeps to its amplitudes and relocates the worst
square to an empty probe point."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle_lib import f64, ptr

U23 = 1.0 / 8388608.0


def u01f(r):
    return np.float32((np.float32(r >> 9) + np.float32(0.5)) * np.float32(U23))


def state(rng, n, L):
    x = rng.uniform(0.6, L - 0.6, n).astype(np.float32)
    y = rng.uniform(0.6, L - 0.6, n).astype(np.float32)
    a = rng.uniform(0, math.pi / 2, n)
    return x, y, np.cos(a).astype(np.float32), np.sin(a).astype(np.float32)


def draws(oracle, key, replica, sweep_end, n):
    o1 = oracle.philox([0, replica, sweep_end, 5], key)
    anchor = min(n - 1, int(u01f(o1[0]) * np.float32(n)))
    mode = 1 + min(5, int(u01f(o1[1]) * np.float32(6.0)))
    return anchor, mode


def e64(oracle, x, y, c, s, L):
    x, y = np.asarray(x, np.float64), np.asarray(y, np.float64)
    c, s = np.asarray(c, np.float64), np.asarray(s, np.float64)
    h = 1 / np.sqrt(c * c + s * s)
    c, s = np.ascontiguousarray(c * h), np.ascontiguousarray(s * h)
    return oracle.lib.o_total64_cs(len(x), ptr(np.ascontiguousarray(x), f64), ptr(np.ascontiguousarray(y), f64),
                                   ptr(c, f64), ptr(s, f64), L)


def find_mode(oracle, key, replica, n, mode_wanted):
    for sweep_end in range(1, 4000):
        anchor, mode = draws(oracle, key, replica, sweep_end, n)
        if mode == mode_wanted:
            return sweep_end, anchor
    raise AssertionError("mode never drawn")


@pytest.mark.parametrize("mode", [1, 2, 3, 4, 5, 6])
def test_collective_move_geometry(oracle, mode):
    rng = np.random.default_rng(mode)
    n, L, w = 60, 12.0, 0.7075
    key = [7, 1]
    sweep_end, anchor = find_mode(oracle, key, 3, n, mode)
    x, y, c, s = state(rng, n, L)
    x0, y0, c0, s0 = x.copy(), y.copy(), c.copy(), s.copy()
    # accept everything (infinite temperature): the move is visible in the state
    acc, e = oracle.collective_model(x, y, c, s, L, w, 0.3, 0.3, 0.0, key[0], key[1], 3, sweep_end, 1)
    assert acc == 1
    moved = (x != x0) | (y != y0) | (c != c0) | (s != s0)
    assert moved.any() and moved[anchor] == (mode != 6 or max(abs(c0[anchor]), abs(s0[anchor])) < 0.9992)
    dx, dy = x - x0, y - y0
    clamped = (x <= 0.45) | (x >= L - 0.45) | (y <= 0.45) | (y >= L - 0.45)
    free = moved & ~clamped
    ang = np.arctan2(s0 * c - c0 * s, c0 * c + s0 * s)  # rotation applied to each square
    if moved.sum() > 1:
        assert np.ptp(ang[moved]) < 1e-5  # one angle for the whole group
    if mode in (1, 2, 5):
        assert np.ptp(dx[free]) < 1e-5 and np.ptp(dy[free]) < 1e-5
    elif mode in (3, 4):
        idx = np.flatnonzero(free)
        for a in idx[:8]:
            for b in idx[:8]:
                d0 = math.hypot(x0[a] - x0[b], y0[a] - y0[b])
                d1 = math.hypot(x[a] - x[b], y[a] - y[b])
                assert abs(d0 - d1) < 2e-5
    else:
        idx = np.flatnonzero(free)
        even, odd = idx[idx % 2 == 0], idx[idx % 2 == 1]
        if len(even) and len(odd):
            assert abs(dx[even][0] + dx[odd][0]) < 1e-5 and abs(dy[even][0] + dy[odd][0]) < 1e-5
    # membership: who moved is the mode's group around the anchor
    ax, ay = x0[anchor], y0[anchor]
    if mode == 1:
        assert np.array_equal(moved, np.abs(y0 - ay) < 0.55)
    elif mode == 2:
        assert np.array_equal(moved, np.abs(x0 - ax) < 0.55)
    elif mode == 3:
        assert np.array_equal(moved, (x0 - ax) ** 2 + (y0 - ay) ** 2 < 2.5 + 1e-5) or \
            np.sum(moved ^ ((x0 - ax) ** 2 + (y0 - ay) ** 2 < 2.5)) <= 1
    assert abs(e - e64(oracle, x, y, c, s, L)) <= 1e-5 * max(1.0, e)


def test_collective_metropolis_rejects_restore_state(oracle):
    """At zero temperature a move that raises the energy is undone exactly."""
    rng = np.random.default_rng(11)
    n, L, w = 40, 7.0, 0.7075
    x, y, c, s = state(rng, n, L)
    e_start = oracle.energy32(x, y, c, s, L, w)
    x0, y0 = x.copy(), y.copy()
    acc, e = oracle.collective_model(x, y, c, s, L, w, 0.4, 0.4, 1e30, 5, 0, 0, 17, 16)
    assert e <= e_start
    if acc == 0:
        assert np.array_equal(x, x0) and np.array_equal(y, y0)
    assert e == oracle.energy32(x, y, c, s, L, w)


def test_kick_amplitudes_and_relocation(oracle):
    rng = np.random.default_rng(3)
    n, L, w = 50, 8.0, 0.7075
    x, y, c, s = state(rng, n, L)
    # one square deep inside another: the worst square, far from empty space
    x[7], y[7] = x[8] + 0.1, y[8] + 0.1
    x0, y0, c0, s0 = x.copy(), y.copy(), c.copy(), s.copy()
    k, e = oracle.kick_model(x, y, c, s, L, w, 9, 0, 2, 41, True)
    assert k == 1
    d = np.maximum(np.abs(x - x0), np.abs(y - y0))
    da = np.abs(np.arctan2(s0 * c - c0 * s, c0 * c + s0 * s))
    moved_far = np.flatnonzero(d > 0.08 + 1e-6)
    assert len(moved_far) <= 1  # only the relocated square jumps
    assert np.all(d <= 0.08 + 1e-6) or len(moved_far) == 1
    rest = np.setdiff1d(np.arange(n), moved_far)
    assert np.all(da[rest] <= 0.18 + 1e-5)
    if len(moved_far):
        j = moved_far[0]
        others = np.delete(np.arange(n), j)
        gap = np.min(np.hypot(x[others] - x[j], y[others] - y[j]))
        gap0 = np.min(np.hypot(x0[others] - x0[j], y0[others] - y0[j]))
        assert gap > gap0  # it went to emptier space
    assert abs(e - e64(oracle, x, y, c, s, L)) <= 1e-5 * max(1.0, e)


def test_kick_leaves_zero_energy_replicas(oracle):
    n, L, w = 9, 3.2, 0.7075
    i = np.arange(n)
    x = (1.01 * (i % 3) + 0.55).astype(np.float32)
    y = (1.01 * (i // 3) + 0.55).astype(np.float32)
    c = np.ones(n, np.float32)
    s = np.zeros(n, np.float32)
    x0 = x.copy()
    k, e = oracle.kick_model(x, y, c, s, L, w, 1, 2, 0, 3, True)
    assert k == 0 and e == 0.0 and np.array_equal(x, x0)
