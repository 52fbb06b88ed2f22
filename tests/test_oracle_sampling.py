"""The CPU model of the FP32 sweep (o_sweep32_model) samples exp(-E/T), and each
of its moves' FP32 energy changes matches an FP64 recompute (no GPU).  The
device's sweep equals the model bit for bit (tests/test_gpu_parity.py), so these
properties hold for the device too.

- Detailed balance: one square in a box with walls; cells restrict moves for
  blocks of k sweeps under a random grid shift (k = 1 and 4).  Its time
  averages match the Boltzmann distribution exp(-E_wall/T) computed by
  quadrature (the reference's Metropolis step, cuda/packing.cu:80, 123).
- Per-move energy change (cuda/tests/globaltest.cu:20-38, within 1e-5
  relative): every proposal's FP32 stencil energy change against the
  FP64 change over every pair and the wall, from the same FP32 inputs."""
from __future__ import annotations

import math

import numpy as np
import pytest

LOG2E = 1.4426950408889634


def _exact_wall_moments(L, T, m=20001):
    """Box [0, L] with walls, one unit square: E = ux^2 + uy^2 where ux is the
    overshoot of the square's extent e(theta) = (|cos| + |sin|)/2 past a wall.
    Per axis: Z = (L - 2e) + sqrt(pi T) erf(e / sqrt T),
    W = 2 int_0^e u^2 exp(-u^2/T) du.  <E> = int 2 W Z / int Z^2 over theta, and
    the marginal density of x is int exp(-ux(x)^2/T) Z(theta) / int Z^2."""
    th = (np.arange(m) + 0.5) * (math.pi / 2) / m
    e = 0.5 * (np.abs(np.cos(th)) + np.abs(np.sin(th)))
    from scipy.special import erf
    rt = math.sqrt(T)
    Z = (L - 2 * e) + math.sqrt(math.pi * T) * erf(e / rt)
    W = T * (0.5 * math.sqrt(math.pi * T) * erf(e / rt) - e * np.exp(-e * e / T))
    return float(np.sum(2 * W * Z) / np.sum(Z * Z)), th, e, Z


def _batch_se(v, batches=50):
    k = len(v) // batches
    means = np.asarray(v[: k * batches]).reshape(batches, k).mean(axis=1)
    return means.std(ddof=1) / math.sqrt(batches)


@pytest.mark.parametrize("T", [0.01, 0.04])
@pytest.mark.parametrize("k", [1, 4])
def test_sweep_samples_boltzmann_distribution(oracle, T, k):
    L, w = 1.4, 0.7075
    x = np.array([0.7], np.float32)
    y = np.array([0.7], np.float32)
    c = np.array([1.0], np.float32)
    s = np.array([0.0], np.float32)
    sweeps, burn = 40000, 500
    es, xs = [], []
    for sw in range(sweeps):
        oracle.sweep_model(x, y, c, s, L, w, 0.2, 0.4, LOG2E / T, 11, 3, 0, sw, k)
        if sw >= burn:
            es.append(float(oracle.lib.o_wall32_cs(x[0], y[0], c[0], s[0], L)))
            xs.append(float(x[0]))
    es = np.asarray(es)
    xs = np.asarray(xs)
    mean_e, th, e, Z = _exact_wall_moments(L, T)
    se = _batch_se(es)
    assert abs(es.mean() - mean_e) < 5 * se + 1e-3 * mean_e, (es.mean(), mean_e, se)
    # marginal of x in 14 bins
    edges = np.linspace(0, L, 15)
    grid = np.linspace(0, L, 2801)
    ux = np.maximum(0, np.abs(grid[:, None] - L / 2) + e[None, :] - L / 2)
    dens = (np.exp(-ux * ux / T) * Z[None, :]).sum(axis=1)
    for b in range(14):
        sel = (grid >= edges[b]) & (grid < edges[b + 1])
        p = dens[sel].sum() / dens.sum()
        ind = ((xs >= edges[b]) & (xs < edges[b + 1])).astype(np.float64)
        assert abs(ind.mean() - p) < 5 * _batch_se(ind) + 2e-3, (b, ind.mean(), p)


def _per_move(oracle, x, y, c, s, L, T, sweeps, sxy=0.1, sth=0.1, shift_every=4):
    rows = []
    for sw in range(sweeps):
        _acc, _cnt, (d32, d64, dep) = oracle.sweep_model(x, y, c, s, L, 0.7075, sxy, sth, LOG2E / T, 7, 0, 3, sw,
                                                         shift_every, per_move=True)
        m = np.isfinite(d32)
        assert np.array_equal(m, np.isfinite(d64))
        rows.append(np.stack([d32[m].astype(np.float64), d64[m], dep[m]], 1))
    return np.concatenate(rows)


def _check_moves(r):
    """|dE32 - dE64| <= 1e-5 |dE64| + 2e-6 D for every move, D = the move's
    summed overlap depths (old and new pose, pairs and walls).  The floor: a
    depth computed in FP32 from centres below 64 (their differences exact for
    near pairs) carries an absolute error of a few 1e-7 (measured at most
    6.6e-7 per unit of D), so a squared term is off by about 2 depth x 3e-7;
    where old and new terms nearly cancel, that floor exceeds 1e-5 of the
    change.  Returns (moves with interactions, those within plain 1e-5)."""
    err = np.abs(r[:, 0] - r[:, 1])
    assert np.all(err <= 1e-5 * np.abs(r[:, 1]) + 2e-6 * r[:, 2] + 1e-30)
    # moves that touch nothing change nothing, exactly
    assert np.all(r[r[:, 2] == 0, 0] == 0.0)
    nz = r[:, 2] > 0
    return int((err[nz] <= 1e-5 * np.abs(r[nz, 1])).sum()), int(nz.sum())


def test_per_move_energy_change_on_the_selftest_corpus(oracle, corpus):
    """The reference's selftest configurations (cuda/tests/selftest.cu:33-38,
    50 squares, centres in (-0.2, 12.2), shifted into a box [0, 12.4])."""
    q = corpus["selftest"]
    n_rel = n_all = 0
    for b in range(0, 100, 10):
        x = np.ascontiguousarray(q[b, :, 0] + np.float32(0.2))
        y = np.ascontiguousarray(q[b, :, 1] + np.float32(0.2))
        c = np.ascontiguousarray(np.cos(q[b, :, 2]).astype(np.float32))
        s = np.ascontiguousarray(np.sin(q[b, :, 2]).astype(np.float32))
        a, t = _check_moves(_per_move(oracle, x, y, c, s, 12.4, 0.05, 4))
        n_rel += a
        n_all += t
    assert n_all > 1000 and n_rel >= 0.8 * n_all


@pytest.mark.parametrize("n,L,T", [(50, 7.5, 0.05), (50, 9.0, 1e-3), (200, 14.0, 0.01)])
def test_per_move_energy_change_dense_random(oracle, n, L, T):
    """Random starts at and above the plain GPU workload's density (n = 50 in
    the box of side 9 and 7.5)."""
    rng = np.random.default_rng(n)
    x = rng.uniform(0.3, L - 0.3, n).astype(np.float32)
    y = rng.uniform(0.3, L - 0.3, n).astype(np.float32)
    a = rng.uniform(0, math.pi / 2, n)
    c, s = np.cos(a).astype(np.float32), np.sin(a).astype(np.float32)
    n_rel, n_all = _check_moves(_per_move(oracle, x, y, c, s, L, T, 8))
    assert n_all > 300 and n_rel >= 0.8 * n_all


def test_per_move_energy_change_precision_workload(oracle):
    """The precision workload's start: n = 500 on the grid of side 23 with
    +-0.02 centre and +-0.01 rad angle jitter, one sweep at two temperatures.
    Neighbours sit at contact, so a move changes small overlaps into similar
    small ones and about 40 % of the changes fall under the FP32 floor."""
    rng = np.random.default_rng(500)
    n, side = 500, 23
    k = np.arange(n)
    x = ((k % side) + 0.5 + rng.uniform(-0.02, 0.02, n)).astype(np.float32)
    y = ((k // side) + 0.5 + rng.uniform(-0.02, 0.02, n)).astype(np.float32)
    a = rng.uniform(-0.01, 0.01, n)
    c, s = np.cos(a).astype(np.float32), np.sin(a).astype(np.float32)
    for T in (1e-4, 1e-2):
        n_rel, n_all = _check_moves(_per_move(oracle, x.copy(), y.copy(), c.copy(), s.copy(), 23.0, T, 1,
                                              sxy=0.02, sth=0.01))
        assert n_all > 400 and n_rel >= 0.5 * n_all
