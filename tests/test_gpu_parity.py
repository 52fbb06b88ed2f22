"""GPU parity: the gfx950 kernels, called through the C ABI, against the CPU
oracle on the reference's corpora and seeded configurations."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle_lib import f32, f64, ptr

pytestmark = pytest.mark.gpu

LOG2E = 1.4426950408889634


def say(*a):
    print("[gpu-case]", *a, flush=True)


def _corpus_pairs(q):
    """All ordered pairs of every selftest block as (dx, dy, ci, si, cj, sj)."""
    B, n, _ = q.shape
    c = np.cos(q[:, :, 2]).astype(np.float32)
    s = np.sin(q[:, :, 2]).astype(np.float32)
    rows = []
    for b in range(B):
        dx = (q[b, None, :, 0] - q[b, :, None, 0]).astype(np.float32)  # x_j - x_i
        dy = (q[b, None, :, 1] - q[b, :, None, 1]).astype(np.float32)
        ci = np.broadcast_to(c[b, :, None], (n, n))
        si = np.broadcast_to(s[b, :, None], (n, n))
        cj = np.broadcast_to(c[b, None, :], (n, n))
        sj = np.broadcast_to(s[b, None, :], (n, n))
        rows.append(np.stack([dx, dy, ci, si, cj, sj], axis=-1).reshape(-1, 6))
    return np.ascontiguousarray(np.concatenate(rows), dtype=np.float32)


def test_depth32_bit_exact_and_fp64_tolerance(sis, oracle, corpus):
    """250 000 selftest pairs: device == oracle restatement bit for bit; both
    within 2e-6 of the FP64 axis oracle (cuda/tests/selftest.cu:53)."""
    q = corpus["selftest"]
    pairs = _corpus_pairs(q)
    say("depth32", pairs.shape)
    d, d2 = sis.dev_depth32(pairs)
    L = oracle.lib
    idx = np.random.default_rng(0).choice(len(pairs), 40000, replace=False)
    for k in idx:
        row = [float(v) for v in pairs[k]]
        assert d[k] == np.float32(L.o_depth32_cs(*row))
        assert d2[k] == np.float32(L.o_depth2_32_cs(*row))
    worst = 0.0
    B, n, _ = q.shape
    for b in range(0, B, 10):
        for i in range(n):
            for j in range(n):
                if i == j:
                    continue
                pi = np.ascontiguousarray(q[b, i]); pj = np.ascontiguousarray(q[b, j])
                ref = L.o_sat_oracle_depth_f(ptr(pi, f32), ptr(pj, f32))
                worst = max(worst, abs(float(d[(b * n + i) * n + j]) - ref))
    assert worst <= 2e-6, worst
    # edge fixtures of selftest.cu:38: touch / separated / overlapping
    assert d[1] == 0.0 and d[2] == 0.0 and d[3] > 0.0


def test_cell_binning_bit_exact(sis, oracle):
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-3, 40, 20000), np.arange(0, 32, 0.25), [0.0, 2.0, 1.9999999, 31.999998]])
    x = x.astype(np.float32)
    say("cellcoord", x.shape)
    ref = sis.dev_cellcoord(x, 0.0, 2.0, 0, 15)  # the reference's binning
    for k in range(len(x)):
        assert ref[k] == min(15, max(0, math.floor(float(x[k]) / 2)))
    for shift, w in ((0.0, 0.7075), (0.31, 0.7075), (0.7, 0.75), (0.05, 1.0)):
        out = sis.dev_cellcoord(x, shift, w, 0, 60)
        for k in range(0, len(x), 7):
            assert out[k] == oracle.lib.o_cellcoord(x[k], shift, w, 0, 60)


def test_accept_exp2_bit_exact(sis, oracle):
    """The sweep's 2^x for the Metropolis test equals the oracle's FP32 form
    bit for bit, round-to-nearest ties, -0 and the cut-off at -100 included."""
    from test_oracle_geometry import _accept_exp2_inputs
    x = _accept_exp2_inputs(np.random.default_rng(5))
    say("accept_exp2", x.shape)
    d = sis.dev_accept_exp2(x)
    f = oracle.lib.o_accept_exp2
    ref = np.array([f(float(v)) for v in x], dtype=np.float32)
    assert np.array_equal(d.view(np.uint32), ref.view(np.uint32))


def _seeded_state(rng, R, n, L):
    x = rng.uniform(0.2, L - 0.2, (R, n))
    y = rng.uniform(0.2, L - 0.2, (R, n))
    a = rng.uniform(0, math.pi / 2, (R, n))
    st = np.stack([x, y, np.cos(a), np.sin(a)], axis=-1).astype(np.float32)
    return np.ascontiguousarray(st)


def _model_run(oracle, st, L, w, ik, sxy, sth, key0, key1, sweep0, sweeps, shift_every=4):
    st = st.copy()
    R, n, _ = st.shape
    acc = np.zeros(R, np.int64)
    cnt = np.zeros(3, np.uint64)
    en = np.zeros(R, np.float32)
    for r in range(R):
        x, y, c, s = (np.ascontiguousarray(st[r, :, k]) for k in range(4))
        for sw in range(sweep0, sweep0 + sweeps):
            a, cc = oracle.sweep_model(x, y, c, s, L, w, sxy[r], sth[r], ik[r], key0, key1, r, sw, shift_every)
            acc[r] += a
            cnt += cc
        st[r, :, 0], st[r, :, 1], st[r, :, 2], st[r, :, 3] = x, y, c, s
        en[r] = oracle.energy32(x, y, c, s, L, w)
    return st, acc, en, cnt


@pytest.mark.parametrize("n,L,R,sweeps,group", [(11, 5.0, 4, 6, 1), (50, 9.0, 6, 5, 3), (90, 11.0, 3, 3, 1),
                                                (20, 6.0, 6, 4, 3), (10, 4.5, 8, 5, 6), (16, 5.0, 4, 4, 4),
                                                (64, 10.0, 2, 3, 1), (60, 3.0, 2, 3, 1),
                                                (100, 11.0, 4, 3, 4), (120, 12.5, 3, 3, 2),
                                                (300, 17.5, 2, 3, 1), (300, 17.5, 4, 2, 2),
                                                (400, 21.0, 2, 2, 2), (700, 27.0, 1, 2, 1)])
def test_sweeps_match_cpu_model(sis, oracle, n, L, R, sweeps, group):
    """k_sweep == the CPU model of the same colour-checkerboard sweep: same
    states, accept counts and canonical energies, bit for bit.  Below n = 96
    one-wavefront tiles: mover records with the register-resident set-up (up
    to 64 squares per tile, one or several replicas, n = 60 in a 3 x 3 box
    filling the serial group of ranks >= 4), mover records with the LDS
    set-up (n = 90 alone) and staged proposal arrays (150 squares in tiles of
    three replicas).  n >= 96 runs the four-wavefront tiles, with the
    proposals staged in LDS (up to 512 squares per tile, and 700 in a tile of
    one replica) and drawn inline (600 squares in tiles of two replicas)."""
    rng = np.random.default_rng(n)
    st = _seeded_state(rng, R, n, L)
    T = np.geomspace(1e-3, 1e-1, R)
    ik = (LOG2E / T).astype(np.float32)
    sxy = np.full(R, 0.12, np.float32)
    sth = np.full(R, 0.1, np.float32)
    say("sweeps", n, L, R, sweeps, group)
    dst, dacc, den, dcnt = sis.dev_sweeps(st, L, 0.7075, ik, sxy, sth, 7, 0, 11, sweeps, group=group)
    mst, macc, men, mcnt = _model_run(oracle, st, L, 0.7075, ik, sxy, sth, 7, 0, 11, sweeps)
    assert np.array_equal(dacc, macc)
    assert np.array_equal(dst, mst)
    assert np.array_equal(den, men)
    assert dcnt[0] == R * n * sweeps and np.array_equal(dcnt, mcnt)


@pytest.mark.parametrize("k", [1, 3, 8])
@pytest.mark.parametrize("n,L,R,group", [(20, 6.0, 6, 3), (90, 11.0, 3, 1), (50, 9.0, 6, 3), (150, 14.0, 4, 2),
                                         (300, 17.5, 4, 2)])
def test_sweeps_match_cpu_model_for_any_shift_period(sis, oracle, n, L, R, group, k):
    """Blocks of k sweeps share one grid shift; the sweeps after a block's
    first keep its cell lists, ranks and levels and only draw new moves.  The
    model builds everything from scratch every sweep with the same shift, so
    equality holds only if no kept list went stale.  Seven sweeps from sweep 5
    (a launch starting inside a block) on every set-up path: register set-up
    with records (60 squares), LDS set-up with records (90), staged arrays in
    one-wavefront tiles (150), four-wavefront tiles staged (300) and inline
    (600)."""
    rng = np.random.default_rng(500 + n + k)
    st = _seeded_state(rng, R, n, L)
    T = np.geomspace(1e-3, 1e-1, R)
    ik = (LOG2E / T).astype(np.float32)
    sxy = np.full(R, 0.12, np.float32)
    sth = np.full(R, 0.1, np.float32)
    say("sweeps shift period", n, L, R, group, k)
    dst, dacc, den, dcnt = sis.dev_sweeps(st, L, 0.7075, ik, sxy, sth, 7, 0, 5, 7, group=group, shift_every=k)
    mst, macc, men, mcnt = _model_run(oracle, st, L, 0.7075, ik, sxy, sth, 7, 0, 5, 7, shift_every=k)
    assert np.array_equal(dacc, macc)
    assert np.array_equal(dst, mst)
    assert np.array_equal(den, men)
    assert np.array_equal(dcnt, mcnt)


@pytest.mark.parametrize("cols,rows,R,group", [(8, 8, 2, 1), (8, 5, 2, 1), (6, 4, 3, 1), (4, 3, 4, 1),
                                              (4, 4, 4, 2), (8, 4, 2, 2)])
def test_sweeps_match_cpu_model_with_crowded_levels(sis, oracle, cols, rows, R, group):
    """Squares on a lattice of pitch three cells all share one colour class
    under any grid shift, so one level holds (nearly) every square of a tile:
    12 to 64 movers in one level of a one-wavefront tile with mover records,
    which runs the record pass at 4, 2 and 1 lanes per mover (random packings
    of these sizes never have more than about 15).  Device == model bit for
    bit."""
    w = 0.7075
    pitch = 3.0 * w
    n = cols * rows
    L = 2 * 1.2 + pitch * (max(cols, rows) - 1)
    rng = np.random.default_rng(900 + n + group)
    gx, gy = np.meshgrid(1.2 + pitch * np.arange(cols), 1.2 + pitch * np.arange(rows))
    x = np.broadcast_to(gx.ravel(), (R, n)) + rng.uniform(-0.01, 0.01, (R, n))
    y = np.broadcast_to(gy.ravel(), (R, n)) + rng.uniform(-0.01, 0.01, (R, n))
    a = rng.uniform(0, math.pi / 2, (R, n))
    st = np.ascontiguousarray(np.stack([x, y, np.cos(a), np.sin(a)], axis=-1).astype(np.float32))
    T = np.geomspace(1e-3, 1e-1, R)
    ik = (LOG2E / T).astype(np.float32)
    # short steps: the squares stay near the lattice, so the levels stay crowded
    sxy = np.full(R, 0.02, np.float32)
    sth = np.full(R, 0.1, np.float32)
    say("sweeps crowded levels", cols, rows, R, group)
    dst, dacc, den, dcnt = sis.dev_sweeps(st, L, w, ik, sxy, sth, 7, 0, 3, 6, group=group)
    mst, macc, men, mcnt = _model_run(oracle, st, L, w, ik, sxy, sth, 7, 0, 3, 6)
    assert np.array_equal(dacc, macc)
    assert np.array_equal(dst, mst)
    assert np.array_equal(den, men)
    assert np.array_equal(dcnt, mcnt)


@pytest.mark.parametrize("n,L,R,sweeps,group,w", [(1, 1.5, 4, 5, 1, 0.7075), (2, 2.5, 4, 5, 2, 0.7075),
                                                  (7, 5.0, 3, 4, 1, 2.0), (30, 8.0, 2, 3, 1, 4.0),
                                                  (1000, 64.0, 1, 1, 1, 0.7075)])
def test_sweeps_match_cpu_model_at_the_limits(sis, oracle, n, L, R, sweeps, group, w):
    """The input limits of cuda/packing.cu:772-775 (n 1..1000, L up to 64) and the
    cell widths the library accepts (the reference's own width 2, the maximum 4):
    one and two squares (a stencil holding only the mover), n = 1000 in the
    largest box (96 cells per side, sparse levels), wide cells with many
    squares per cell (occupancy ranks up to the serial group)."""
    rng = np.random.default_rng(1000 + n)
    st = _seeded_state(rng, R, n, L)
    T = np.geomspace(1e-3, 1e-1, R)
    ik = (LOG2E / T).astype(np.float32)
    sxy = np.full(R, 0.12, np.float32)
    sth = np.full(R, 0.1, np.float32)
    say("sweeps at the limits", n, L, R, sweeps, group, w)
    dst, dacc, den, dcnt = sis.dev_sweeps(st, L, w, ik, sxy, sth, 7, 0, 11, sweeps, group=group)
    mst, macc, men, mcnt = _model_run(oracle, st, L, w, ik, sxy, sth, 7, 0, 11, sweeps)
    assert np.array_equal(dacc, macc)
    assert np.array_equal(dst, mst)
    assert np.array_equal(den, men)
    assert dcnt[0] == R * n * sweeps and np.array_equal(dcnt, mcnt)


@pytest.mark.parametrize("n,L,R,sweeps,group,moves,kick_every,interval", [
    (40, 8.0, 6, 3, 2, 4, 1, 0),      # kick 1: relocating
    (40, 8.0, 4, 2, 1, 2, 1, 1),      # kick 2: jitter only
    (300, 17.5, 2, 2, 1, 3, 2, 1),    # four-wavefront tiles, relocating kick
])
def test_collective_moves_and_kick_match_cpu_model(sis, oracle, n, L, R, sweeps, group, moves, kick_every,
                                                   interval):
    """Sweeps, then collective moves (cuda/packing.cu:564-616), then the kick
    (:696-745) on the device == the oracle's model, bit for bit: states,
    energies and the collective / kick counters."""
    rng = np.random.default_rng(n + moves)
    st = _seeded_state(rng, R, n, L)
    T = np.geomspace(1e-3, 1e-1, R)
    ik = (LOG2E / T).astype(np.float32)
    sxy = np.full(R, 0.12, np.float32)
    sth = np.full(R, 0.1, np.float32)
    say("collective+kick", n, R, moves, kick_every, interval)
    dst, dacc, den, dcnt = sis.dev_sweeps(st, L, 0.7075, ik, sxy, sth, 7, 0, 11, sweeps, group=group,
                                          collective=moves, kick_every=kick_every, interval=interval)
    mst, macc, _, _ = _model_run(oracle, st, L, 0.7075, ik, sxy, sth, 7, 0, 11, sweeps)
    coll_acc, kicked = 0, 0
    men = np.zeros(R, np.float32)
    relocate = ((interval + 1) // kick_every) % 2 == 1
    for r in range(R):
        x, y, c, s = (np.ascontiguousarray(mst[r, :, k]) for k in range(4))
        a, e = oracle.collective_model(x, y, c, s, L, 0.7075, sxy[r], sth[r], ik[r], 7, 0, r, 11 + sweeps, moves)
        coll_acc += a
        if (interval + 1) % kick_every == 0:
            k, e = oracle.kick_model(x, y, c, s, L, 0.7075, 7, 0, r, interval, relocate)
            kicked += k
        mst[r, :, 0], mst[r, :, 1], mst[r, :, 2], mst[r, :, 3] = x, y, c, s
        men[r] = e
    assert np.array_equal(dacc, macc)
    assert np.array_equal(dst, mst)
    assert np.array_equal(den, men)
    assert dcnt[3] == R * moves and dcnt[4] == coll_acc and dcnt[5] == kicked
    assert kicked > 0


def test_sweep_energy_vs_fp64_recompute(sis, oracle):
    """Energy after device sweeps equals the FP64 all-pairs recompute of the
    same state within 1e-5 relative (cuda/tests/globaltest.cu:49)."""
    rng = np.random.default_rng(4)
    R, n, L = 8, 80, 10.0
    st = _seeded_state(rng, R, n, L)
    T = np.geomspace(1e-4, 5e-2, R)
    ik = (LOG2E / T).astype(np.float32)
    say("energy-vs-fp64", R, n)
    for rnd in range(3):
        st, acc, en, _ = sis.dev_sweeps(st, L, 0.7075, ik, np.full(R, 0.1, np.float32), np.full(R, 0.08, np.float32),
                                        3, 1, rnd * 4, 4, group=4)
        for r in range(R):
            x, y, c, s = (np.ascontiguousarray(st[r, :, k], np.float64) for k in range(4))
            h = 1 / np.sqrt(c * c + s * s)
            c, s = np.ascontiguousarray(c * h), np.ascontiguousarray(s * h)
            e64 = oracle.lib.o_total64_cs(n, ptr(x, f64), ptr(y, f64), ptr(c, f64), ptr(s, f64), L)
            assert abs(en[r] - e64) <= 1e-5 * max(1.0, e64), (r, en[r], e64)


def _xycs(x, y, a):
    return np.ascontiguousarray(np.stack([x, y, np.cos(a), np.sin(a)], axis=-1).astype(np.float32))


def _o_total64_xycs(oracle, xycs, L):
    x, y, c, s = (np.ascontiguousarray(xycs[:, k], np.float64) for k in range(4))
    h = 1 / np.sqrt(c * c + s * s)
    c, s = np.ascontiguousarray(c * h), np.ascontiguousarray(s * h)
    return oracle.lib.o_total64_cs(len(x), ptr(x, f64), ptr(y, f64), ptr(c, f64), ptr(s, f64), L)


def test_zero_overlap_grid_test_equals_all_pairs(sis, oracle):
    """The exchange kernel's FP64 zero-overlap test (a grid of width 1.5)
    accepts a configuration exactly when the oracle's all-pairs FP64 energy of
    cuda/packing.cu:625-632 is zero, and the device's all-pairs energy equals
    the oracle's: touching grids, grids overlapping by 1e-9, one displaced
    square (every neighbour direction, across cell boundaries), squares through
    a wall, the reference's own raw outputs and random dense configurations."""
    import json
    from conftest import ROOT
    cases = []
    for n, pitch in ((25, 1.0), (25, 1.0 - 1e-9), (1000, 1.000001), (1000, 1.0), (400, 1.0 - 1e-7)):
        k = int(math.isqrt(n - 1)) + 1
        i = np.arange(n)
        x, y = (i % k + 0.5) * pitch, (i // k + 0.5) * pitch
        cases.append((_xycs(x, y, np.zeros(n)), k * pitch))
    rng = np.random.default_rng(17)
    for trial in range(60):
        # a touching 12 x 12 grid at pitch 1.5 with one square pushed towards a neighbour
        k, pitch = 12, 1.5
        n = k * k
        i = np.arange(n)
        x, y = (i % k + 0.75) * pitch - 0.375, (i // k + 0.75) * pitch - 0.375
        a = rng.uniform(0, 0.02, n)
        j = int(rng.integers(0, n))
        ang = rng.uniform(0, 2 * math.pi)
        d = rng.choice([0.2, 0.499999, 0.5, 0.500001, 0.7])
        x[j] += d * math.cos(ang)
        y[j] += d * math.sin(ang)
        cases.append((_xycs(x, y, a), k * pitch + 0.2))
    for trial in range(20):
        n = int(rng.integers(2, 300))
        L = math.sqrt(n) * rng.uniform(1.0, 2.5)
        x, y = rng.uniform(0.5, L - 0.5, n), rng.uniform(0.5, L - 0.5, n)
        cases.append((_xycs(x, y, rng.uniform(0, math.pi / 2, n)), L))
    x = np.array([0.2, 3.0, 5.5]); y = np.array([0.5, 3.0, 5.5])
    cases.append((_xycs(x, y, np.zeros(3)), 6.0))  # first square through the wall
    for nn in (5, 10, 11, 17, 50):
        raw = json.loads((ROOT / "tests/golden/raw" / f"raw{nn}.json").read_text())
        q = np.asarray(raw["q"], dtype=np.float64)
        cases.append((_xycs(q[:, 0], q[:, 1], q[:, 2]), float(raw["target_L"])))
    zeros = nonzero = 0
    for k, (xycs, L) in enumerate(cases):
        say("zero64", k, xycs.shape[0], L)
        z, e = sis.dev_zero64(xycs, L)
        eo = _o_total64_xycs(oracle, xycs, L)
        assert z == (eo == 0.0), (k, z, eo)
        assert abs(e - eo) <= 1e-12 * eo, (k, e, eo)
        zeros += z
        nonzero += not z
    assert zeros >= 3 and nonzero >= 10, (zeros, nonzero)


def test_replicas_per_group_do_not_change_results(sis):
    """Determinism: the tile shape (replicas per workgroup) never changes any
    replica's arithmetic."""
    rng = np.random.default_rng(9)
    R, n, L = 12, 40, 7.5
    st = _seeded_state(rng, R, n, L)
    ik = (LOG2E / np.geomspace(1e-3, 1e-1, R)).astype(np.float32)
    outs = []
    for g in (1, 4, 12):
        say("group", g)
        outs.append(sis.dev_sweeps(st, L, 0.7075, ik, np.full(R, 0.1, np.float32), np.full(R, 0.1, np.float32),
                                   5, 6, 0, 8, group=g))
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("pitch,L", [(1.000001, 32.0001), (0.99, 31.7)])
def test_globaltest_grid_n1000(sis, oracle, pitch, L):
    """cuda/tests/globaltest.cu:6-38 analogue: the n=1000 grid (pitch 1.000001
    in L=32.0001, and compressed to pitch 0.99 so every neighbour pair and
    diagonal overlaps), 8 replicas, sweeps; the device's FP32 energy equals the
    FP64 host recompute within 1e-5 relative (cuda/tests/globaltest.cu:49 uses
    1e-5 absolute)."""
    n, R = 1000, 8
    i = np.arange(n)
    st = np.zeros((R, n, 4), np.float32)
    st[:, :, 0] = (i % 32 + 0.5) * pitch
    st[:, :, 1] = (i // 32 + 0.5) * pitch
    st[:, :, 2] = 1.0
    ik = (LOG2E / np.geomspace(1e-7, 1e-2, R)).astype(np.float32)
    say("globaltest n=1000")
    st2, acc, en, cnt = sis.dev_sweeps(st, L, 0.7075, ik, np.full(R, 0.015, np.float32),
                                       np.full(R, 0.02, np.float32), 142, 0, 0, 3, group=1)
    assert cnt[0] == 3 * n * R
    for r in range(R):
        x, y, c, s = (np.ascontiguousarray(st2[r, :, k], np.float64) for k in range(4))
        h = 1 / np.sqrt(c * c + s * s)
        c, s = np.ascontiguousarray(c * h), np.ascontiguousarray(s * h)
        e64 = oracle.lib.o_total64_cs(n, ptr(x, f64), ptr(y, f64), ptr(c, f64), ptr(s, f64), L)
        assert abs(en[r] - e64) <= 1e-5 * e64, (r, en[r], e64)
        if pitch < 1:
            assert e64 > 0.1


def _expected_export(oracle, q0, oq):
    """The export rule: the certified side before the polish unless the one
    after it is smaller (cuda/packing.cu:506-507, 689-693)."""
    s0, l0, o0 = oracle.certified_side(q0)
    s1, l1, o1 = oracle.certified_side(oq)
    if s1 < s0 or (s0 >= 1e8 and s1 >= 1e8):
        return s1, l1, o1
    return s0, l0, o0


@pytest.mark.parametrize("n,k,iters", [(16, 4, 200), (30, 6, 200), (300, 18, 20), (500, 23, 20), (1000, 32, 20)])
def test_polish_and_certified_side_match_oracle(sis, oracle, n, k, iters):
    """FP64 polish (cuda/packing.cu:624-687) and certified side (:254-297) on the
    device against the oracle.  n = 300, 500 and 1000 start from the precision
    workload's grid of side ceil(sqrt n) with +-0.02 centre and +-0.01 rad
    jitter; above 256 squares every thread of the polish workgroup holds
    several squares."""
    rng = np.random.default_rng(2 + n)
    if n <= 30:
        grid = np.zeros((n, 3))
        oracle.lib.o_grid_packing(n, ptr(grid, f64))
        q = grid.copy()
        q[:, :2] = q[:, :2] * 0.99 + rng.uniform(-0.01, 0.01, (n, 2))
        q[:, 2] = rng.uniform(0, 0.03, n)
        L = k * 0.995
    else:
        i = np.arange(n)
        q = np.zeros((n, 3))
        q[:, 0] = i % k + 0.5 + rng.uniform(-0.02, 0.02, n)
        q[:, 1] = i // k + 0.5 + rng.uniform(-0.02, 0.02, n)
        q[:, 2] = np.mod(rng.uniform(-0.01, 0.01, n), math.pi / 2)
        L = float(k)
    say("polish", n, iters)
    dev = sis.dev_polish(q, L, iters)
    oq, oe, _ = oracle.polish64(q, L, iters)
    assert abs(dev["e_before"] - oracle.total64(q, L)) <= 1e-12 * max(1, dev["e_before"])
    assert dev["e_after"] <= dev["e_before"]
    assert abs(dev["e_after"] - oe) <= 1e-9 * max(1e-12, oe) + 1e-20
    side, lam, out = _expected_export(oracle, q, oq)
    assert abs(dev["lam"] - lam) <= 1e-9
    if side < 1e8:
        assert abs(dev["L_cert"] - side) <= 1e-9
        assert np.max(np.abs(dev["q"] - out)) <= 1e-9


def test_solve_n11_certifies(sis, oracle, ref_certifier):
    """End to end (the n=11, 8-replica CPU workload's shape on the GPU): the exported packing is
    overlap-free in FP64 and passes the reference's interval certifier after
    the dilation ladder of cuda/demo.py:31."""
    from refcert import certify_q
    say("solve n=11")
    r = sis.solve(11, 5.0, None, replicas=8, tmin=1e-4, tmax=1e-1, sweeps=4000, sweeps_per_exchange=10,
                  shrink=1e-3, rng_seed=42, init_mode=1)
    assert r.proposals == 8 * 11 * 4000
    assert r.feasible == 1 and r.shrinks > 0
    assert math.sqrt(11) - 1e-7 <= r.L_cert < 5.0
    q = r.q
    for i in range(11):
        for j in range(i + 1, 11):
            assert oracle.lib.o_sat_oracle64(ptr(q[i].copy(), f64), ptr(q[j].copy(), f64)) == 0.0
    dil, L50 = certify_q(ref_certifier, q)
    assert dil is not None
    assert float(L50) <= r.L_cert * dil + 1e-9


def test_portfolio_start_reaches_goebel_side(sis, oracle, ref_certifier):
    """init_mode 2 (portfolio starts, as the reference's `packing` seeds its
    chains): n = 10 from the 4 x 4 grid seed reaches Goebel's side 3 + 1/sqrt2
    (one of the starting topologies) within the box-shrink budget, and the
    export certifies."""
    from refcert import certify_q
    n = 10
    i = np.arange(n)
    q = np.stack([i % 4 + 0.5, i // 4 + 0.5, np.zeros(n)], axis=1)
    say("portfolio n=10")
    r = sis.solve(n, 4.0, q, replicas=64, ladder=8, sweeps=3000, shrink=1e-2, rng_seed=9, init_mode=2,
                  tmin=1e-6, tmax=1e-2)
    assert r.feasible == 1
    assert r.L_cert <= 3 + 1 / math.sqrt(2) + 0.02
    for a in range(n):
        for b in range(a + 1, n):
            assert oracle.lib.o_sat_oracle64(ptr(r.q[a].copy(), f64), ptr(r.q[b].copy(), f64)) == 0.0
    dil, L50 = certify_q(ref_certifier, r.q)
    assert dil is not None


def test_batch_equals_alone(sis):
    """Alone vs in the mix: each instance's result is bit-identical whether it
    runs alone (plain) or inside a batch with others of different n."""
    # n = 270 runs on four-wavefront tiles, launched beside the others
    insts = [(13, 5.0, None), (37, 8.0, None), (60, 9.5, None), (270, 17.0, None)]
    ps = [sis.default_params(replicas=r, sweeps=s, rng_seed=seed, init_mode=1, tmin=1e-4, tmax=3e-2)
          for r, s, seed in ((16, 300, 1), (32, 200, 2), (8, 250, 3), (6, 40, 4))]
    say("batch")
    mix = sis.solve_batch(insts, ps)
    for (n, L, q), p, m in zip(insts, ps, mix):
        say("alone", n)
        a = sis.solve(n, L, q, params=p)
        assert np.array_equal(a.q, m.q)
        for k in ("proposals", "accepted", "shrinks", "exchanges_accepted", "L_cert", "target_L", "e64_after"):
            assert a.fields[k] == m.fields[k], k


def test_plain_runs_stop_at_the_budget_and_read_status(sis):
    """The plain schedule waits for every interval's end (one interval queued
    ahead) and reads the instances' status at the end of a run:
    a run longer than the budget stops at it, a run after the end launches
    nothing, the status the host read equals the device's descriptor, every
    eighth launch of a call is timed, and stopping early changes no result
    (same packing as the graph schedule run to the end)."""
    insts = [(13, 5.0, None), (37, 8.0, None)]
    ps = [sis.default_params(replicas=8, sweeps=s, rng_seed=seed, init_mode=1, tmin=1e-4, tmax=3e-2, improve=1)
          for s, seed in ((120, 11), (45, 12))]  # 12 and 5 intervals of 10 sweeps
    say("plain early stop")
    b = sis.BatchSession(insts, ps)
    assert [b.status(k) for k in range(2)] == [b.state(k)["desc"] for k in range(2)]  # before any run
    sm, launches, _ = b.run(3, mode=0)
    assert launches == 1 and sm > 0  # intervals 0 of the call timed (every eighth)
    st = [b.status(k) for k in range(2)]
    assert [d["sweeps_done"] for d in st] == [30, 30] and [d["interval"] for d in st] == [3, 3]
    assert st == [b.state(k)["desc"] for k in range(2)]
    _, launches, _ = b.run(100, mode=0)  # 9 more intervals end both budgets
    assert launches == 2  # calls' intervals 0 and 8 timed
    st = [b.status(k) for k in range(2)]
    assert [d["sweeps_done"] for d in st] == [120, 45] and [d["interval"] for d in st] == [12, 5]
    assert st == [b.state(k)["desc"] for k in range(2)]
    _, launches, _ = b.run(5, mode=0)
    assert launches == 0
    res = b.finalize()
    b.close()
    ref = sis.solve_batch(insts, ps)
    for r, m in zip(res, ref):
        assert np.array_equal(r.q, m.q) and r.L_cert == m.L_cert and r.proposals == m.proposals


def test_near_pair_counting_changes_nothing_else(sis):
    """SIS_FLAG_NEAR_PAIRS only adds counting: the same packing, counters and
    certified side with and without it (plain, batched with a four-wavefront
    instance, and allocation-free forms); near pairs are counted only with it.
    Each near pair is a candidate seen from one of two positions, so with it
    0 < near_pairs <= 2 candidates."""
    insts = [(37, 8.0, None), (120, 13.0, None)]
    ps = [sis.default_params(replicas=r, sweeps=s, rng_seed=seed, init_mode=1, tmin=1e-4, tmax=3e-2,
                             collective_moves=2, kick_every=3, improve=1)
          for r, s, seed in ((16, 200, 5), (8, 60, 6))]
    say("near pairs: batch")
    off = sis.solve_batch(insts, ps)
    on = sis.solve_batch(insts, [sis.with_near_pairs(p) for p in ps])
    say("near pairs: alone / no-alloc")
    alone = sis.solve(*insts[0], params=sis.with_near_pairs(ps[0]))
    noalloc = sis.solve_noalloc(*insts[0], params=sis.with_near_pairs(ps[0]))
    for a, b in list(zip(off, on)) + [(off[0], alone), (off[0], noalloc)]:
        assert np.array_equal(a.q, b.q)
        for k in ("proposals", "accepted", "candidates", "collective_tried", "collective_accepted", "kicks",
                  "exchanges_accepted", "shrinks", "L_cert", "target_L", "feasible_checkpoints"):
            assert a.fields[k] == b.fields[k], k
        assert a.near_pairs == 0 and 0 < b.near_pairs <= 2 * b.candidates
    assert alone.near_pairs == on[0].near_pairs == noalloc.near_pairs
    with pytest.raises(sis.SisError):
        sis.solve(*insts[0], params=sis.default_params(flags=2, sweeps=10))


def test_autotune_returns_a_candidate_that_runs(sis):
    """The autotuner tries every (tile size, sweeps per exchange, cell width)
    candidate on the batch; its choice is one of the candidates and a batch
    run with it proposes the requested sweeps.  A choice that changes the
    trajectory passed the quality probe: its reported mean paired loss of
    relative side reduction is what two runs of the batch (the candidate and
    the caller's settings, polish off, to the same sweep) give, and with two
    standard errors added is at most 0.001; a kept trajectory after a check
    means the last candidate checked failed that bound."""
    import math

    import numpy as np

    say("autotune")
    inst = [(30, 7.0, None), (120, 13.0, None), (60, 9.5, None), (45, 8.0, None), (80, 10.0, None)]
    kw = dict(replicas=16, sweeps=100, init_mode=1, tmin=1e-4, tmax=3e-2, improve=1, shrink=1e-3)
    p = sis.default_params(**kw)
    t = sis.autotune(inst, p, probe_sweeps=20)
    say("autotune", t)
    assert t["candidates_tried"] == 30
    assert t["replicas_per_group"] in (0, 1, 2, 4, 8)
    assert t["sweeps_per_exchange"] in (5, 10, 20)
    assert t["cell"] in (0.7075, 0.8)
    assert t["proposals_per_s"] > 0
    assert 0 <= t["quality_checked"] <= 3
    if t["trajectory_kept"]:
        assert (t["sweeps_per_exchange"], t["cell"]) == (p.sweeps_per_exchange, p.cell)
    else:
        assert t["quality_checked"] >= 1
        assert (t["checked_sweeps_per_exchange"], t["checked_cell"]) == (t["sweeps_per_exchange"], t["cell"])
    if t["quality_checked"]:
        # the last candidate checked, recomputed from two runs of the batch
        assert t["quality_sweeps"] == 1000
        spe, cell = t["checked_sweeps_per_exchange"], t["checked_cell"]
        lcm = spe * p.sweeps_per_exchange // math.gcd(spe, p.sweeps_per_exchange)
        sw = max(lcm, 1000 // lcm * lcm)
        own = sis.solve_batch(inst, sis.default_params(**{**kw, "sweeps": sw, "polish_iters": 0}))
        cand = sis.solve_batch(inst, sis.default_params(**{**kw, "sweeps": sw, "polish_iters": 0,
                                                           "sweeps_per_exchange": spe, "cell": cell}))
        d = np.array([(1 - min(a.L_cert, L) / L) - (1 - min(b.L_cert, L) / L)
                      for a, b, (_, L, _) in zip(own, cand, inst)])
        say("autotune probe", d)
        assert d.mean() == pytest.approx(t["quality_delta"], abs=1e-12)
        assert d.std(ddof=1) / math.sqrt(len(d)) == pytest.approx(t["quality_se"], rel=1e-9, abs=1e-15)
        # taken exactly when the upper bound of the loss is within the margin
        assert (t["quality_delta"] + 2 * t["quality_se"] <= 1e-3) == (not t["trajectory_kept"])
    else:
        assert t["trajectory_kept"] and t["quality_delta"] == 0.0
    tuned = sis.default_params(**kw, replicas_per_group=t["replicas_per_group"],
                               sweeps_per_exchange=t["sweeps_per_exchange"], cell=t["cell"])
    res = sis.solve_batch(inst, tuned)
    for (n, _, _), r in zip(inst, res):
        assert r.proposals == n * 16 * 100


def test_equal_temperatures_accept_every_exchange(sis):
    """Forced exchange (cuda/tests/globaltest.cu:39-46): with tmin == tmax the
    criterion is log u < 0, so every attempted swap is taken."""
    say("forced exchange")
    r = sis.solve(20, 6.0, None, replicas=16, tmin=1e-2, tmax=1e-2, sweeps=200, sweeps_per_exchange=5,
                  init_mode=1, shrink=0.0)
    assert r.exchanges_tried == 20 * 8 + 20 * 7
    assert r.exchanges_accepted == r.exchanges_tried


def test_packing_cli_end_to_end(tmp_path, ref_certifier):
    """`packing seed.txt epochs steps chains rng_seed out.json` keeps the raw
    JSON contract (cuda/packing.cu:859-865) and its output certifies through
    the export gate of cuda/demo.py:25-40."""
    import json
    import subprocess

    from conftest import PKG
    from refcert import certify_q
    n = 11
    w = math.isqrt(n - 1) + 1
    f = 1.00002
    rows = [((i % w + 0.5) * f, (i // w + 0.5) * f, 0.0) for i in range(n)]
    seed = tmp_path / "seed11.txt"
    seed.write_text(f"{n} {w * f:.12g}\n" + "".join(f"{x:.12g} {y:.12g} {a:.12g}\n" for x, y, a in rows))
    out = tmp_path / "raw11.json"
    say("packing cli")
    r = subprocess.run([str(PKG / "bin" / "packing"), str(seed), "20", "32", "64", "42", str(out)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    raw = json.loads(out.read_text())
    for k in ("n", "target_L", "seconds", "chain_steps", "proposals", "feasible_checkpoints", "shrinks", "chains",
              "sm_count", "blocks_per_sm", "wave_chains", "device", "q"):
        assert k in raw, k
    assert raw["n"] == n and len(raw["q"]) == n and raw["chains"] >= 64 and raw["chains"] % 8 == 0
    assert raw["target_L"] <= w * f + 1e-12
    dil, L50 = certify_q(ref_certifier, np.asarray(raw["q"]))
    assert dil is not None


def test_seeded_cli_end_to_end(tmp_path, ref_certifier):
    """`seeded seed.txt epochs steps chains seed out.json` keeps the JSON
    fields of cuda/seeded.cu:950-969 and its output certifies; with
    SEED_STRUCTURED=0 and SEED_COLLECTIVE=0 it still runs and certifies."""
    import json
    import os
    import subprocess

    from conftest import PKG
    from refcert import certify_q
    n = 10
    f = 1.00002
    rows = [((i % 4 + 0.5) * f, (i // 4 + 0.5) * f, 0.0) for i in range(n)]
    seed = tmp_path / "seed10.txt"
    seed.write_text(f"{n} {4 * f:.12g}\n" + "".join(f"{x:.12g} {y:.12g} {a:.12g}\n" for x, y, a in rows))
    for env in ({}, {"SEED_STRUCTURED": "0", "SEED_COLLECTIVE": "0", "SEED_POLISH": "0"}):
        out = tmp_path / "raw10.json"
        say("seeded cli", env)
        e = {k: v for k, v in os.environ.items() if not k.startswith("SEED_")}
        e.update(env)
        r = subprocess.run([str(PKG / "bin" / "seeded"), str(seed), "12", "32", "64", "7", str(out)],
                           capture_output=True, text=True, timeout=300, env=e)
        assert r.returncode == 0, r.stderr
        raw = json.loads(out.read_text())
        for k in ("n", "target_L", "seconds", "shrinks", "epochs", "steps", "chains", "device", "sm_count",
                  "blocks_per_sm", "wave_chains", "collective", "polish", "structured", "proposals", "accepted",
                  "exchange_attempts", "exchange_accepts", "q"):
            assert k in raw, k
        assert raw["collective"] == int(env.get("SEED_COLLECTIVE", 1))
        assert raw["structured"] == int(env.get("SEED_STRUCTURED", 1))
        assert raw["target_L"] <= 4 * f + 1e-12 and len(raw["q"]) == n
        dil, L50 = certify_q(ref_certifier, np.asarray(raw["q"]))
        assert dil is not None
