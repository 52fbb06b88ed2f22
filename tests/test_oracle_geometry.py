"""The CPU oracle pinned against the reference's own device-test corpora,
tolerances and special cases (no GPU).

cuda/tests/selftest.cu:33-53   FP32 SAT vs FP64 oracle (2e-6), cell list vs all pairs (1e-4)
cuda/tests/seeded_test.cu:15-30 FP64 pair vs oracle (1e-12), gradient vs central FD (1e-7)
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle_lib import f32, f64, ptr


def _fa(x):
    return np.ascontiguousarray(x, dtype=np.float32)


def test_selftest_edge_fixtures(oracle):
    """selftest.cu:37-38: touching, just separated, just overlapping at a cell edge."""
    L = oracle.lib
    a = _fa([2, 2, 0])
    touch, sep, over = _fa([3, 2, 0]), _fa([3.000001, 2, 0]), _fa([2.999999, 2, 0])
    assert L.o_ref_depth32(ptr(a, f32), ptr(touch, f32)) == 0.0
    assert L.o_ref_depth32(ptr(a, f32), ptr(sep, f32)) == 0.0
    d = L.o_ref_depth32(ptr(a, f32), ptr(over, f32))
    assert 0.5e-6 < d < 2e-6
    # the product's closed form on the same fixtures, cos/sin of 0 exact
    assert L.o_depth32_cs(1.0, 0.0, 1.0, 0.0, 1.0, 0.0) == 0.0
    assert L.o_depth32_cs(np.float32(3.000001) - 2, 0.0, 1.0, 0.0, 1.0, 0.0) == 0.0
    assert L.o_depth32_cs(np.float32(2.999999) - 2, 0.0, 1.0, 0.0, 1.0, 0.0) > 0.0


def test_selftest_corpus_depth_vs_fp64(oracle, corpus):
    """Both FP32 formulations within 2e-6 of the FP64 axis oracle on 250 000 pairs."""
    q = corpus["selftest"]
    L = oracle.lib
    worst_ref = worst_cs = 0.0
    disagree = 0
    for b in range(0, 100, 5):  # 20 blocks x 2500 ordered pairs = 50 000 pairs
        blk = q[b]
        cs = np.cos(blk[:, 2].astype(np.float32)).astype(np.float32)
        sn = np.sin(blk[:, 2].astype(np.float32)).astype(np.float32)
        for i in range(50):
            pi = _fa(blk[i])
            for j in range(50):
                pj = _fa(blk[j])
                ref64 = L.o_sat_oracle_depth_f(ptr(pi, f32), ptr(pj, f32))
                d_ref = L.o_ref_depth32(ptr(pi, f32), ptr(pj, f32))
                dx = np.float32(blk[j, 0] - blk[i, 0])
                dy = np.float32(blk[j, 1] - blk[i, 1])
                d_cs = L.o_depth32_cs(dx, dy, cs[i], sn[i], cs[j], sn[j]) if i != j else d_ref
                worst_ref = max(worst_ref, abs(d_ref - ref64))
                worst_cs = max(worst_cs, abs(d_cs - ref64))
                if abs(ref64) > 1e-5 and ((d_cs > 0) != (ref64 > 0)):
                    disagree += 1
    assert worst_ref <= 2e-6, worst_ref
    assert worst_cs <= 2e-6, worst_cs
    assert disagree == 0


def test_closed_form_radius_equals_per_axis_radius(oracle):
    """R = 1/2 (1 + |cos d| + |sin d|) is the reference's per-axis projected
    radius (cuda/packing.cu:29-30) on all four axes."""
    rng = np.random.default_rng(5)
    for _ in range(2000):
        a, b = rng.uniform(-4, 4, 2)
        c, s, d, t = math.cos(a), math.sin(a), math.cos(b), math.sin(b)
        axes = [(c, s), (-s, c), (d, t), (-t, d)]
        radii = [0.5 * (abs(ux * c + uy * s) + abs(-ux * s + uy * c) + abs(ux * d + uy * t) + abs(-ux * t + uy * d))
                 for ux, uy in axes]
        R = 0.5 + 0.5 * (abs(c * d + s * t) + abs(s * d - c * t))
        assert max(abs(r - R) for r in radii) < 1e-14


def _corner_overlap(p, q):
    """Independent construction: project the 8 corners on each of the 4 edge
    normals; the penetration along an axis is the shortest 1-D push that
    separates the two projected intervals; depth = min over the axes."""
    def corners(x, y, a):
        c, s = math.cos(a), math.sin(a)
        return [(x + 0.5 * (sx * c - sy * s), y + 0.5 * (sx * s + sy * c)) for sx, sy in
                ((1, 1), (1, -1), (-1, 1), (-1, -1))]
    cp, cq = corners(*p), corners(*q)
    best = 1e9
    for a in (p[2], q[2]):
        for ux, uy in ((math.cos(a), math.sin(a)), (-math.sin(a), math.cos(a))):
            P = [x * ux + y * uy for x, y in cp]
            Q = [x * ux + y * uy for x, y in cq]
            best = min(best, min(max(P) - min(Q), max(Q) - min(P)))
    return max(0.0, best)


def test_sat_matches_corner_projection_construction(oracle):
    rng = np.random.default_rng(11)
    for _ in range(3000):
        p = np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-3, 3)])
        q = np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-3, 3)])
        e64 = oracle.lib.o_sat_oracle64(ptr(p, f64), ptr(q, f64))
        assert abs(math.sqrt(e64) - _corner_overlap(p, q)) < 1e-12
        e, _ = oracle.pair64(p, q)
        assert abs(e - e64) <= 1e-12


def test_axis_aligned_reduces_to_interval_overlap(oracle):
    rng = np.random.default_rng(3)
    for _ in range(2000):
        dx, dy = rng.uniform(-1.3, 1.3, 2).astype(np.float32)
        k1, k2 = rng.integers(0, 4, 2)
        c1, s1 = [(1, 0), (0, 1), (-1, 0), (0, -1)][k1]
        c2, s2 = [(1, 0), (0, 1), (-1, 0), (0, -1)][k2]
        d = oracle.lib.o_depth32_cs(dx, dy, c1, s1, c2, s2)
        expect = max(0.0, 1.0 - max(abs(float(dx)), abs(float(dy))))
        if dx * dx + dy * dy >= 2:
            expect = 0.0
        assert abs(d - expect) < 1e-6


def test_far_pairs_have_zero_depth(oracle):
    rng = np.random.default_rng(9)
    for _ in range(2000):
        r = rng.uniform(math.sqrt(2) + 1e-4, 3)
        phi = rng.uniform(0, 2 * math.pi)
        a, b = rng.uniform(0, 2 * math.pi, 2)
        d = oracle.lib.o_depth32_cs(r * math.cos(phi), r * math.sin(phi), math.cos(a), math.sin(a),
                                    math.cos(b), math.sin(b))
        assert d == 0.0


def test_cell_list_equals_all_pairs(oracle, corpus):
    """selftest.cu:19-30: 3x3-stencil local energy vs all pairs, tolerance 1e-4."""
    q = corpus["selftest"]
    worst = 0.0
    for b in range(100):
        blk = np.ascontiguousarray(q[b].reshape(-1), dtype=np.float32)
        loc = np.zeros(50, np.float32)
        full = np.zeros(50, np.float32)
        oracle.lib.o_ref_local32(50, ptr(blk, f32), 12.0, ptr(loc, f32), ptr(full, f32))
        worst = max(worst, float(np.max(np.abs(loc - full))))
    assert worst <= 1e-4, worst


def test_reference_cell_binning_closed_form(oracle):
    """cuda/packing.cu:22-25: clamp(floor(x/2), 0, 15) per axis, row-major."""
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-5, 40, 5000), np.arange(-2, 34, 0.5)]).astype(np.float32)
    for x in xs:
        y = np.float32(x * 0.37 + 1)
        c = oracle.lib.o_ref_cell(x, y)
        cx = min(15, max(0, math.floor(float(x) / 2)))
        cy = min(15, max(0, math.floor(float(y) / 2)))
        assert c == cy * 16 + cx
        assert oracle.lib.o_cellcoord(x, 0.0, 2.0, 0, 15) == cx


def test_seeded_corpus_pair_and_gradient(oracle, corpus):
    """seeded_test.cu:15-30: pair vs oracle 1e-12, gradient vs FD (h=1e-7) 1e-7."""
    qs = corpus["seeded"]
    err = jerr = 0.0
    for i in range(0, 10000, 4):
        p, z = qs[i, 0].copy(), qs[i, 1].copy()
        e, g = oracle.pair64(p, z)
        err = max(err, abs(e - oracle.lib.o_sat_oracle64(ptr(p, f64), ptr(z, f64))))
        for k in range(3):
            a, b = p.copy(), p.copy()
            a[k] += 1e-7
            b[k] -= 1e-7
            fd = (oracle.pair64(a, z)[0] - oracle.pair64(b, z)[0]) / 2e-7
            jerr = max(jerr, abs(fd - g[k]))
    assert err <= 1e-12, err
    assert jerr <= 1e-7, jerr


def test_boundary_gradient_finite_difference(oracle):
    rng = np.random.default_rng(4)
    L = 3.0
    for _ in range(500):
        p = np.array([rng.uniform(-0.2, 3.2), rng.uniform(-0.2, 3.2), rng.uniform(0.1, 1.4)])
        g = np.zeros(3)
        oracle.lib.o_boundary64(ptr(p, f64), L, ptr(g, f64))
        for k in range(3):
            a, b = p.copy(), p.copy()
            a[k] += 1e-7
            b[k] -= 1e-7
            fd = (oracle.lib.o_boundary64(ptr(a, f64), L, None) - oracle.lib.o_boundary64(ptr(b, f64), L, None)) / 2e-7
            assert abs(fd - g[k]) < 1e-6


def test_product_wall_equals_reference_wall(oracle):
    """FP32 wall (cuda/packing.cu:35-39) == product wall with the same cos/sin."""
    rng = np.random.default_rng(8)
    for _ in range(3000):
        x, y = rng.uniform(-1, 7, 2).astype(np.float32)
        a = np.float32(rng.uniform(-3, 3))
        p = _fa([x, y, a])
        ref = oracle.lib.o_ref_wall32(ptr(p, f32), 6.0)
        c, s = np.float32(math.cos(a)), np.float32(math.sin(a))
        mine = oracle.lib.o_wall32_cs(x, y, np.cos(a).astype(np.float32), np.sin(a).astype(np.float32), 6.0)
        assert abs(ref - mine) <= 1e-5 * max(1.0, ref)
        del c, s


def test_rotation_polynomial(oracle):
    rng = np.random.default_rng(12)
    import ctypes
    c_out, s_out = ctypes.c_float(), ctypes.c_float()
    for _ in range(2000):
        a = rng.uniform(-4, 4)
        t = rng.uniform(-0.78, 0.78)
        oracle.lib.o_rot32(math.cos(a), math.sin(a), t, ctypes.byref(c_out), ctypes.byref(s_out))
        assert abs(c_out.value - math.cos(a + t)) < 4e-7
        assert abs(s_out.value - math.sin(a + t)) < 4e-7


def test_certified_side_separates_every_pair(oracle):
    """cuda/packing.cu:254-297: after the lambda dilation every near pair is
    separated in FP64, lambda >= 1, and L >= sqrt(n)."""
    rng = np.random.default_rng(6)
    grid = np.zeros((9, 3))
    oracle.lib.o_grid_packing(9, ptr(grid, f64))
    for trial in range(20):
        q = grid.copy()
        q[:, :2] += rng.uniform(-1e-4, 1e-4, (9, 2))
        q[:, 2] += rng.uniform(-1e-5, 1e-5, 9)
        side, lam, out = oracle.certified_side(q)
        assert lam >= 1.0 and side < 1e8
        assert side >= 3.0 - 1e-7
        for i in range(9):
            for j in range(i + 1, 9):
                assert oracle.lib.o_sat_oracle64(ptr(out[i].copy(), f64), ptr(out[j].copy(), f64)) == 0.0


def test_certified_side_rejects_deep_overlap(oracle):
    q = np.array([[0.5, 0.5, 0.0], [1.2, 0.5, 0.0]])
    side, lam, _ = oracle.certified_side(q)
    assert side == 1e9 and lam > 1.008


def test_polish_never_increases_energy(oracle):
    rng = np.random.default_rng(2)
    grid = np.zeros((16, 3))
    oracle.lib.o_grid_packing(16, ptr(grid, f64))
    q = grid.copy()
    q[:, :2] *= 0.985
    q[:, :2] += rng.uniform(-0.02, 0.02, (16, 2))
    q[:, 2] = rng.uniform(0, 0.05, 16)
    e0 = oracle.total64(q, 3.98)
    q2, e1, it = oracle.polish64(q, 3.98, 200)
    assert e1 <= e0
    assert abs(oracle.total64(q2, 3.98) - e1) <= 1e-15 * max(1, e1)


def test_grid_packing_matches_reference_grid(oracle):
    """src/squares/core.py:27-35: width = isqrt(n-1)+1, centres at half integers."""
    for n in (1, 2, 3, 4, 5, 9, 10, 16, 31):
        q = np.zeros((n, 3))
        oracle.lib.o_grid_packing(n, ptr(q, f64))
        width = math.isqrt(n - 1) + 1
        for i in range(n):
            assert q[i, 0] == i % width + 0.5 and q[i, 1] == i // width + 0.5 and q[i, 2] == 0.0


def _accept_exp2_inputs(rng):
    """x <= 0 over the whole useful range, with the round-to-nearest ties of
    the range reduction (half integers), integers, -0, the cut-off at -100 and
    values just inside and outside it."""
    x = np.concatenate([-rng.exponential(3.0, 200000), -rng.uniform(0, 110, 50000),
                        -np.arange(0, 101, 0.5), -np.arange(0, 1, 1 / 64),
                        [-0.0, -1e-30, -1e-7, -99.99999, -100.0, -100.00001, -120.0]])
    return x.astype(np.float32)


def test_accept_exp2_against_two_to_the_x(oracle):
    """The Metropolis test's 2^x (o_accept_exp2, the device's accept_exp2):
    relative error below 1.5e-7 of 2^x in FP64 over [-100, 0]; exact powers of
    two at the integers (the polynomial is 1 at f = 0 and the scaling is
    exact); 0 below -100, where every accept uniform (>= 2^-24) rejects."""
    f = oracle.lib.o_accept_exp2
    rng = np.random.default_rng(5)
    x = _accept_exp2_inputs(rng)
    worst = 0.0
    for v in x:
        got = f(float(v))
        if v < -100.0:
            assert got == 0.0
            continue
        ref = 2.0 ** float(v)
        worst = max(worst, abs(got - ref) / ref)
    assert worst < 1.5e-7, worst
    for k in range(0, 101):
        assert f(-float(k)) == 2.0 ** -k
    assert f(float("nan")) == 0.0
    assert f(-25.0) < 2.0 ** -24  # from here down no accept uniform passes

