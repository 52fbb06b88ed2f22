"""The oracle's exchange step (o_ladder_model) pinned by what the mathematics
fixes: the fixed FP64 2^x and the reheating factor against pow, the exchange
criterion of cuda/packing.cu:141-152, 750-765 by hand cases a sign error fails
and by the stationary distribution of the slot permutation (detailed balance),
the step adaptation by its closed form, the improvement path of
cuda/packing.cu:503-504, 831-851 by its bookkeeping, the zero-overlap shrink by
configurations with and without overlap.  The device's exchange kernel equals
this model bit for bit (tests/test_gpu_ladder.py)."""
from __future__ import annotations

import itertools
import math

import numpy as np
import pytest

from oracle_lib import REP_DTYPE, SLOT_DTYPE

LOG2E = 1.4426950408889634
MULTS = [0.02, 0.1, 0.4, 1.0, 2.5, 5.0, 10.0, 20.0]


def make_ladder(n, R, lad, T, *, improve=0, cool=0, spe=10, L=6.0, seed=1, step=0.1):
    """Records of one instance as the engine initialises them."""
    desc = dict(n=n, R=R, ladder=lad, improve=improve, cool_period=cool, spe=spe, sweeps_total=10 ** 9,
                sweeps_done=0, key0=seed, key1=0, interval=0, shrink=1e-3, target_acc=0.3, scale=1.0,
                Lf=np.float32(L), L=L, L_arch=L, L_box_arch=L, arch_valid=0, shrinks=0, feasible=0,
                fp64_rejects=0, ex_tried=0, ex_acc=0, certs=0)
    reps = np.zeros(R, REP_DTYPE)
    reps["slot"] = np.arange(R)
    reps["Lc"] = 1e9
    reps["lam"] = 1.0
    slots = np.zeros(R, SLOT_DTYPE)
    slots["T"] = np.asarray(T, np.float32)
    slots["base_xy"] = step
    slots["base_th"] = step
    slots["step_xy"] = step
    slots["step_th"] = step
    slots["L"] = L
    slots["inv_t_log2e"] = (LOG2E / slots["T"].astype(np.float64)).astype(np.float32)
    ros = np.arange(R, dtype=np.int32)
    xycs = np.zeros((R, n, 4), np.float32)
    arch = np.zeros((n, 4), np.float32)
    return desc, reps, slots, ros, xycs, arch


def test_exp2_fixed64_and_cooling_factor(oracle):
    """The fixed 2^x is within 2 ulp of 2^x on [-60, 60]; the reheating factor is
    0.002^((k mod P)/(P-1)) (cuda/packing.cu:516) and its square root."""
    xs = np.concatenate([np.linspace(-60, 60, 4001), np.random.default_rng(0).uniform(-10, 10, 4000)])
    for x in xs:
        got = oracle.lib.o_exp2_fixed64(float(x))
        assert abs(got - 2.0 ** x) <= 4.5e-16 * 2.0 ** x, x
    for P in (2, 3, 200, 1000):
        for k in (0, 1, P // 2, P - 1, P, P + 7, 3 * P - 1):
            c, r = oracle.cool_factor(k, P)
            ref = 0.002 ** ((k % P) / (P - 1))
            assert abs(c - ref) <= 1e-15 * ref and abs(r - math.sqrt(ref)) <= 1e-15 * math.sqrt(ref), (P, k)
    assert oracle.cool_factor(5, 0) == (1.0, 1.0)
    assert oracle.cool_factor(0, 200)[0] == 1.0 and abs(oracle.cool_factor(199, 200)[0] - 0.002) < 1e-17


def test_exchange_colder_higher_energy_always_swaps(oracle):
    """Slot 0 colder than slot 1 and the colder replica at the higher energy: the
    criterion (1/Ta - 1/Tb)(Ea - Eb) > 0, every attempt swaps (a sign error
    would reject nearly all)."""
    for seed in range(200):
        d, reps, slots, ros, xy, ar = make_ladder(3, 2, 2, [1e-3, 1e-2], seed=seed)
        reps["E"] = [0.05, 0.01]
        d2, r2, s2, ros2, *_ = oracle.ladder(d, reps, slots, ros, xy, ar)
        assert d2["ex_tried"] == 1 and d2["ex_acc"] == 1
        assert list(ros2) == [1, 0] and list(r2["slot"]) == [1, 0]


def test_exchange_reverse_order_swaps_with_boltzmann_probability(oracle):
    """Colder replica at the lower energy: a swap costs
    D = (1/Ta - 1/Tb)(Eb - Ea) and happens with probability exp(-D)."""
    Ta, Tb, Ea, Eb = 1e-2, 2e-2, 0.010, 0.018
    p = math.exp(-(1 / Ta - 1 / Tb) * (Eb - Ea))  # 0.67
    trials, swaps = 4000, 0
    for seed in range(trials):
        d, reps, slots, ros, xy, ar = make_ladder(3, 2, 2, [Ta, Tb], seed=seed)
        reps["E"] = [Ea, Eb]
        d2, *_ = oracle.ladder(d, reps, slots, ros, xy, ar)
        swaps += d2["ex_acc"]
    sd = math.sqrt(p * (1 - p) / trials)
    assert abs(swaps / trials - p) < 5 * sd, (swaps / trials, p)


def test_exchange_pairs_only_adjacent_slots_of_one_ladder(oracle):
    """Two ladders of four: even intervals pair slots (0,1), (2,3), (4,5), (6,7);
    odd intervals (1,2), (5,6) and never (3,4), which lie in different ladders."""
    T = [1e-3, 2e-3, 4e-3, 8e-3] * 2
    d, reps, slots, ros, xy, ar = make_ladder(3, 8, 4, T)
    reps["E"] = [0.9, 0.8, 0.7, 0.6, 0.5, 0.4, 0.3, 0.2]  # each colder slot higher: every attempt swaps
    d2, r2, s2, ros2, *_ = oracle.ladder(d, reps, slots, ros, xy, ar)
    assert d2["ex_tried"] == 4 and list(ros2) == [1, 0, 3, 2, 5, 4, 7, 6]
    d, reps, slots, ros, xy, ar = make_ladder(3, 8, 4, T)
    d["interval"] = 1
    reps["E"] = [0.9, 0.8, 0.7, 0.6, 0.5, 0.4, 0.3, 0.2]
    d2, r2, s2, ros2, *_ = oracle.ladder(d, reps, slots, ros, xy, ar)
    assert d2["ex_tried"] == 2 and list(ros2) == [0, 2, 1, 3, 4, 6, 5, 7]
    assert all(ros2[r2["slot"][r]] == r for r in range(8))


def test_exchange_detailed_balance_over_slot_permutations(oracle):
    """Four replicas of fixed energies on a four-temperature ladder, exchange
    steps only: the slot permutation's stationary distribution is
    exp(-sum_s E(slot s) / T_s) / Z (parallel tempering's product measure).
    Histogram of 60 000 steps against the 24 exact probabilities."""
    T = np.array([0.01, 0.02, 0.04, 0.08])
    E = np.array([0.00, 0.01, 0.03, 0.06])  # replica energies
    d, reps, slots, ros, xy, ar = make_ladder(2, 4, 4, T, seed=9)
    reps["E"] = E + 1e-7  # nonzero: no zero-overlap test
    E = reps["E"].astype(np.float64)
    perms = list(itertools.permutations(range(4)))
    w = np.array([math.exp(-sum(E[p[s]] / T[s] for s in range(4))) for p in perms])
    p_exact = w / w.sum()
    counts = dict.fromkeys(perms, 0)
    steps = 60000
    for _ in range(steps):
        d, reps, slots, ros, xy, ar, _imp = oracle.ladder(d, reps, slots, ros, xy, ar)
        slots["T"] = T  # keep the ladder (no cooling: unchanged anyway)
        counts[tuple(int(v) for v in ros)] += 1
    freq = np.array([counts[p] for p in perms]) / steps
    # exchange chains decorrelate in a few steps; allow an integrated
    # autocorrelation time of 8 steps
    sd = np.sqrt(p_exact * (1 - p_exact) * 8 / steps)
    assert np.all(np.abs(freq - p_exact) < 5 * sd + 1e-3), (freq, p_exact)
    # the reversed criterion would favour the hot-low assignment instead
    worst = perms[int(np.argmin(p_exact))]
    assert freq[perms.index(worst)] < 0.02


def test_adaptation_closed_form_and_clamps(oracle):
    """Each slot's step bases scale by e^((rate - target)/2) of the replica that
    held it (cuda/packing.cu:618), clamped to [sqrt(3) 1e-7, sqrt(3)/2] (xy) and
    [sqrt(3) 1e-7, 0.7794229] (angle); the next interval's steps are the bases."""
    n, R = 10, 6
    d, reps, slots, ros, xy, ar = make_ladder(n, R, R, np.geomspace(1e-3, 1e-1, R), spe=10)
    reps["E"] = np.linspace(1, 2, R)
    reps["acc"] = [0, 30, 60, 100, 0, 100]
    slots["base_xy"] = [0.1, 0.1, 0.1, 0.1, 2e-7, 0.86]
    slots["base_th"] = [0.1, 0.1, 0.1, 0.1, 2e-7, 0.77]
    # replica r holds slot (r + 1) % R
    reps["slot"] = (np.arange(R) + 1) % R
    ros = np.argsort(reps["slot"]).astype(np.int32)
    d2, r2, s2, *_ = oracle.ladder(d, reps, slots, ros, xy, ar)
    for r in range(R):
        s = int(reps["slot"][r])
        rate = reps["acc"][r] / (n * 10)
        f = math.exp((rate - 0.3) / 2)
        exy = min(max(slots["base_xy"][s] * f, 1.7320508e-7), 0.8660254)
        eth = min(max(slots["base_th"][s] * f, 1.7320508e-7), 0.7794229)
        assert abs(s2["base_xy"][s] - exy) <= 3e-7 * exy and abs(s2["base_th"][s] - eth) <= 3e-7 * eth
        assert s2["step_xy"][s] == s2["base_xy"][s] and s2["step_th"][s] == s2["base_th"][s]


def test_cooling_sets_next_interval_temperature_and_steps(oracle):
    """With a period P the slots leave step k with 1/T of T * cool(k + 1) and
    steps base * sqrt(cool(k + 1)) (cuda/packing.cu:516-517, 523)."""
    R, P = 4, 5
    T = np.array([1e-3, 1e-2, 1e-1, 1.0])
    d, reps, slots, ros, xy, ar = make_ladder(3, R, R, T, cool=P)
    reps["E"] = 1.0
    reps["acc"] = 9  # rate 0.3: bases unchanged
    for k in range(2 * P + 1):
        d, reps, slots, ros, xy, ar, _ = oracle.ladder(d, reps, slots, ros, xy, ar)
        c = 0.002 ** (((k + 1) % P) / (P - 1))
        got_t = LOG2E / slots["inv_t_log2e"].astype(np.float64)
        assert np.allclose(got_t, slots["T"].astype(np.float64) * c, rtol=2e-7)
        assert np.allclose(slots["step_xy"], slots["base_xy"] * math.sqrt(c), rtol=2e-7)
    assert d["interval"] == 2 * P + 1


def test_improvement_path_bookkeeping(oracle):
    """cuda/packing.cu:831-851 restated on temperature slots: the smallest
    certified side (lowest replica on ties) below best - 1e-11 becomes the best
    side and the archive; every slot's box is max(sqrt n + 1e-6, best - shrink *
    mults[8 pos / ladder]); the coldest slot of each ladder restarts from the
    archive dilated by lambda, translated by the lower corner and scaled to its
    box; no exchange that interval."""
    n, R, lad = 4, 16, 8
    rng = np.random.default_rng(3)
    d, reps, slots, ros, xy, ar = make_ladder(n, R, lad, np.tile(np.geomspace(1e-3, 1e-1, lad), 2),
                                              improve=1, L=4.0)
    xy[:] = rng.uniform(0.6, 3.4, xy.shape).astype(np.float32)
    xy[:, :, 2], xy[:, :, 3] = np.float32(1.0), np.float32(0.0)
    reps["E"] = rng.uniform(0.01, 0.1, R)
    # permute the slots so that slot owners differ from replica indices
    perm = rng.permutation(R)
    reps["slot"] = perm
    ros = np.argsort(perm).astype(np.int32)
    reps["Lc"] = 1e9
    reps["Lc"][[5, 9]] = 3.9998  # tie: replica 5 wins
    reps["lam"][[5, 9]] = 1.0001
    reps["minx"][5], reps["miny"][5] = 0.05, 0.07
    d2, r2, s2, ros2, xy2, ar2, imp = oracle.ladder(d, reps, slots, ros, xy, ar)
    assert imp and d2["L"] == 3.9998 and d2["L_arch"] == 3.9998 and d2["arch_valid"] == 1
    assert d2["shrinks"] == 1 and d2["feasible"] == 1 and d2["ex_tried"] == 0
    assert d2["L_box_arch"] == float(slots["L"][perm[5]])
    assert np.array_equal(ar2, xy[5])
    boxes = [max(2 + 1e-6, 3.9998 - 1e-3 * MULTS[(s % lad) * 8 // lad]) for s in range(R)]
    assert np.array_equal(s2["L"], np.asarray(boxes, np.float32))
    heads = {int(ros[0]), int(ros[lad])}
    for r in range(R):
        if r in heads:
            s = int(perm[r])
            sc = boxes[s] / 3.9998
            ex = ((xy[5, :, 0].astype(np.float64) * 1.0001 - 0.05) * sc).astype(np.float32)
            assert np.array_equal(xy2[r, :, 0], ex), r
            assert np.array_equal(xy2[r, :, 2:], xy[5, :, 2:])
        else:
            assert np.array_equal(xy2[r], xy[r]), r
    # gain 2e-4 < 5e-4: only heads.  A gain above 5e-4 restarts the whole first ladder.
    reps["Lc"][[5, 9]] = 3.99
    d3, r3, s3, ros3, xy3, *_ = oracle.ladder(d, reps, slots, ros, xy, ar)
    changed = {r for r in range(R) if not np.array_equal(xy3[r], xy[r])}
    assert changed == {int(ros[s]) for s in range(lad)} | {int(ros[lad])}
    # no improvement (within 1e-11 of the best): exchange as usual
    reps["Lc"][[5, 9]] = 4.0 - 5e-12
    d4, *_rest, imp4 = oracle.ladder(d, reps, slots, ros, xy, ar)
    assert not imp4 and d4["L"] == 4.0 and d4["ex_tried"] > 0


def test_zero_overlap_shrink_and_fp64_reject(oracle):
    """improve = 0: the coldest replica with FP32 energy exactly zero is tested in
    FP64; a configuration without overlap is archived and the box shrinks by
    `shrink` (centres rescaled by the ratio at the next staging); one whose FP64
    energy is positive is counted as a rejection and nothing moves."""
    n, R = 4, 4
    d, reps, slots, ros, xy, ar = make_ladder(n, R, R, [1e-3, 2e-3, 4e-3, 8e-3], L=3.0)
    grid = np.array([[0.75, 0.75], [2.25, 0.75], [0.75, 2.25], [2.25, 2.25]], np.float32)
    xy[:, :, :2] = grid
    xy[:, :, 2], xy[:, :, 3] = np.float32(1.0), np.float32(0.0)
    reps["E"] = [0.1, 0.0, 0.0, 0.2]
    d2, r2, s2, ros2, xy2, ar2, _ = oracle.ladder(d, reps, slots, ros, xy, ar)
    assert d2["arch_valid"] == 1 and d2["shrinks"] == 1 and d2["L"] == pytest.approx(2.999)
    assert d2["scale"] == np.float32(2.999 / 3.0) and np.all(s2["L"] == np.float32(2.999))
    assert d2["fp64_rejects"] == 0 and np.array_equal(ar2, xy[int(ros2[1])])
    xy[:, 1, :2] = [1.5, 0.75]  # squares 0 and 1 overlap in FP64 (the FP32 energy claims 0)
    d3, r3, s3, *_ = oracle.ladder(d, reps, slots, ros, xy, ar)
    assert d3["fp64_rejects"] == 1 and d3["shrinks"] == 0 and d3["arch_valid"] == 0 and d3["L"] == 3.0
