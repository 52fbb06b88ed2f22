"""GPU parity of the exchange step and of whole PT-SA trajectories.

- k_ladder (sis_dev_ladder) equals the oracle's o_ladder_model bit for bit:
  step adaptation, exchange decisions (cuda/packing.cu:141-152, 750-765), the
  FP64 zero-overlap shrink, the improvement path (:503-504, 831-851) and the
  cooling cycle (:516-517, 756-758), over 60 chained intervals of unequal
  ladders.
- The sweep launch's FP64 certified side of every replica equals the oracle's
  all-pairs restatement of cuda/packing.cu:254-297.
- A resident batch run interval by interval (sweeps, certified sides, exchange)
  equals the oracle's sweep model + ladder model at every interval."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle_lib import REP_DTYPE, SLOT_DTYPE

pytestmark = pytest.mark.gpu

LOG2E = 1.4426950408889634


def say(*a):
    print("[gpu-case]", *a, flush=True)


def _grid_state(rng, R, n, L, pitch):
    """Replicas on a grid (zero overlap when pitch >= 1 and the grid fits) with
    a little jitter in some replicas."""
    k = int(math.isqrt(n - 1)) + 1
    i = np.arange(n)
    st = np.zeros((R, n, 4), np.float32)
    st[:, :, 0] = (i % k + 0.5) * pitch
    st[:, :, 1] = (i // k + 0.5) * pitch
    a = np.zeros((R, n))
    a[R // 2:] = rng.uniform(0, 0.3, (R - R // 2, n))  # the upper half overlap
    st[:, :, 2] = np.cos(a)
    st[:, :, 3] = np.sin(a)
    assert st[:, :, :2].max() <= L
    return st


def _ladder_case(n, R, lad, improve, cool, L, seed):
    rng = np.random.default_rng(seed)
    desc = dict(n=n, R=R, ladder=lad, improve=improve, cool_period=cool, spe=10, sweeps_total=10 ** 6,
                sweeps_done=0, key0=seed, key1=7, interval=0, shrink=np.float32(2e-3), target_acc=np.float32(0.3),
                scale=1.0, Lf=np.float32(L), L=L, L_arch=L, L_box_arch=L, arch_valid=0, shrinks=0, feasible=0,
                fp64_rejects=0, ex_tried=0, ex_acc=0, certs=0)
    reps = np.zeros(R, REP_DTYPE)
    perm = rng.permutation(R)
    reps["slot"] = perm
    reps["Lc"] = 1e9
    reps["lam"] = 1.0
    ros = np.argsort(perm).astype(np.int32)
    slots = np.zeros(R, SLOT_DTYPE)
    pos = np.arange(R) % lad
    slots["T"] = (1e-4 * (1e3 ** (pos / max(1, lad - 1)))).astype(np.float32)
    slots["base_xy"] = rng.uniform(0.01, 0.3, R)
    slots["base_th"] = rng.uniform(0.01, 0.3, R)
    slots["step_xy"] = slots["base_xy"]
    slots["step_th"] = slots["base_th"]
    slots["L"] = L
    slots["inv_t_log2e"] = (LOG2E / slots["T"].astype(np.float64)).astype(np.float32)
    st = _grid_state(rng, R, n, L, 1.0 + 1e-6)
    arch = st[0].copy()
    return rng, desc, reps, slots, ros, st, arch


def _randomise(rng, desc, reps, st, step):
    """New sweep results for the next step: energies (a few exactly zero),
    accept counts and, with improve, certified sides (some below the best)."""
    R, n = desc["R"], desc["n"]
    reps["E"] = rng.uniform(0, 0.05, R).astype(np.float32)
    reps["E"][rng.random(R) < 0.25] = 0.0
    reps["acc"] = rng.integers(0, n * desc["spe"] + 1, R)
    if desc["improve"]:
        lc = desc["L"] - rng.uniform(-0.02, 0.004, R)
        lc[rng.random(R) < 0.5] = 1e9
        if step % 7 == 3:  # a tie between two replicas at the smallest side
            lc[1] = lc[5] = desc["L"] - 0.005
        reps["Lc"] = lc
        reps["lam"] = 1.0 + rng.uniform(0, 1e-3, R)
        reps["minx"] = rng.uniform(-0.01, 0.01, R)
        reps["miny"] = rng.uniform(-0.01, 0.01, R)
    # moves of the upper half (they overlap); the lower half keeps its grid
    st[R // 2:, :, :2] += rng.uniform(-0.01, 0.01, st[R // 2:, :, :2].shape).astype(np.float32)
    st[:, :, :2] = np.clip(st[:, :, :2], 0, desc["Lf"])


@pytest.mark.parametrize("n,R,lad,improve,cool,L", [
    (20, 32, 8, 0, 0, 5.2),      # four ladders, zero-overlap shrink (all-pairs FP64 test)
    (20, 32, 8, 1, 7, 5.2),      # improvement path with cooling
    (120, 24, 24, 0, 200, 11.2),  # one ladder, grid FP64 test (n >= 96)
    (50, 16, 16, 1, 0, 8.2),     # one ladder, improvement path, gains above 5e-4
    (130, 32, 32, 1, 0, 12.2),   # restarts with one thread per square
    (300, 16, 16, 1, 0, 18.2),   # restarts with squares looped over the threads (n > 256)
    (12, 600, 8, 1, 0, 4.2),     # R > 2 x 256: several replicas, slots and pairs per thread
    (12, 600, 24, 0, 5, 4.2),    # the same with the zero-overlap shrink
])
def test_ladder_kernel_matches_model(sis, oracle, n, R, lad, improve, cool, L):
    rng, desc, reps, slots, ros, st, arch = _ladder_case(n, R, lad, improve, cool, L, 100 + n + improve)
    improved = zero_tests = 0
    for step in range(60):
        _randomise(rng, desc, reps, st, step)
        say("ladder", n, R, lad, improve, cool, step)
        dd, dr, ds, dros, dst, dar = sis.dev_ladder(desc, reps, slots, ros, st, arch)
        od, orr, osl, oros, ost, oar, imp = oracle.ladder(desc, reps, slots, ros, st, arch)
        for k in od:
            assert dd[k] == od[k], (step, k, dd[k], od[k])
        assert np.array_equal(dr, orr), step
        assert np.array_equal(ds.view(np.uint32), osl.view(np.uint32)), step
        assert np.array_equal(dros, oros) and np.array_equal(dst, ost) and np.array_equal(dar, oar), step
        improved += imp
        zero_tests += (od["fp64_rejects"] + od["shrinks"]) > (desc["fp64_rejects"] + desc["shrinks"])
        desc, reps, slots, ros, st, arch = od, orr, osl, oros, ost, oar
        desc["Lf"] = np.float32(desc["Lf"])
    # an interval that improves does no exchange (with 600 replicas every one does)
    assert desc["interval"] == 60 and (desc["ex_tried"] > 0 or improved == 60)
    if improve:
        assert improved >= 3
    else:
        assert zero_tests >= 3


@pytest.mark.parametrize("n,L,R,group", [(11, 4.2, 4, 1), (50, 8.3, 6, 3), (150, 13.5, 3, 1), (700, 27.5, 2, 1)])
def test_launch_certified_side_matches_oracle(sis, oracle, n, L, R, group):
    """Per-replica boxes below the grid side: the sweeps equal the model run in
    each replica's own box, and the FP64 certified side of every replica at the
    end of the launch equals the oracle's all-pairs one (cuda/packing.cu:254-297);
    its dilation and corner are defined where it certifies."""
    rng = np.random.default_rng(n)
    k = int(math.isqrt(n - 1)) + 1
    pitch = (L - 0.1) / k
    i = np.arange(n)
    st = np.zeros((R, n, 4), np.float32)
    # replica 0 exactly on the grid, the middle ones jittered without overlap,
    # the last one compressed by 3 % (dilation above 1.008: no certified side)
    jit = np.full((R, 1), 0.005)
    jit[0] = 0.0
    st[:, :, 0] = (i % k + 0.5) * pitch + rng.uniform(-1, 1, (R, n)) * jit
    st[:, :, 1] = (i // k + 0.5) * pitch + rng.uniform(-1, 1, (R, n)) * jit
    if R > 1:
        st[-1, :, :2] *= 0.97
    a = rng.uniform(-1, 1, (R, n)) * jit
    st[:, :, 2], st[:, :, 3] = np.cos(a), np.sin(a)
    box = np.float32(L) - rng.uniform(0, 0.05, R).astype(np.float32)
    box[0] = np.float32(L)
    st[:, :, :2] = np.minimum(st[:, :, :2], box[:, None, None])
    st[:, :, :2] = np.maximum(st[:, :, :2], 0)
    T = np.geomspace(1e-5, 1e-3, R)
    ik = (LOG2E / T).astype(np.float32)
    sxy = np.full(R, 0.01, np.float32)
    sth = np.full(R, 0.01, np.float32)
    say("launch certify", n, L, R, group)
    dst, dacc, den, dcnt, dreps = sis.dev_launch(st, L, box, 0.7075, ik, sxy, sth, 5, 1, 40, 6, group=group,
                                                 certify=True)
    certified = 0
    for r in range(R):
        x, y, c, s = (np.ascontiguousarray(st[r, :, j]) for j in range(4))
        acc = 0
        for sw in range(40, 46):
            a_, _ = oracle.sweep_model(x, y, c, s, float(box[r]), 0.7075, sxy[r], sth[r], ik[r], 5, 1, r, sw, 4)
            acc += a_
        assert dacc[r] == acc
        m = np.stack([x, y, c, s], -1)
        assert np.array_equal(dst[r], m), r
        assert den[r] == oracle.energy32(x, y, c, s, float(box[r]), 0.7075)
        side, lam, mx, my = oracle.certify_xycs(dst[r])
        if side < 1e8:
            certified += 1
            assert dreps["lam"][r] == pytest.approx(lam, rel=1e-14, abs=0)
            assert dreps["Lc"][r] == pytest.approx(side, rel=1e-14)
            assert dreps["minx"][r] == pytest.approx(mx, rel=1e-13, abs=1e-13)
            assert dreps["miny"][r] == pytest.approx(my, rel=1e-13, abs=1e-13)
        else:
            assert dreps["Lc"][r] == 1e9
    assert certified >= 1 and (R == 1 or dreps["Lc"][R - 1] == 1e9)


@pytest.mark.parametrize("n,L,R,group", [(50, 8.3, 6, 3), (150, 13.5, 4, 1)])
def test_launch_certifies_before_the_kick(sis, oracle, n, L, R, group):
    """A launch with collective moves, a kick and certified sides, in the
    reference's order (cuda/packing.cu:689-745: checkpoint, then kick): the
    certified sides are those of the state after the collective moves, a
    replica they certify is not kicked (its state stays the certified one),
    the others are kicked; device == oracle model bit for bit."""
    rng = np.random.default_rng(n + 1)
    k = int(math.isqrt(n - 1)) + 1
    pitch = (L - 0.1) / k
    i = np.arange(n)
    st = np.zeros((R, n, 4), np.float32)
    jit = np.full((R, 1), 0.004)
    st[:, :, 0] = (i % k + 0.5) * pitch + rng.uniform(-1, 1, (R, n)) * jit
    st[:, :, 1] = (i // k + 0.5) * pitch + rng.uniform(-1, 1, (R, n)) * jit
    st[R // 2:, :, :2] *= 0.93  # overlapping beyond a certifiable depth: kicked
    a = rng.uniform(-1, 1, (R, n)) * jit
    st[:, :, 2], st[:, :, 3] = np.cos(a), np.sin(a)
    box = np.full(R, np.float32(L))
    st[:, :, :2] = np.clip(st[:, :, :2], 0, L)
    T = np.geomspace(1e-6, 1e-5, R)
    ik = (LOG2E / T).astype(np.float32)
    sxy = np.full(R, 0.004, np.float32)
    sth = np.full(R, 0.004, np.float32)
    say("certify before kick", n, R, group)
    dst, dacc, den, dcnt, dreps = sis.dev_launch(st, L, box, 0.7075, ik, sxy, sth, 9, 2, 40, 6, group=group,
                                                 collective=2, kick_every=1, interval=0, certify=True)
    kicked = certified = 0
    for r in range(R):
        x, y, c, s = (np.ascontiguousarray(st[r, :, j]) for j in range(4))
        for sw in range(40, 46):
            oracle.sweep_model(x, y, c, s, L, 0.7075, sxy[r], sth[r], ik[r], 9, 2, r, sw, 4)
        _, e = oracle.collective_model(x, y, c, s, L, 0.7075, sxy[r], sth[r], ik[r], 9, 2, r, 46, 2)
        side, lam, mx, my = oracle.certify_xycs(np.stack([x, y, c, s], -1))
        before = np.stack([x, y, c, s], -1).copy()
        if side < 1e8:
            certified += 1
            assert dreps["Lc"][r] == pytest.approx(side, rel=1e-14)
            assert np.array_equal(dst[r], before), r  # certified: not kicked
        else:
            assert dreps["Lc"][r] == 1e9
            kk, e = oracle.kick_model(x, y, c, s, L, 0.7075, 9, 2, r, 0, True)
            kicked += kk
        assert np.array_equal(dst[r], np.stack([x, y, c, s], -1)), r
        assert den[r] == e, r
    assert dcnt[5] == kicked and certified >= 1 and kicked >= 1


def _model_interval(oracle, s, w, spsh, certify_every):
    """One interval of the oracle: staging (scale, clamp to the slot box),
    sweeps, canonical energy, certified side (improve = 1, every
    certify_every-th launch and the last; absent, 1e9, otherwise), exchange
    step."""
    d = s["desc"]
    R, n = d["R"], d["n"]
    left = d["sweeps_total"] - d["sweeps_done"]
    nsw = min(d["spe"], left)
    sweep0 = d["sweeps_done"]
    certify = d["improve"] and ((d["interval"] + 1) % certify_every == 0 or nsw >= left)
    st, reps, slots = s["state"].copy(), s["reps"].copy(), s["slots"]
    scale = np.float32(d["scale"])
    for r in range(R):
        sl = slots[reps["slot"][r]]
        Lr = sl["L"]
        x = np.ascontiguousarray(np.minimum(np.maximum(st[r, :, 0] * scale, np.float32(0)), Lr))
        y = np.ascontiguousarray(np.minimum(np.maximum(st[r, :, 1] * scale, np.float32(0)), Lr))
        c, sn = np.ascontiguousarray(st[r, :, 2]), np.ascontiguousarray(st[r, :, 3])
        acc = 0
        for sw in range(sweep0, sweep0 + nsw):
            a_, _ = oracle.sweep_model(x, y, c, sn, float(Lr), w, sl["step_xy"], sl["step_th"], sl["inv_t_log2e"],
                                       d["key0"], d["key1"], r, sw, spsh)
            acc += a_
        st[r] = np.stack([x, y, c, sn], -1)
        reps["E"][r] = oracle.energy32(x, y, c, sn, float(Lr), w)
        reps["acc"][r] = acc
        if certify:
            side, lam, mx, my = oracle.certify_xycs(st[r])
            reps["Lc"][r], reps["lam"][r], reps["minx"][r], reps["miny"][r] = side, lam, mx, my
        elif d["improve"]:
            reps["Lc"][r] = 1e9
    return st, reps


@pytest.mark.parametrize("case", ["improve", "shrink"])
def test_batch_trajectory_matches_model(sis, oracle, case):
    """A resident instance run one interval at a time equals the oracle run the
    same way, at every interval: replica states, energies, accept counts,
    slots, step sizes, boxes, exchanges, archive, best side.  The certified
    sides agree within 1e-14 relative (FP64 square roots); the model then
    continues from the device's values so the discrete decisions that follow
    are compared bit for bit."""
    if case == "improve":
        n, L, R = 24, 7.0, 16
        p = sis.default_params(replicas=R, ladder=8, sweeps=5 * 40, sweeps_per_exchange=5, tmin=1e-4, tmax=1e-2,
                               shrink=2e-3, improve=1, cool_period=20, certify_every=3, init_mode=1, rng_seed=31,
                               polish_iters=0)
    else:
        n, L, R = 16, 6.0, 12
        p = sis.default_params(replicas=R, ladder=12, sweeps=4 * 40, sweeps_per_exchange=4, tmin=1e-5, tmax=1e-3,
                               shrink=5e-3, improve=0, cool_period=0, init_mode=1, rng_seed=5, polish_iters=0)
    b = sis.BatchSession([(n, L, None)], p)
    s = b.state(0)
    w, spsh = p.cell, p.sweeps_per_shift
    improved = 0
    for iv in range(40):
        say("trajectory", case, iv)
        b.run(1, 0)
        dev = b.state(0)
        # the host's status read back after the interval is the device's descriptor
        assert b.status(0) == dev["desc"], iv
        st, reps = _model_interval(oracle, s, w, spsh, p.certify_every)
        # energies and accept counts come from the sweeps; the states are
        # compared after the exchange step, which may restart replicas
        assert np.array_equal(dev["reps"]["E"], reps["E"]) and np.array_equal(dev["reps"]["acc"], reps["acc"]), iv
        if p.improve:
            fin = reps["Lc"] < 1e8
            assert np.array_equal(dev["reps"]["Lc"] < 1e8, fin), iv
            for f in ("Lc", "lam", "minx", "miny"):
                a_, b_ = dev["reps"][f][fin], reps[f][fin]  # defined where certified
                assert np.all(np.abs(a_ - b_) <= 1e-14 * np.maximum(1, np.abs(b_))), (iv, f)
                reps[f] = dev["reps"][f]
        d, r2, sl2, ros2, st2, ar2, imp = oracle.ladder(s["desc"], reps, s["slots"], s["ros"], st, s["arch"])
        improved += imp
        assert np.array_equal(dev["reps"], r2), iv
        assert np.array_equal(dev["slots"].view(np.uint32), sl2.view(np.uint32)), iv
        assert np.array_equal(dev["ros"], ros2) and np.array_equal(dev["arch"], ar2), iv
        assert np.array_equal(dev["state"], st2), iv
        for k, v in d.items():
            assert dev["desc"][k] == v, (iv, k, dev["desc"][k], v)
        s = dev
    if p.improve:
        assert improved >= 2 and s["desc"]["L"] < L
    else:
        assert s["desc"]["shrinks"] >= 3 and s["desc"]["L"] < L
    b.close()


def test_seed_kept_when_nothing_improves(sis, ref_certifier):
    """The improvement path when no replica certifies below the seed's side
    (the 2 x 2 grid of four squares, side 2.00004, searched at T = 1, where
    every replica overlaps far beyond a dilation of 1.008): the result is the
    seed itself, at the seed's side, and it certifies (the reference writes
    the seed and its L when no chain improves on it, cuda/packing.cu:820,
    859-865)."""
    from refcert import certify_q
    f = 1.00002
    q = np.array([[0.5 * f, 0.5 * f, 0.0], [1.5 * f, 0.5 * f, 0.0], [0.5 * f, 1.5 * f, 0.0], [1.5 * f, 1.5 * f, 0.0]])
    say("seed kept")
    for mode in (0, 2):
        r = sis.solve(4, 2 * f, q, replicas=16, ladder=8, sweeps=200, improve=1, certify_every=1, init_mode=mode,
                      tmin=1.0, tmax=1.0, shrink=1e-3, polish_iters=20, rng_seed=3)
        assert r.feasible == 0 and r.target_L == 2 * f
        assert r.L_cert <= 2 * f + 1e-12
        # no sweep evaluated the exported seed: no FP32 energy and no gap
        assert r.e32 == 0 and r.e_gap == 0 and r.e64_before >= 0
        dil, L50 = certify_q(ref_certifier, r.q)
        assert dil is not None and float(L50) <= 2 * f + 1e-6
