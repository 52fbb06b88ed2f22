#!/usr/bin/env python3
"""Benchmark of the PT-SA square-packing hot path on one MI355X.

Workload (the plain GPU PT-SA configuration): n = 50 unit squares, 256 replicas
on one geometric temperature ladder 1e-5..5e-2, translation and rotation step
sigma 0.05 adapted to 30 % acceptance, Philox seed 7, centres uniform in
[0.5, L0-0.5]^2 with L0 = ceil(sqrt(50)) + 1 = 9, angles uniform, replica
exchange every 10 sweeps, the reference's improvement path (best certified
side, per-temperature targets best - 1e-4 * mults, ladder heads restarted
from the best configuration), final FP64 polish and certified-side export.

One step = 1000 sweeps of every replica (100 exchange intervals of 10 sweeps,
run in the plain form: one launch pair per interval, the host waiting for every
interval's end with the next interval already queued).  One sweep = n proposals
per replica, so one step is 256 * 50 * 1000 = 1.28e7 proposals.  --steps K times exactly K steps after
--warmup W untimed steps.

Prints one JSON line: proposals/s, the sweep kernel's roofline entry (FP32
VALU: algorithmic FLOP per proposal 2*(6C + 28P) + 45 with C and P counted on
the device), and the CPU baseline (the oracle's FP64 C++ PT-SA on a bounded
sample of the same workload, timed on this host).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "squares-in-squares_amd"))

N, REPLICAS, L0 = 50, 256, float(math.ceil(math.sqrt(50)) + 1)
SWEEPS_PER_STEP, SPE = 1000, 10
PEAK_FP32_TFLOPS = 157.3


def c2_params(sis, sweeps_total):
    return sis.default_params(tmin=1e-5, tmax=5e-2, move_sigma=0.05, rot_sigma=0.05, target_accept=0.3,
                              shrink=1e-4, replicas=REPLICAS, ladder=0, sweeps=sweeps_total,
                              sweeps_per_exchange=SPE, polish_iters=200, rng_seed=7, init_mode=1, improve=1)


def full_length_leg(sis, sweeps: int):
    """The plain workload run to its full length (1e6 sweeps: the timed
    steps are a window of it) in the same plain form, with the polish and the
    export: certified L, the wall time and the rate over the whole run."""
    import numpy as np

    sess = sis.BatchSession([(N, L0, None)], c2_params(sis, sweeps))
    t0 = time.perf_counter()
    sweep_ms, timed_launches, _ = sess.run(0, mode=0)
    wall = time.perf_counter() - t0
    r = sess.finalize()[0]
    sess.close()
    try:
        dil, Ls = sis.export_certify(np.asarray(r.q))
    except Exception as e:  # reported, not hidden
        dil, Ls = None, f"not certified: {e}"
    return {"sweeps": r.sweeps, "proposals": r.proposals, "seconds": wall, "proposals_per_s": r.proposals / wall,
            "sweep_kernel_ms_avg": sweep_ms / max(1, timed_launches), "feasible": r.feasible, "shrinks": r.shrinks, "target_L": r.target_L,
            "L_cert": r.L_cert, "certified_L": Ls, "dilation": dil, "e32_vs_e64_gap": r.e_gap,
            "feasible_checkpoints": r.feasible_checkpoints, "grid_L": float(math.ceil(math.sqrt(N)))}


def host_cpu():
    """Host cores visible to this process and the CPU model."""
    model = "unknown"
    try:
        for ln in Path("/proc/cpuinfo").read_text().splitlines():
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"nproc": avail, "cpu_count": os.cpu_count(), "model": model}


def cpu_baseline(target_s: float = 12.0):
    """The oracle's C++ FP64 PT-SA (reference semantics) on the same workload
    shape, bounded to about target_s seconds of host time."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib

    o = oracle_lib.load()
    threads = max(1, min(16, host_cpu()["nproc"]))
    probe = 20
    r = o.ptsa(n=N, L0=L0, replicas=REPLICAS, tmin=1e-5, tmax=5e-2, sweeps=probe, swap_every=SPE, shrink=1e-4,
               seed=7, threads=threads)
    per_sweep = max(r["seconds"], 1e-6) / probe
    sweeps = int(max(probe, min(20000, target_s / per_sweep)) // SPE * SPE)
    r = o.ptsa(n=N, L0=L0, replicas=REPLICAS, tmin=1e-5, tmax=5e-2, sweeps=sweeps, swap_every=SPE, shrink=1e-4,
               seed=7, threads=threads)
    return {"value": r["proposals"] / r["seconds"], "unit": "proposals/s", "cores": threads, "kind": "port",
            "sample": f"n=50, 256 replicas, {sweeps} sweeps ({r['proposals']} proposals) of the FP64 C++ "
                      f"PT-SA in {r['seconds']:.2f} s on {threads} threads", "host": host_cpu()}


def reference_certifier_check(packings):
    """Every served packing through the reference's own interval certifier
    (built from its source into oracle/_ref), outside any timed region, as the
    checker: the export gate of cuda/demo.py:25-40 (centres dilated by
    1.000002, 1.00001, 1.0001 until accepted).  One CSV holds one puzzle per n,
    so the packings go in rounds of distinct n.  Returns counts, or None when
    the certifier was not built."""
    import subprocess
    import tempfile

    exe = ROOT / "oracle" / "_ref" / "certifier"
    if not exe.exists():
        return None
    pending = list(range(len(packings)))
    passed = {}
    threads = str(max(1, min(16, host_cpu()["nproc"])))
    t0 = time.perf_counter()
    for dil in (1.000002, 1.00001, 1.0001):
        left = pending
        while left:
            seen, batch, rest = set(), [], []
            for k in left:
                n = packings[k].shape[0]
                (rest if n in seen else batch).append(k)
                seen.add(n)
            with tempfile.TemporaryDirectory() as d:
                csv, det = Path(d) / "p.csv", Path(d) / "d.json"
                lines, row = ["id,n,square_id,x,y,angle"], 0
                for k in batch:
                    q = packings[k]
                    for i, (x, y, a) in enumerate(q):
                        lines.append(f"{row},{q.shape[0]},{i},{x * dil:.17g},{y * dil:.17g},{a:.17g}")
                        row += 1
                csv.write_text("\n".join(lines) + "\n")
                maxn = max(packings[k].shape[0] for k in batch)
                subprocess.run([str(exe), str(csv), "--max-n", str(maxn), "--threads", threads, "--details",
                                str(det)], check=True, capture_output=True, timeout=600)
                res = json.loads(det.read_text())["results"]
                for k in batch:
                    r = res[packings[k].shape[0] - 1]
                    if r["L"] is not None:
                        passed[k] = (dil, r["L"])
            left = rest
        pending = [k for k in pending if k not in passed]
        if not pending:
            break
    return {"checked": len(packings), "passed": len(passed), "seconds": time.perf_counter() - t0,
            "dilations": {str(d): sum(1 for v in passed.values() if v[0] == d) for d in (1.000002, 1.00001, 1.0001)}}


def near_pair_replay(sis, insts, params, intervals_runs=None):
    """Near pairs per proposal of a workload from an untimed replay with
    near-pair counting on (FLAG_NEAR_PAIRS).  Counting changes no arithmetic,
    so the replay makes the same proposals as the timed run, which counts only
    the candidates.  intervals_runs: the run() calls of the timed session
    (default: the whole schedule at once).  Returns (P per proposal, results)."""
    sess = sis.BatchSession(insts, [sis.with_near_pairs(p) for p in params])
    for iv, mode in intervals_runs or [(0, 1)]:
        sess.run(iv, mode=mode)
    res = sess.finalize()
    sess.close()
    props = sum(r.proposals for r in res)
    return sum(r.near_pairs for r in res) / max(1, props), res


def traffic_per_launch():
    """HBM bytes per sweep launch of the plain n=50 workload, from the
    FETCH_SIZE and WRITE_SIZE counter runs summarised in
    profiles/k_sweep_traffic.json (each scaled by the same counter's reading
    for a copy of known size), or None when no summary is committed."""
    p = ROOT / "profiles" / "k_sweep_traffic.json"
    if not p.exists():
        return None
    return json.loads(p.read_text()).get("traffic_bytes_per_launch")


def cpu_c1(sis):
    """The CPU configuration run exactly: n=11, L0=5, centres U[0.5, 4.5]^2,
    angles U[0, pi/2), 8 replicas on a geometric ladder 1e-4..1e-1, 2e5 sweeps,
    exchange every 10 sweeps, L -= 1e-4 on a zero-overlap configuration, seed
    42; one thread per replica and one thread alone; then the exact certificate
    of the best configuration (dilation ladder)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib

    o = oracle_lib.load()
    runs = {}
    for threads in (8, 1):
        runs[threads] = o.ptsa(n=11, L0=5.0, replicas=8, tmin=1e-4, tmax=1e-1, sweeps=200000, swap_every=10,
                               shrink=1e-4, seed=42, threads=threads)
    r = runs[8]
    t0 = time.perf_counter()
    try:
        dil, L50 = sis.export_certify(r["best_q"])
    except Exception as e:  # an uncertifiable export is reported, not hidden
        dil, L50 = None, f"not certified: {e}"
    cert_s = time.perf_counter() - t0
    return {
        "workload": "n=11, L0=5, 8 replicas, T 1e-4..1e-1, 2e5 sweeps (1.76e7 proposals), exchange every 10, "
                    "shrink 1e-4, seed 42",
        "threads_8": {"seconds": runs[8]["seconds"], "proposals_per_s": runs[8]["proposals"] / runs[8]["seconds"]},
        "threads_1": {"seconds": runs[1]["seconds"], "proposals_per_s": runs[1]["proposals"] / runs[1]["seconds"]},
        "L_feasible": r["L_feasible"], "shrinks": r["shrinks"], "certified_L": L50, "dilation": dil,
        "certify_seconds": cert_s,
    }


def c1_gpu_leg(sis):
    """The CPU configuration (n=11, L0=5, centres U[0.5, 4.5]^2, angles
    U[0, pi/2), 8 replicas on 1e-4..1e-1, 2e5 sweeps, exchange every 10, seed
    42) on the GPU, in the plain form, under its own rule (L -= 1e-4 on a
    zero-overlap configuration) and under the reference's improvement path
    with the 200-interval reheating cycle (the `packing` CLI's search); each
    result certified exactly."""
    import numpy as np

    out = {}
    for name, improve, cool in (("shrink_on_zero_overlap", 0, 0), ("improvement_path", 1, 200)):
        t0 = time.perf_counter()
        r = sis.solve(11, 5.0, None, replicas=8, sweeps=200000, sweeps_per_exchange=10, tmin=1e-4, tmax=1e-1,
                      shrink=1e-4, rng_seed=42, init_mode=1, improve=improve, cool_period=cool, polish_iters=200)
        wall = time.perf_counter() - t0
        try:
            dil, L50 = sis.export_certify(np.asarray(r.q))
        except Exception as e:  # reported, not hidden
            dil, L50 = None, f"not certified: {e}"
        out[name] = {"certified_L": L50, "dilation": dil, "L_cert": r.L_cert, "shrinks": r.shrinks,
                     "seconds": wall, "proposals_per_s": r.proposals / wall}
    out["grid_L"] = 4.0
    out["reference_recorded_target_L"] = 3.890074968  # cuda/results/raw11.json
    return out


def precision_leg(sis, sweeps: int = 10000):
    """n = 500 and 1000 started from the grid of side ceil(sqrt(n)) with
    jitter U(+-0.02) on the centres and U(+-0.01 rad) on the angles, 128
    replicas, FP32 sweeps, 200 FP64 polish iterations, FP64 L-shrink test and
    certified side.  Reports the FP32-vs-FP64 energy gap of the export, the
    FP32 zero-overlap configurations the FP64 test rejected, the certified
    slack and the exact certificate."""
    import numpy as np

    out = []
    for n in (500, 1000):
        side = math.isqrt(n - 1) + 1
        rng = np.random.default_rng(n)
        i = np.arange(n)
        q = np.zeros((n, 3))
        q[:, 0] = i % side + 0.5 + rng.uniform(-0.02, 0.02, n)
        q[:, 1] = i // side + 0.5 + rng.uniform(-0.02, 0.02, n)
        q[:, 2] = rng.uniform(-0.01, 0.01, n)
        t0 = time.perf_counter()
        r = sis.solve(n, float(side), q, replicas=128, sweeps=sweeps, polish_iters=200, init_mode=0,
                      init_jitter=0.0, tmin=1e-6, tmax=1e-2, shrink=1e-3, rng_seed=n)
        wall = time.perf_counter() - t0
        try:
            dil, L50 = sis.export_certify(r.q)
        except Exception as e:
            dil, L50 = None, f"not certified: {e}"
        grid_L = float(side)  # the unjittered grid: squares in contact, certified side exactly `side`
        out.append({"n": n, "L0": side, "sweeps": sweeps, "replicas": 128, "target_L": r.target_L,
                    "L_cert": r.L_cert, "certified_slack": r.L_cert - r.target_L,
                    "L_cert_vs_L0": r.L_cert / side - 1.0,
                    "note": "the box is the grid's own side: the unjittered grid fills it exactly, and a row of "
                            "squares tilted by 0.01 rad is 1 %% wider than the box, so zero overlap at L0 needs a "
                            "global rearrangement into the %d free cells; the certified side of the unjittered "
                            "grid is %g" % (side * side - n, grid_L),
                    "lambda": r.fields["lambda"],
                    "e32": r.e32, "e64_before": r.e64_before, "e32_vs_e64_gap": r.e_gap,
                    "fp64_rejects": r.fp64_rejects, "feasible": r.feasible, "shrinks": r.shrinks,
                    "certified_L": L50, "dilation": dil, "proposals_per_s": r.proposals / wall,
                    "seconds": wall})
    return out


def batched_leg(sis, n_inst: int, sweeps: int, plain_sample: int, autotune: bool = True):
    """Mixed-n stream (512 instances, n ~ U{20..200} with seed 3, 64 replicas
    each): all instances in one HIP-graph schedule vs the plain form (one
    instance per launch sequence, the host waiting for every interval) on a sample of the
    same stream.  Returns proposals/s of both and the sweep-kernel roofline of
    the batched form; with the autotuner, also the rate and search quality of
    the whole stream under the untuned settings."""
    import numpy as np

    rng = np.random.default_rng(3)
    ns = rng.integers(20, 201, n_inst)
    insts = [(int(n), float(math.ceil(math.sqrt(n)) + 1), None) for n in ns]
    params = [sis.default_params(tmin=1e-5, tmax=5e-2, replicas=64, sweeps=sweeps, sweeps_per_exchange=SPE,
                                 rng_seed=1000 + k, init_mode=1, polish_iters=200, shrink=1e-4, improve=1)
              for k in range(n_inst)]

    def stream(ps):
        spe = ps[0].sweeps_per_exchange
        sess = sis.BatchSession(insts, ps)
        warm = max(1, 50 // spe)
        sess.run(warm, mode=1)
        t0 = time.perf_counter()
        _, _, dev_ms = sess.run(0, mode=1)
        wall = time.perf_counter() - t0
        res = sess.finalize()
        sess.close()
        return res, wall, dev_ms, sum(n * 64 * (sweeps - warm * spe) for n, _, _ in insts)

    def reduction(res):
        gain = [1.0 - r.L_cert / L for r, (_, L, _) in zip(res, insts) if r.L_cert < 1e8]
        return float(np.mean(gain)) if gain else None

    tune, untuned = None, None
    if autotune:
        # autotuner on a 64-instance slice of the stream, 20 probe sweeps per candidate
        tune = sis.autotune(insts[:64], params[:64], probe_sweeps=20)
        # the untuned settings on the whole stream, for the tuner's speed and
        # search quality at full length (the timed run itself when the tuner
        # kept them)
        own = (params[0].replicas_per_group, params[0].sweeps_per_exchange, params[0].cell)
        if own == (tune["replicas_per_group"], tune["sweeps_per_exchange"], tune["cell"]):
            untuned = {"same_settings_as_tuned": True}
        else:
            u_res, u_wall, _, u_timed = stream(params)
            untuned = {"replicas_per_group": own[0], "sweeps_per_exchange": own[1], "cell": own[2],
                       "proposals_per_s": u_timed / u_wall, "feasible_instances": sum(r.feasible for r in u_res),
                       "mean_relative_side_reduction": reduction(u_res)}
        for p in params:
            p.replicas_per_group = tune["replicas_per_group"]
            p.sweeps_per_exchange = tune["sweeps_per_exchange"]
            p.cell = tune["cell"]
    spe = params[0].sweeps_per_exchange
    res, wall, dev_ms, timed = stream(params)
    batched = timed / wall
    props = sum(r.proposals for r in res)
    c_per = sum(r.candidates for r in res) / props
    # near pairs from an untimed replay of the stream's first 64 instances
    # (counting changes no trajectory)
    p_per, _ = near_pair_replay(sis, insts[:64], params[:64])
    feas = sum(r.feasible for r in res)
    # plain form on a sample of the same stream, same sweeps: same results
    plain_props, plain_wall, sweep_ms, same = 0, 0.0, 0.0, 0
    plain_L = []
    for k in range(plain_sample):
        s = sis.BatchSession([insts[k]], params[k])
        t1 = time.perf_counter()
        sm, _, _ = s.run(0, mode=0)
        plain_wall += time.perf_counter() - t1
        plain_props += insts[k][0] * 64 * sweeps
        sweep_ms += sm
        pr = s.finalize()[0]
        s.close()
        plain_L.append(pr.L_cert)
        same += bool(np.array_equal(pr.q, res[k].q) and pr.L_cert == res[k].L_cert)
    plain = plain_props / plain_wall if plain_wall > 0 else None
    return {
        "workload": f"{n_inst} instances, n ~ U{{20..200}} (seed 3), 64 replicas each, {sweeps} sweeps each, "
                    f"exchange every {spe} sweeps (the autotuner's choice below; {SPE} untuned), random starts in a box of side ceil(sqrt n) + 1, "
                    "the reference's improvement path; batched = one HIP-graph schedule over all instances",
        "batched_proposals_per_s": batched, "plain_proposals_per_s": plain,
        "ratio": batched / plain if plain else None,
        "plain_sample_instances": plain_sample, "timed_proposals": timed, "wall_s": wall,
        "device_ms": dev_ms, "feasible_instances": feas, "autotuned": tune, "untuned": untuned,
        "quality": {"instances_improved": feas, "mean_relative_side_reduction": reduction(res),
                    "mean_L_cert_sample_batched": float(np.mean([r.L_cert for r in res[:plain_sample]]))
                    if plain_sample else None,
                    "mean_L_cert_sample_plain": float(np.mean(plain_L)) if plain_L else None,
                    "plain_sample_identical": same},
        "flop_per_proposal": 12.0 * c_per + 28.0 * p_per + 45.0, "C_per_proposal": c_per,
        "P_per_proposal": p_per, "P_source": "untimed replay of the first 64 instances with near-pair counting",
        "achieved_tflops_wall": timed * (12.0 * c_per + 28.0 * p_per + 45.0) / wall / 1e12,
    }


def grid_seed(n: int):
    """Certified start of a request: the half-integer grid of side
    ceil(sqrt(n)) with centres and side dilated by 1.00002, as the reference
    seeds its searches (cuda/demo.py:14-23)."""
    import numpy as np

    w = math.isqrt(n - 1) + 1
    i = np.arange(n)
    q = np.zeros((n, 3))
    q[:, 0] = (i % w + 0.5) * 1.00002
    q[:, 1] = (i // w + 0.5) * 1.00002
    return float(w * 1.00002), q


SIZE_CLASSES = ((10, 63), (64, 255), (256, 1000))


def served_leg(sis, rate: float, duration: float, sweep_scale: float, plain_seconds: float, seed: int = 5):
    """Serving stream: Poisson arrivals at `rate`/s for `duration` s; n
    log-uniform in 10..1000; sweeps U(5e4, 5e5) x sweep_scale; replicas
    U{32, 64, ..., 256}; half the requests with a 200-iteration FP64 polish,
    half without; every request starts from its certified grid seed, runs the
    reference's improvement path and is certified exactly.  Reports p50/p99
    time to a certified L and aggregate proposals/s, and the plain form
    (requests one at a time, the host waiting for every interval) on the same requests,
    about `plain_seconds` split over three size classes (n < 64, 64-255,
    >= 256), each class in arrival order.  Returns the line and the packings."""
    import numpy as np

    rng = np.random.default_rng(seed)
    reqs = []
    t = 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        n = int(round(math.exp(rng.uniform(math.log(10), math.log(1000)))))
        sweeps = max(SPE, int(rng.uniform(5e4, 5e5) * sweep_scale) // SPE * SPE)
        R = int(rng.integers(1, 9)) * 32
        polish = 200 if rng.random() < 0.5 else 0
        reqs.append((t, n, sweeps, R, polish, 100 + len(reqs)))
    cpu = max(1, min(16, host_cpu()["nproc"]))
    # 16 GB of device arena: the stream holds at most about 4 GB of replica state
    srv = sis.Server(max_instances=1024, intervals_per_round=5, certifier_threads=cpu, certify=True,
                     hbm_budget_bytes=16e9)

    def params(n, sweeps, R, polish, s):
        return sis.default_params(replicas=R, sweeps=sweeps, polish_iters=polish, rng_seed=s, init_mode=0,
                                  tmin=1e-5, tmax=5e-2, shrink=1e-4, improve=1)

    tickets = []
    t0 = time.perf_counter()
    for (ta, n, sweeps, R, polish, s) in reqs:
        while time.perf_counter() - t0 < ta:
            time.sleep(0.0005)
        L0, q0 = grid_seed(n)
        tickets.append(srv.submit(n, L0, q0, params=params(n, sweeps, R, polish, s)))
    results = {}
    t_note = time.perf_counter()
    while len(results) < len(tickets):
        for tk in tickets:
            if tk not in results:
                r = srv.poll(tk)
                if r is not None:
                    results[tk] = r
        time.sleep(0.002)
        if time.perf_counter() - t_note > 20.0:
            t_note = time.perf_counter()
            print(f"served: {len(results)}/{len(tickets)} requests done after {t_note - t0:.0f} s",
                  file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    lat, cert_ok, done_at = [], 0, []
    for tk, (ta, *_rest) in zip(tickets, reqs):
        _, _, tc, _L50 = srv.times(tk)
        lat.append(tc)
        done_at.append(ta + tc)
        cert_ok += results[tk].status == 0
        srv.release(tk)
    srv.close()
    props = sum(results[tk].proposals for tk in tickets)
    # near pairs per proposal from an untimed replay, with near-pair
    # counting, of every k-th request (at most 32) as one batch
    step = max(1, -(-len(reqs) // 32))
    sample = reqs[::step]
    p_per, _ = near_pair_replay(sis, [(n,) + grid_seed(n) for (_, n, *_r) in sample],
                                [params(n, sweeps, R, polish, s_) for (_, n, sweeps, R, polish, s_) in sample])
    flops = sum(12.0 * results[tk].candidates + (28.0 * p_per + 45.0) * results[tk].proposals for tk in tickets)
    offered = sum(n * R * sweeps for (_, n, sweeps, R, _, _) in reqs)
    lat = np.array(lat)
    served = props / wall
    # plain form on the same requests: about plain_seconds / 3 per size class,
    # in arrival order; results identical to the served ones
    classes = []
    pp_all, pw_all, same_all, k_all = 0, 0.0, 0, 0
    for lo, hi in SIZE_CLASSES:
        idx = [k for k, r in enumerate(reqs) if lo <= r[1] <= hi]
        pp, pw, same, done = 0, 0.0, 0, 0
        for k in idx:
            if pw >= plain_seconds / len(SIZE_CLASSES):
                break
            ta, n, sweeps, R, polish, s = reqs[k]
            L0, q0 = grid_seed(n)
            t1 = time.perf_counter()
            r = sis.solve(n, L0, q0, params=params(n, sweeps, R, polish, s))
            pw += time.perf_counter() - t1
            pp += r.proposals
            same += bool(np.array_equal(r.q, results[tickets[k]].q) and r.L_cert == results[tickets[k]].L_cert)
            done += 1
        cprops = sum(results[tickets[k]].proposals for k in idx)
        span = (max(done_at[k] for k in idx) - min(reqs[k][0] for k in idx)) if idx else 0.0
        c_served = cprops / span if span > 0 else None
        c_plain = pp / pw if pw > 0 else None
        classes.append({"n": f"{lo}-{hi}", "requests": len(idx), "proposals": cprops,
                        "served_proposals_per_s": c_served, "plain_proposals_per_s": c_plain,
                        "ratio_served_vs_plain": c_served / c_plain if c_served and c_plain else None,
                        "p50_time_to_certified_L_s": float(np.percentile(lat[idx], 50)) if idx else None,
                        "plain_requests": done, "plain_results_identical": same})
        pp_all += pp
        pw_all += pw
        same_all += same
        k_all += done
    # the stream's own mix, at the plain rates of its classes
    plain_mix = None
    if all(c["plain_proposals_per_s"] for c in classes if c["requests"]):
        plain_mix = props / sum(c["proposals"] / c["plain_proposals_per_s"] for c in classes if c["requests"])
    line = {
        "workload": f"Poisson arrivals {rate}/s for {duration} s ({len(reqs)} requests), n log-uniform 10..1000, "
                    f"sweeps U(5e4,5e5) x {sweep_scale}, replicas U{{32..256}}, polish 0/200 mixed, grid seeds, "
                    "the reference's improvement path, every result certified",
        "requests": len(reqs), "certified": cert_ok,
        "offered_proposals_per_s": offered / duration, "served_proposals_per_s": served,
        "p50_time_to_certified_L_s": float(np.percentile(lat, 50)),
        "p99_time_to_certified_L_s": float(np.percentile(lat, 99)),
        "makespan_s": wall, "plain_proposals_per_s": pp_all / pw_all if pw_all else None,
        "ratio_served_vs_plain": served / (pp_all / pw_all) if pw_all else None,
        "plain_proposals_per_s_stream_mix": plain_mix,
        "ratio_served_vs_plain_stream_mix": served / plain_mix if plain_mix else None,
        "plain_requests": k_all, "plain_results_identical": same_all,
        "size_classes": classes,
        "P_per_proposal": p_per, "P_source": f"untimed replay of {len(sample)} requests with near-pair counting",
        "served_tflops": flops / wall / 1e12,
        "fraction_of_fp32_peak": flops / wall / (PEAK_FP32_TFLOPS * 1e12),
    }
    return line, [results[tk].q for tk in tickets]


def tilted_five_rows():
    """The tilted five of the reference's demo seed for n = 5
    (src/squares/optimizer.py:13-36): four corner squares and a central one at
    45 degrees in a box of side 2 + 1/sqrt(2), centres dilated by 1.000000001."""
    side = 2 + 1 / math.sqrt(2)
    f = 1.000000001
    pts = [(0.5, 0.5, 0.0), (side - 0.5, 0.5, 0.0), (0.5, side - 0.5, 0.0), (side - 0.5, side - 0.5, 0.0),
           (side / 2, side / 2, math.pi / 4)]
    return [(x * f, y * f, a) for x, y, a in pts]


def reference_seed_leg(sis):
    """The drop-in `packing` binary on the reference's demo workload
    (cuda/demo.py:11-23, 72-86: grid seeds, the tilted five for n = 5,
    coordinates and side x 1.00002, `packing seed 20000 32 256 42`): certified
    L (this engine's exact certifier) beside the recorded target_L of
    cuda/results/raw{n}.json, and with `steps` counted as sweeps."""
    import subprocess
    import tempfile

    import numpy as np

    exe = ROOT / "squares-in-squares_amd" / "bin" / "packing"
    rows_out = []
    with tempfile.TemporaryDirectory() as d:
        for n in (5, 10, 11, 17, 50):
            if n == 5:
                q = np.asarray(tilted_five_rows())
            else:
                w = math.isqrt(n - 1) + 1
                q = np.asarray([((i % w) + 0.5, (i // w) + 0.5, 0.0) for i in range(n)])
            side = float(sis.certify_strings([tuple(format(float(v), ".17g") for v in r) for r in q]))
            seed = Path(d) / f"seed{n}.txt"
            seed.write_text(f"{n} {side * 1.00002:.12g}\n" + "".join(
                f"{x * 1.00002:.12g} {y * 1.00002:.12g} {a:.12g}\n" for x, y, a in q))
            rec = json.loads((ROOT / "tests" / "golden" / "raw" / f"raw{n}.json").read_text())
            row = {"n": n, "seed_L": side * 1.00002, "recorded_target_L": rec["target_L"]}
            for name, env in (("reference_argv", {}), ("steps_as_sweeps", {"PACK_SWEEP_STEPS": "1"})):
                out = Path(d) / f"raw{n}_{name}.json"
                t0 = time.perf_counter()
                subprocess.run([str(exe), str(seed), "20000", "32", "256", "42", str(out)], check=True,
                               env={**os.environ, **env}, capture_output=True, timeout=600)
                wall = time.perf_counter() - t0
                raw = json.loads(out.read_text())
                try:
                    dil, L = sis.export_certify(np.asarray(raw["q"]))
                except Exception as e:  # reported, not hidden
                    dil, L = None, f"not certified: {e}"
                row[name] = {"certified_L": L, "dilation": dil, "seconds": wall, "proposals": raw["proposals"]}
            rows_out.append(row)
    return rows_out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-vs-plain leg")
    ap.add_argument("--batched-instances", type=int, default=512)
    ap.add_argument("--batched-sweeps", type=int, default=200000)
    ap.add_argument("--plain-sample", type=int, default=8)
    ap.add_argument("--no-served", action="store_true", help="skip the serving leg")
    ap.add_argument("--batched-only", action="store_true", help="run only the batched leg (profiling)")
    ap.add_argument("--copy-calibration", action="store_true", help="one 256 MiB device copy (profiling)")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-seconds", type=float, default=10.0)
    ap.add_argument("--serve-sweep-scale", type=float, default=0.1)
    ap.add_argument("--plain-seconds", type=float, default=8.0)
    ap.add_argument("--no-precision", action="store_true", help="skip the n=500/1000 precision leg")
    ap.add_argument("--precision-sweeps", type=int, default=10000)
    ap.add_argument("--no-reference-seeds", action="store_true", help="skip the demo-seed quality leg")
    ap.add_argument("--full-sweeps", type=int, default=1000000,
                    help="sweeps of the untimed full-length run of the plain workload (0: skip)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {a.gpus} is not supported (requests shard across GPUs "
              "by instance, with no inter-GPU traffic during the search)", file=sys.stderr)
        sys.exit(2)

    if a.copy_calibration:
        # profiling aid: one device copy of exactly 256 MiB (read 256 MiB, write
        # 256 MiB) to scale the FETCH_SIZE / WRITE_SIZE counters
        import torch

        x = torch.ones(1 << 26, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": x.numel() * 4}), flush=True)
        return

    import numpy as np
    import sisamd as sis

    cus, dev = sis.device_info()
    if a.batched_only:
        # profiling aid: the batched leg alone, no plain sample, no CPU leg
        print(json.dumps(batched_leg(sis, a.batched_instances, a.batched_sweeps, 0, autotune=False)), flush=True)
        return
    total_steps = a.warmup + a.steps
    p = c2_params(sis, total_steps * SWEEPS_PER_STEP)
    sess = sis.BatchSession([(N, L0, None)], p)
    intervals_per_step = SWEEPS_PER_STEP // SPE
    for _ in range(a.warmup):
        sess.run(intervals_per_step, mode=0)
    t0 = time.perf_counter()
    sweep_ms = 0.0
    launches = 0
    for _ in range(a.steps):
        sm, nl, _ = sess.run(intervals_per_step, mode=0)
        sweep_ms += sm
        launches += nl
    wall = time.perf_counter() - t0
    res = sess.finalize()[0]
    sess.close()

    timed_props = a.steps * SWEEPS_PER_STEP * N * REPLICAS
    value = timed_props / wall
    # algorithmic FLOP per proposal from the device counters (whole run); the
    # near pairs (old + new position) from an untimed replay of the same run
    c_per = res.candidates / res.proposals
    p_per, rep = near_pair_replay(sis, [(N, L0, None)], [p], [(intervals_per_step, 0)] * total_steps)
    replay_same = bool(rep[0].L_cert == res.L_cert and rep[0].accepted == res.accepted
                       and np.array_equal(rep[0].q, res.q))
    flop_per_prop = 12.0 * c_per + 28.0 * p_per + 45.0
    # every launch of the timed region sweeps SPE sweeps of every replica; the
    # plain loop times every eighth one with HIP events on its stream
    launches_total = a.steps * intervals_per_step
    avg_ms = sweep_ms / launches if launches else None
    flop_per_launch = timed_props / launches_total * flop_per_prop
    achieved = flop_per_launch / (avg_ms * 1e-3) / 1e12 if avg_ms else None
    line = {
        "metric": "MC proposals/s; p50/p99 time-to-certified-L, served batched vs plain GPU",
        "value": value,
        "unit": "proposals/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded uniform random initial packings)",
        "config": {
            "workload": "plain PT-SA on GPU: n=50, 256 replicas, geometric ladder 1e-5..5e-2, sigma 0.05/0.05 "
                        "adapted to 30% acceptance, exchange every 10 sweeps, Philox seed 7, FP64 polish at end; "
                        "interaction cells of side sqrt(2) realised as 5x5 stencils of colour cells of side "
                        "0.7075",
            "n": N, "replicas": REPLICAS, "L0": L0, "sweeps_per_step": SWEEPS_PER_STEP,
            "sweeps_timed": a.steps * SWEEPS_PER_STEP, "schedule": "plain (host waits for every interval's end, one interval queued ahead)",
            "device": dev, "compute_units": cus,
        },
        "roofline": {
            "bound": "valu", "kernel": "k_sweep<64, false, 1> (one-wavefront tiles; compiled for one wave per SIMD)",
            "roof": "FP32 vector ALU peak (157.3 TFLOP/s); no dense contraction on this path, so no matrix-core "
                    "work; HBM traffic is far below its roof; at this load (one wavefront per compute unit) the "
                    "launch is latency-bound",
            "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": (achieved / PEAK_FP32_TFLOPS) if achieved else None,
            "traffic": traffic_per_launch(),
            "flop_per_proposal": flop_per_prop, "C_per_proposal": c_per, "P_per_proposal": p_per,
            "P_source": "untimed replay of the same run with near-pair counting (identical: %s)" % replay_same,
            "sweep_kernel_ms_avg": avg_ms, "sweep_launches_timed": launches, "sweep_launches": launches_total,
            "flop_per_launch": flop_per_launch,
        },
        "quality": {
            "feasible": res.feasible, "target_L": res.target_L, "L_cert": res.L_cert, "shrinks": res.shrinks,
            "e32": res.e32, "e64_before_polish": res.e64_before, "e32_vs_e64_gap": res.e_gap,
            "accept_rate": res.accepted / res.proposals,
            "exchange_rate": res.exchanges_accepted / max(1, res.exchanges_tried),
        },
    }
    if a.full_sweeps > 0:
        line["full_length"] = full_length_leg(sis, a.full_sweeps)
    if not a.no_batched:
        line["batched"] = batched_leg(sis, a.batched_instances, a.batched_sweeps, a.plain_sample)
    packings = []
    if not a.no_served:
        line["served"], packings = served_leg(sis, a.serve_rate, a.serve_seconds, a.serve_sweep_scale,
                                              a.plain_seconds)
    if not a.no_precision:
        line["precision"] = precision_leg(sis, a.precision_sweeps)
    if not a.no_reference_seeds:
        line["reference_seeds"] = reference_seed_leg(sis)
        line["cpu_config_on_gpu"] = c1_gpu_leg(sis)
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(a.cpu_seconds)
        line["cpu_baseline"]["c1_exact"] = cpu_c1(sis)
        if packings:
            line["cpu_baseline"]["served_reference_certifier"] = reference_certifier_check(packings)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
