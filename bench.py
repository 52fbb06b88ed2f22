#!/usr/bin/env python3
"""Benchmark of the PT-SA square-packing hot path on one MI355X.

Workload (the plain GPU PT-SA configuration): n = 50 unit squares, 256 replicas
on one geometric temperature ladder 1e-5..5e-2, translation and rotation step
sigma 0.05 adapted to 30 % acceptance, Philox seed 7, centres uniform in
[0.5, L0-0.5]^2 with L0 = ceil(sqrt(50)) + 1 = 9, angles uniform, replica
exchange every 10 sweeps, box shrink 1e-4 on zero overlap, final FP64 polish
and certified-side export.

One step = 1000 sweeps of every replica (100 exchange intervals of 10 sweeps,
run in the plain form: one launch sequence, host synchronisation after every
interval).  One sweep = n proposals per replica, so one step is
256 * 50 * 1000 = 1.28e7 proposals.  --steps K times exactly K steps after
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
                              sweeps_per_exchange=SPE, polish_iters=200, rng_seed=7, init_mode=1)


def cpu_baseline(target_s: float = 12.0):
    """The oracle's C++ FP64 PT-SA (reference semantics) on the same workload
    shape, bounded to about target_s seconds of host time."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib

    o = oracle_lib.load()
    threads = max(1, min(16, os.cpu_count() or 1))
    probe = 20
    r = o.ptsa(n=N, L0=L0, replicas=REPLICAS, tmin=1e-5, tmax=5e-2, sweeps=probe, swap_every=SPE, shrink=1e-4,
               seed=7, threads=threads)
    per_sweep = max(r["seconds"], 1e-6) / probe
    sweeps = int(max(probe, min(20000, target_s / per_sweep)) // SPE * SPE)
    r = o.ptsa(n=N, L0=L0, replicas=REPLICAS, tmin=1e-5, tmax=5e-2, sweeps=sweeps, swap_every=SPE, shrink=1e-4,
               seed=7, threads=threads)
    return {"value": r["proposals"] / r["seconds"], "unit": "proposals/s", "cores": threads, "kind": "port",
            "sample": f"n=50, 256 replicas, {sweeps} sweeps ({r['proposals']} proposals) of the FP64 C++ "
                      f"PT-SA in {r['seconds']:.2f} s on {threads} threads"}


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
        out.append({"n": n, "L0": side, "sweeps": sweeps, "replicas": 128, "target_L": r.target_L,
                    "L_cert": r.L_cert, "certified_slack": r.L_cert - r.target_L, "lambda": r.fields["lambda"],
                    "e32": r.e32, "e64_before": r.e64_before, "e32_vs_e64_gap": r.e_gap,
                    "fp64_rejects": r.fp64_rejects, "feasible": r.feasible, "shrinks": r.shrinks,
                    "certified_L": L50, "dilation": dil, "proposals_per_s": r.proposals / wall,
                    "seconds": wall})
    return out


def batched_leg(sis, n_inst: int, sweeps: int, plain_sample: int, autotune: bool = True):
    """Mixed-n stream (512 instances, n ~ U{20..200} with seed 3, 64 replicas
    each): all instances in one HIP-graph schedule vs the plain form (one
    instance per launch sequence, host sync per interval) on a sample of the
    same stream.  Returns proposals/s of both and the sweep-kernel roofline of
    the batched form."""
    import numpy as np

    rng = np.random.default_rng(3)
    ns = rng.integers(20, 201, n_inst)
    insts = [(int(n), float(math.ceil(math.sqrt(n)) + 1), None) for n in ns]
    params = [sis.default_params(tmin=1e-5, tmax=5e-2, replicas=64, sweeps=sweeps, sweeps_per_exchange=SPE,
                                 rng_seed=1000 + k, init_mode=1, polish_iters=200) for k in range(n_inst)]
    tune = None
    if autotune:
        # autotuner on a 64-instance slice of the stream, 20 probe sweeps per candidate
        tune = sis.autotune(insts[:64], params[:64], probe_sweeps=20)
        for p in params:
            p.replicas_per_group = tune["replicas_per_group"]
            p.sweeps_per_exchange = tune["sweeps_per_exchange"]
            p.cell = tune["cell"]
    spe = params[0].sweeps_per_exchange
    sess = sis.BatchSession(insts, params)
    warm = max(1, 50 // spe)
    sess.run(warm, mode=1)
    t0 = time.perf_counter()
    _, _, dev_ms = sess.run(0, mode=1)
    wall = time.perf_counter() - t0
    res = sess.finalize()
    sess.close()
    timed = sum(n * 64 * (sweeps - warm * spe) for n, _, _ in insts)
    batched = timed / wall
    props = sum(r.proposals for r in res)
    c_per = sum(r.candidates for r in res) / props
    p_per = sum(r.near_pairs for r in res) / props
    feas = sum(r.feasible for r in res)
    # plain form on a sample of the same stream
    plain_props, plain_wall, sweep_ms = 0, 0.0, 0.0
    for k in range(plain_sample):
        s = sis.BatchSession([insts[k]], params[k])
        t1 = time.perf_counter()
        sm, _, _ = s.run(0, mode=0)
        plain_wall += time.perf_counter() - t1
        plain_props += insts[k][0] * 64 * sweeps
        sweep_ms += sm
        s.close()
    plain = plain_props / plain_wall if plain_wall > 0 else None
    return {
        "workload": f"{n_inst} instances, n ~ U{{20..200}} (seed 3), 64 replicas each, {sweeps} sweeps each, "
                    "exchange every 10 sweeps; batched = one HIP-graph schedule over all instances",
        "batched_proposals_per_s": batched, "plain_proposals_per_s": plain,
        "ratio": batched / plain if plain else None,
        "plain_sample_instances": plain_sample, "timed_proposals": timed, "wall_s": wall,
        "device_ms": dev_ms, "feasible_instances": feas, "autotuned": tune,
        "flop_per_proposal": 12.0 * c_per + 28.0 * p_per + 45.0,
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


def served_leg(sis, rate: float, duration: float, sweep_scale: float, plain_seconds: float, seed: int = 5):
    """Serving stream: Poisson arrivals at `rate`/s for `duration` s; n
    log-uniform in 10..1000; sweeps U(5e4, 5e5) x sweep_scale; replicas
    U{32, 64, ..., 256}; half the requests with a 200-iteration FP64 polish,
    half without; every request starts from its certified grid seed and every
    result is certified exactly.  Reports p50/p99 time to a certified L and
    aggregate proposals/s, and the plain form (requests one at a time, host
    sync per interval) on the same requests in arrival order for about
    `plain_seconds`."""
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
    cpu = max(1, min(16, os.cpu_count() or 1))
    srv = sis.Server(max_instances=1024, intervals_per_round=5, certifier_threads=cpu, certify=True,
                     hbm_budget_bytes=200e9)
    tickets = []
    t0 = time.perf_counter()
    for (ta, n, sweeps, R, polish, s) in reqs:
        while time.perf_counter() - t0 < ta:
            time.sleep(0.0005)
        L0, q0 = grid_seed(n)
        tickets.append(srv.submit(n, L0, q0, replicas=R, sweeps=sweeps, polish_iters=polish, rng_seed=s,
                                  init_mode=0, tmin=1e-5, tmax=5e-2))
    results = {}
    while len(results) < len(tickets):
        for tk in tickets:
            if tk not in results:
                r = srv.poll(tk)
                if r is not None:
                    results[tk] = r
        time.sleep(0.002)
    wall = time.perf_counter() - t0
    lat, cert_ok, Ls = [], 0, {}
    for tk in tickets:
        _, _, tc, L50 = srv.times(tk)
        lat.append(tc)
        cert_ok += results[tk].status == 0
        Ls[tk] = L50
        srv.release(tk)
    srv.close()
    props = sum(results[tk].proposals for tk in tickets)
    flops = sum(12.0 * results[tk].candidates + 28.0 * results[tk].near_pairs + 45.0 * results[tk].proposals
                for tk in tickets)
    offered = sum(n * R * sweeps for (_, n, sweeps, R, _, _) in reqs)
    lat = np.array(lat)
    # plain form on the same requests in arrival order; identical results
    pp, pw, k, same = 0, 0.0, 0, 0
    while k < len(reqs) and pw < plain_seconds:
        ta, n, sweeps, R, polish, s = reqs[k]
        L0, q0 = grid_seed(n)
        t1 = time.perf_counter()
        r = sis.solve(n, L0, q0, replicas=R, sweeps=sweeps, polish_iters=polish, rng_seed=s, init_mode=0,
                      tmin=1e-5, tmax=5e-2)
        pw += time.perf_counter() - t1
        pp += r.proposals
        same += bool(np.array_equal(r.q, results[tickets[k]].q) and r.L_cert == results[tickets[k]].L_cert)
        k += 1
    served = props / wall
    return {
        "workload": f"Poisson arrivals {rate}/s for {duration} s ({len(reqs)} requests), n log-uniform 10..1000, "
                    f"sweeps U(5e4,5e5) x {sweep_scale}, replicas U{{32..256}}, polish 0/200 mixed, grid seeds, "
                    "every result certified",
        "requests": len(reqs), "certified": cert_ok,
        "offered_proposals_per_s": offered / duration, "served_proposals_per_s": served,
        "p50_time_to_certified_L_s": float(np.percentile(lat, 50)),
        "p99_time_to_certified_L_s": float(np.percentile(lat, 99)),
        "makespan_s": wall, "plain_proposals_per_s": pp / pw if pw else None,
        "ratio_served_vs_plain": served / (pp / pw) if pw else None,
        "plain_requests": k, "plain_results_identical": same,
        "served_tflops": flops / wall / 1e12,
        "fraction_of_fp32_peak": flops / wall / (PEAK_FP32_TFLOPS * 1e12),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-vs-plain leg")
    ap.add_argument("--batched-instances", type=int, default=512)
    ap.add_argument("--batched-sweeps", type=int, default=400)
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
    # algorithmic FLOP per proposal from the device counters (whole run)
    c_per = res.candidates / res.proposals
    p_per = res.near_pairs / res.proposals  # old + new position
    flop_per_prop = 12.0 * c_per + 28.0 * p_per + 45.0
    achieved = timed_props * flop_per_prop / (sweep_ms * 1e-3) / 1e12 if sweep_ms > 0 else None
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
            "sweeps_timed": a.steps * SWEEPS_PER_STEP, "schedule": "plain (host sync per interval)",
            "device": dev, "compute_units": cus,
        },
        "roofline": {
            "bound": "mfma", "kernel": "k_sweep<64>",
            "roof": "FP32 compute peak; the sweep runs on the FP32 vector ALUs (no dense contraction on this "
                    "path, no matrix-core work); HBM traffic is far below its roof",
            "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": (achieved / PEAK_FP32_TFLOPS) if achieved else None,
            "traffic": traffic_per_launch(),
            "flop_per_proposal": flop_per_prop, "C_per_proposal": c_per, "P_per_proposal": p_per,
            "sweep_kernel_ms_avg": sweep_ms / max(1, launches), "sweep_launches": launches,
        },
        "quality": {
            "feasible": res.feasible, "target_L": res.target_L, "L_cert": res.L_cert, "shrinks": res.shrinks,
            "e32": res.e32, "e64_before_polish": res.e64_before, "e32_vs_e64_gap": res.e_gap,
            "accept_rate": res.accepted / res.proposals,
            "exchange_rate": res.exchanges_accepted / max(1, res.exchanges_tried),
        },
    }
    if not a.no_batched:
        line["batched"] = batched_leg(sis, a.batched_instances, a.batched_sweeps, a.plain_sample)
    if not a.no_served:
        line["served"] = served_leg(sis, a.serve_rate, a.serve_seconds, a.serve_sweep_scale, a.plain_seconds)
    if not a.no_precision:
        line["precision"] = precision_leg(sis, a.precision_sweeps)
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(a.cpu_seconds)
        line["cpu_baseline"]["c1_exact"] = cpu_c1(sis)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
