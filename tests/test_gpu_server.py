"""Serving on the GPU: continuous batching admits requests of different n while
others run; every result is bit-identical to the same request solved alone,
and passes the exact certifier."""
from __future__ import annotations

import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)


def _wait(srv, tickets, timeout=120):
    t0 = time.time()
    out = {}
    while len(out) < len(tickets):
        for t in tickets:
            if t not in out:
                r = srv.poll(t)
                if r is not None:
                    out[t] = r
        if time.time() - t0 > timeout:
            raise TimeoutError("server did not finish")
        time.sleep(0.005)
    return out


def test_server_results_equal_alone(sis):
    # every size class of the loop (n < 64, n < 256, larger); the last request
    # stays infeasible, so its export is the best replica picked on the device
    reqs = [(15, 5.5, dict(replicas=32, sweeps=300, rng_seed=11)),
            (40, 8.0, dict(replicas=16, sweeps=500, rng_seed=12)),
            (90, 11.0, dict(replicas=8, sweeps=200, rng_seed=13)),
            (25, 6.0, dict(replicas=24, sweeps=150, rng_seed=14)),
            (300, 17.4, dict(replicas=8, sweeps=60, rng_seed=15))]
    say("server mix")
    srv = sis.Server(max_instances=8, intervals_per_round=3, certifier_threads=2, certify=True)
    tickets = []
    for k, (n, L, kw) in enumerate(reqs):
        tickets.append(srv.submit(n, L, None, init_mode=1, tmin=1e-4, tmax=3e-2, **kw))
        if k == 1:
            time.sleep(0.05)  # later arrivals join a running batch
    res = _wait(srv, tickets)
    times = [srv.times(t) for t in tickets]
    srv.close()
    assert not res[tickets[-1]].feasible
    for (n, L, kw), t, (tq, td, tc, L50) in zip(reqs, tickets, times):
        say("alone", n)
        alone = sis.solve(n, L, None, init_mode=1, tmin=1e-4, tmax=3e-2, **kw)
        m = res[t]
        assert np.array_equal(alone.q, m.q)
        for k in ("proposals", "accepted", "shrinks", "exchanges_accepted", "L_cert", "target_L", "e64_after",
                  "e32", "e64_before"):
            assert alone.fields[k] == m.fields[k], k
        assert 0 <= tq <= td <= tc
        if m.feasible:
            d, L_alone = sis.export_certify(alone.q)
            assert L50 == L_alone


def test_server_admission_under_slot_cap(sis):
    """More requests than resident slots: they queue and are admitted as slots
    free up; all finish."""
    say("server queue")
    srv = sis.Server(max_instances=2, intervals_per_round=2, certifier_threads=1, certify=False)
    tickets = [srv.submit(12 + 3 * k, 5.0, None, replicas=8, sweeps=100, init_mode=1, rng_seed=k) for k in range(6)]
    res = _wait(srv, tickets)
    admits = sorted(srv.times(t)[0] for t in tickets)
    assert len(res) == 6
    assert admits[-1] > admits[0]  # the last ones waited for a slot
    # finished tickets can be forgotten; a running one is kept
    late = srv.submit(30, 6.5, None, replicas=8, sweeps=20000, init_mode=1, rng_seed=99)
    assert srv.release(late) is False
    for t in tickets:
        assert srv.release(t) is True
        with pytest.raises(sis.SisError):
            srv.times(t)
    assert _wait(srv, [late])[late] is not None
    assert srv.release(late) is True
    srv.close()


def _outcome(sis, srv, tickets, timeout=120):
    """ticket -> ("ok", result) or ("error", code), polling until all are in."""
    t0 = time.time()
    out = {}
    while len(out) < len(tickets):
        for t in tickets:
            if t in out:
                continue
            try:
                r = srv.poll(t)
            except sis.SisError as e:
                out[t] = ("error", e.code)
                continue
            if r is not None:
                out[t] = ("ok", r)
        if time.time() - t0 > timeout:
            raise TimeoutError("server did not finish")
        time.sleep(0.005)
    return out


def test_lane_fault_fails_only_that_lane(sis):
    """A failed round (forced from the host at lane 0's next sweep launch, so
    that no kernel faults) fails the requests resident in that lane with
    SIS_EDEVICE; every other request (the other lanes', and a lane-0 request
    admitted after the failure) finishes with the result of its plain solve,
    and lane 0 serves the next request normally."""
    say("server lane fault")
    srv = sis.Server(max_instances=8, intervals_per_round=2, certifier_threads=2, certify=True)
    srv.inject_lane_fault(0)
    kw = dict(init_mode=1, tmin=1e-4, tmax=3e-2, improve=1)
    reqs = [(20, 6.0, 16, 200, 40), (21, 6.0, 16, 200, 41), (300, 17.4, 8, 60, 50), (120, 12.0, 8, 100, 51)]
    tickets = [srv.submit(n, L, None, replicas=R, sweeps=sw, rng_seed=seed, **kw) for n, L, R, sw, seed in reqs]
    out = _outcome(sis, srv, tickets)
    failed = [t for t in tickets if out[t][0] == "error"]
    assert failed and failed[0] == tickets[0]  # the first request is in lane 0's first round
    for t, (n, L, R, sw, seed) in zip(tickets, reqs):
        if t in failed:
            assert n < 64 and out[t] == ("error", sis.SIS_EDEVICE)
            assert srv.release(t) is True
            continue
        r = out[t][1]
        alone = sis.solve(n, L, None, replicas=R, sweeps=sw, rng_seed=seed, **kw)
        assert np.array_equal(alone.q, r.q) and alone.L_cert == r.L_cert and r.sweeps == sw
    again = srv.submit(21, 6.0, None, replicas=16, sweeps=200, rng_seed=41, **kw)
    kind, r = _outcome(sis, srv, [again])[again]
    assert kind == "ok" and r.status == 0
    alone = sis.solve(21, 6.0, None, replicas=16, sweeps=200, rng_seed=41, **kw)
    assert np.array_equal(alone.q, r.q)
    assert srv.stats()["resident_requests"] == 0
    srv.close()


def test_deadline_stops_a_running_request(sis):
    """A request past its deadline stops at its lane's next round boundary and
    returns the best configuration so far, certified, with status
    SIS_ETIMEOUT; a request without a deadline on the same server runs out."""
    say("server deadline running")
    srv = sis.Server(max_instances=4, intervals_per_round=5, certifier_threads=2, certify=True,
                     request_timeout_s=0.0)
    L0 = 7.00014
    q0 = np.array([((i % 7) + 0.5, (i // 7) + 0.5, 0.0) for i in range(45)]) * [1.00002, 1.00002, 1]
    t_long = srv.submit(45, L0, q0, replicas=64, sweeps=10**7, init_mode=0, improve=1, rng_seed=3, timeout_s=1.0)
    t_short = srv.submit(30, 6.5, None, replicas=16, sweeps=300, init_mode=1, rng_seed=4)
    t0 = time.time()
    out = _outcome(sis, srv, [t_long, t_short], timeout=60)
    took = time.time() - t0
    kind, r = out[t_long]
    assert kind == "ok" and r.status == sis.SIS_ETIMEOUT
    assert 0 < r.sweeps < 10**7 and r.proposals == r.sweeps * 45 * 64
    assert r.L_cert <= L0 + 1e-9
    _, _, _, L50 = srv.times(t_long)
    assert L50[:2] in ("6.", "7.")  # the exact certificate of the partial result
    assert took < 30
    kind, r2 = out[t_short]
    assert kind == "ok" and r2.status == 0 and r2.sweeps == 300
    srv.close()


def test_deadline_passes_in_the_queue(sis):
    """With one resident slot, a request whose deadline passes while it waits
    behind another is never admitted (sis_poll: SIS_ETIMEOUT, no result); the
    server's request_timeout_s applies to requests submitted without one."""
    say("server deadline queued")
    srv = sis.Server(max_instances=1, intervals_per_round=5, certifier_threads=1, certify=False,
                     request_timeout_s=1.5)
    first = srv.submit(40, 8.0, None, replicas=32, sweeps=10**7, init_mode=1, rng_seed=5)
    waits = srv.submit(20, 6.0, None, replicas=8, sweeps=100, init_mode=1, rng_seed=6, timeout_s=0.3)
    out = _outcome(sis, srv, [first, waits], timeout=60)
    assert out[waits] == ("error", sis.SIS_ETIMEOUT)
    tq, td, tc, _ = srv.times(waits)
    assert tq < 0 and td >= 0.3
    kind, r = out[first]
    assert kind == "ok" and r.status == sis.SIS_ETIMEOUT and r.sweeps < 10**7
    srv.close()


def test_arena_follows_the_budget(sis):
    """The arena follows hbm_budget_bytes (here 6 GiB, past the default
    4 GiB): 70 requests of 1000 squares x 4096 replicas (4.6 GB of replica
    state), each searching until its 15 s deadline (admitting all 70 takes
    several seconds: each uploads 65 MB of replica state), are resident at
    once, and all return their results."""
    say("server arena")
    budget = 6 * 2**30
    srv = sis.Server(max_instances=80, intervals_per_round=1, certifier_threads=1, certify=False,
                     hbm_budget_bytes=budget)
    st = srv.stats()
    assert 0.9 * budget <= st["arena_bytes"] <= budget
    L0 = 32 * 1.00002
    q0 = np.array([((i % 32) + 0.5, (i // 32) + 0.5, 0.0) for i in range(1000)]) * [1.00002, 1.00002, 1]
    tickets = [srv.submit(1000, L0, q0, replicas=4096, sweeps=10**6, polish_iters=0, init_mode=0, rng_seed=k,
                          timeout_s=15.0) for k in range(70)]
    out = _outcome(sis, srv, tickets, timeout=120)
    assert all(kind == "ok" and r.status == sis.SIS_ETIMEOUT and r.L_cert <= L0 for kind, r in out.values())
    st = srv.stats()
    assert st["peak_resident_bytes"] > 4 * 2**30 and st["resident_requests"] == 0
    srv.close()
