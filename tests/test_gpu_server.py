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
