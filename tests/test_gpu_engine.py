"""GPU tests of the batched engine: every job of a mixed batch reproduces the
oracle bitwise, results are the same alone and in a mix, jobs can join and
leave between steps, and admission respects the memory budget."""
from __future__ import annotations

import numpy as np
import pytest

import problems as P
import swm
from oracle_api import Oracle
from parity_util import describe_diff, same_bits

pytestmark = pytest.mark.gpu


def oracle_run(p, steps):
    o = Oracle(p)
    for _ in range(steps):
        dt, _, _ = o.timestep()
        o.step(dt)
    return o.get_U()


def mix():
    return [(P.vortex(20, 1), 3), (P.lake_bumps(33, 2), 2), (P.vortex(47, 3), 2),
            (P.dam_break(26, dry=0.1), 4), (P.lake_flag1(18, True), 3), (P.vortex(16, 4), 1)]


def run_all(eng):
    n = 1
    while n:
        n = eng.step()


def test_batch_matches_oracle_per_job():
    jobs = mix()
    eng = swm.Engine()
    ids = [eng.submit(p, k) for p, k in jobs]
    run_all(eng)
    for (p, k), jid in zip(jobs, ids):
        info = eng.info(jid)
        assert info["state"] == swm.JOB_DONE and info["steps_done"] == k
        U = eng.fetch(jid)
        U_r = oracle_run(p, k)
        assert same_bits(U, U_r), f"{p.kind} n={p.n}: {describe_diff(U, U_r)}"


def test_alone_and_in_mix_are_identical():
    target = P.lake_bumps(40, 9)
    eng1 = swm.Engine()
    a = eng1.submit(target, 3)
    run_all(eng1)
    alone = eng1.fetch(a)
    eng2 = swm.Engine()
    others = [eng2.submit(p, k) for p, k in mix()]
    b = eng2.submit(target, 3)
    run_all(eng2)
    assert same_bits(eng2.fetch(b), alone)
    assert all(eng2.info(j)["state"] == swm.JOB_DONE for j in others)


def test_continuous_admission_and_release():
    eng = swm.Engine()
    first = [(eng.submit(p, k), p, k) for p, k in mix()[:3]]
    eng.step()
    late = [(eng.submit(p, k), p, k) for p, k in mix()[3:]]
    done = {}
    while True:
        n = eng.step()
        for jid, p, k in first + late:
            if jid not in done and eng.info(jid)["state"] == swm.JOB_DONE:
                done[jid] = eng.fetch(jid)
                eng.release(jid)
        if n == 0:
            break
    assert len(done) == len(first) + len(late)
    for jid, p, k in first + late:
        assert same_bits(done[jid], oracle_run(p, k)), (p.kind, p.n)
    assert eng.bytes_in_use() == 0


def test_admission_respects_budget():
    p = P.vortex(64, 5)
    fp = swm.job_footprint_bytes(64, 64, False, False)
    eng = swm.Engine(budget_bytes=int(1.5 * fp))
    a = eng.submit(p, 1)
    assert a is not None
    assert eng.submit(p, 1) is None  # does not fit while the first holds memory
    eng.step()
    eng.release(a)
    b = eng.submit(p, 1)
    assert b is not None
    eng.step()
    assert same_bits(eng.fetch(b), oracle_run(p, 1))


def test_failed_job_does_not_disturb_others():
    eng = swm.Engine()
    dry = eng.submit(P.dam_break(24), 5)  # dt becomes NaN before step 2
    good = P.vortex(24, 8)
    g = eng.submit(good, 3)
    run_all(eng)
    assert eng.info(dry)["state"] == swm.JOB_FAILED
    assert eng.info(g)["state"] == swm.JOB_DONE
    assert same_bits(eng.fetch(g), oracle_run(good, 3))


def test_serve_policies_match_sequential():
    """The continuous-batching loop (FIFO and shortest-work-first with a cap
    on running cells) returns, request by request, the plain solver's
    results bitwise."""
    import serve

    reqs = serve.request_mix(10, seed=3, rate=200.0, nmin=16, nmax=48, smin=2, smax=5)
    serve.sequential(reqs, keep_results=True)
    ref = [r.result.copy() for r in reqs]
    for policy, cap in (("fifo", 0), ("sjf", 3000)):
        for r in reqs:
            r.result, r.done, r.admitted = None, float("nan"), float("nan")
        out = serve.serve(reqs, budget_bytes=0, keep_results=True, policy=policy,
                          max_running_cells=cap)
        assert out["jobs"] == len(reqs) and out["failed"] == 0
        for r, U in zip(reqs, ref):
            assert same_bits(r.result, U), (policy, r.rid, r.problem.kind, r.problem.n)


def test_zero_step_job_is_done_and_keeps_its_state():
    p = P.vortex(20, 2)
    eng = swm.Engine()
    a = eng.submit(p, 0)
    b = eng.submit(P.vortex(24, 3), 2)
    assert eng.info(a)["state"] == swm.JOB_DONE
    while eng.step():
        pass
    assert eng.info(a)["steps_done"] == 0
    assert same_bits(eng.fetch(a), p.U0)
    assert eng.info(b)["state"] == swm.JOB_DONE and eng.info(b)["steps_done"] == 2
