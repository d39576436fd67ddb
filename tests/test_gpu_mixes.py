"""The batched and served configurations at their own shapes.

* Batched: 256 jobs, N ~ logU[32, 512], 20..200 steps, half vortices, half
  well-balanced lakes over bumps (problems.mixed_batch(256, 2024), the bench
  workload), all in one engine: every job's final state is bitwise the
  sequential plain solver's, and the 8 jobs of least work are bitwise the
  oracle's over their full step counts.
* Served: the first requests of the 2000-request stream (N ~ logU[64, 1024],
  50..500 steps, vortex / lake over bumps / shallow dam break, seeded Poisson
  arrivals) through the continuous-batching loop: each result is bitwise the
  same alone (sequential plain solver) and in the mix.
Every job's data are made on the device from its family parameters.
"""
from __future__ import annotations

import numpy as np
import pytest

import problems as P
import serve
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


def test_batched_256_jobs_match_sequential_and_oracle():
    specs = P.mixed_batch(256, 2024)
    print(f"batched: 256 jobs, {sum(p.n * p.rows * p.steps for p in specs) / 1e9:.2f} G "
          "cell-steps", flush=True)
    res, secs, esteps, _, states = serve.run_batch(specs, [p.steps for p in specs])
    assert all(st == swm.JOB_DONE for st in states)
    assert esteps == max(p.steps for p in specs)
    for k, (p, U) in enumerate(zip(specs, res)):
        s = serve.solver_for(p)
        s.run_steps(p.steps)
        assert same_bits(s.download(), U), (k, p.kind, p.n, p.steps)
        s.close()
    small = sorted(range(len(specs)), key=lambda k: specs[k].n * specs[k].rows * specs[k].steps)
    for k in small[:8]:
        p = specs[k]
        U_r = oracle_run(p.build(), p.steps)
        assert same_bits(res[k], U_r), (k, p.kind, p.n, describe_diff(res[k], U_r))


def test_served_stream_prefix_alone_and_in_the_mix():
    reqs = serve.request_mix(2000, 2024, 20.0, nmin=64, nmax=1024, smin=50, smax=500)[:12]
    print(f"served: {len(reqs)} requests, N {[r.problem.n for r in reqs]}", flush=True)
    out = serve.serve(reqs, budget_bytes=0, keep_results=True)
    assert out["jobs"] == len(reqs) and out["failed"] == 0
    mixed = [r.result.copy() for r in reqs]
    for r in reqs:
        r.result, r.done, r.admitted = None, float("nan"), float("nan")
    serve.sequential(reqs, keep_results=True)
    for r, U in zip(reqs, mixed):
        assert np.isfinite(U).all()
        assert same_bits(r.result, U), (r.rid, r.problem.kind, r.problem.n, r.nsteps)
