"""Batched and served mPDeC5 solves on one GPU.

* ``run_batch``: a set of independent solves of uneven sizes in one engine,
  all advanced in lockstep (one shared set of kernel launches per step).
* ``serve``: a continuous-batching request loop.  Requests arrive over time
  (seeded Poisson stream); each is admitted into the running engine as soon
  as it has arrived and its device footprint fits the memory budget, its
  steps run in the same launches as every other running solve, and it leaves
  when done, freeing its memory for the next.  Latency = completion time -
  arrival time.
* ``sequential``: the comparison baseline, one plain solver per request, run
  one after another in arrival order.

The solver itself is the C ABI of include/swm.h (swm.Engine / swm.Solver);
this module only schedules.
"""
from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field

import numpy as np

import problems as P
import swm


@dataclass
class Request:
    rid: int
    arrival: float           # seconds after the stream starts
    problem: P.Problem | P.ReferenceSetup  # host data, or an initial condition made on the device
    nsteps: int
    admitted: float = math.nan
    done: float = math.nan
    result: np.ndarray | None = None
    state: int = -1
    bytes: int = 0

    @property
    def latency(self):
        return self.done - self.arrival


def request_mix(count: int, seed: int, rate: float, nmin: int = 64, nmax: int = 1024,
                smin: int = 50, smax: int = 500) -> list[Request]:
    """Seeded Poisson arrivals at `rate` per second; N log-uniform on
    [nmin, nmax], steps uniform on {smin..smax}; initial states drawn from the
    vortex, lake-at-rest-over-bumps and shallow dam-break families.  A request
    carries only its family's parameters (problems.FamilySpec): its arrays are
    made on the device at admission (spec.build() makes the same ones on the
    host)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(count):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        steps = int(rng.integers(smin, smax + 1))
        fam = int(rng.integers(0, 3))
        s = int(rng.integers(0, 2**31 - 1))
        if fam == 0:
            p = P.vortex_spec(n, s, steps=steps)
        elif fam == 1:
            p = P.lake_bumps_spec(n, s, steps=steps)
        else:
            p = P.dam_break_spec(n, steps=steps, dry=0.1)
        out.append(Request(k, t, p, steps))
    return out


def run_batch(problems, nsteps, budget_bytes: int = 0, profile: bool = False,
              steps_per_call: int = 16):
    """All problems (Problem or FamilySpec) in one engine from the start;
    returns (results, seconds, engine steps, profile).  Each engine call runs
    up to steps_per_call steps with one host read per batch."""
    eng = swm.Engine(budget_bytes)
    ids = [_submit_problem(eng, p, k) for p, k in zip(problems, nsteps)]
    if any(i is None for i in ids):
        raise RuntimeError("batch does not fit the memory budget")
    if profile:
        eng.profile(True)
    t0 = time.perf_counter()
    steps = 0
    while True:
        running, k = eng.run(steps_per_call)
        steps += k
        if not running:
            break
    t1 = time.perf_counter()
    prof = eng.profile_read() if profile else None
    res = [eng.fetch(i) for i in ids]
    states = [eng.info(i)["state"] for i in ids]
    eng.close()
    return res, t1 - t0, steps, prof, states


def reference_mix(count: int, seed: int, rate: float, nmin: int = 64, nmax: int = 1024,
                  smin: int = 50, smax: int = 500) -> list[Request]:
    """Like request_mix, but each request names one of the reference's own
    initial conditions (1-6, the lakes 2 and 3 well-balanced) and carries no
    host data: its initial data are made on the device at admission."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(count):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        steps = int(rng.integers(smin, smax + 1))
        ic = int(rng.integers(1, 7))
        out.append(Request(k, t, P.ReferenceSetup(ic, n, wellbalanced=ic in (2, 3), steps=steps),
                           steps))
    return out


def _submit_problem(eng, p, nsteps):
    if isinstance(p, P.ReferenceSetup):
        return eng.submit_setup(p, nsteps)
    if isinstance(p, P.FamilySpec):
        return eng.submit_family(p, nsteps)
    return eng.submit(p, nsteps)


def _submit(eng, r: Request):
    return _submit_problem(eng, r.problem, r.nsteps)


def solver_for(p):
    """The plain solver holding problem p (Problem, FamilySpec or
    ReferenceSetup), its data made where p says."""
    if isinstance(p, P.ReferenceSetup):
        return swm.Solver.from_setup(p)
    if isinstance(p, P.FamilySpec):
        return swm.Solver.from_family(p)
    return swm.Solver.from_problem(p)


def _solver(r: Request):
    return solver_for(r.problem)


def work(r: Request) -> int:
    """Cell-steps of a request."""
    return r.problem.n * r.problem.rows * r.nsteps


def serve(requests: list[Request], budget_bytes: int, keep_results: bool = True,
          log=None, policy: str = "fifo", max_running_cells: int = 0,
          steps_per_call: int = 4) -> dict:
    """Continuous batching.  Arrived requests are admitted while their device
    footprint fits the budget, stepped together with all running solves, and
    retired as soon as they finish.

    policy "fifo": admission in arrival order (a request that does not fit
    blocks the ones behind it).  policy "sjf": among the arrived requests the
    one with the least work (cells x steps) first, and no admission beyond
    max_running_cells running cells (0: no cap) -- the GPU is saturated well
    below the memory limit, and short requests then finish without waiting
    behind long ones.

    Between admissions the engine runs up to steps_per_call steps of the
    running solves (swm_engine_run: one host read per batch of steps); a
    solve that finishes inside them stops on the device at its own step
    count."""
    if policy not in ("fifo", "sjf"):
        raise ValueError(policy)
    eng = swm.Engine(budget_bytes)
    pending = sorted(requests, key=lambda r: r.arrival)
    arrived: list[Request] = []
    running: dict[int, Request] = {}
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0  # noqa: E731
    steps = 0
    peak_bytes = 0
    peak_running = 0
    cells = 0
    while pending or arrived or running:
        t = now()
        while pending and pending[0].arrival <= t:
            arrived.append(pending.pop(0))
        if policy == "sjf":
            arrived.sort(key=work)
        while arrived:
            r = arrived[0]
            c = r.problem.n * r.problem.rows
            if policy == "sjf" and max_running_cells and running and cells + c > max_running_cells:
                break
            jid = _submit(eng, r)
            if jid is None:
                if not running:
                    raise RuntimeError(f"request {r.rid} alone exceeds the budget")
                break
            arrived.pop(0)
            r.admitted = now()
            r.bytes = eng.info(jid)["bytes"]
            running[jid] = r
            cells += c
        if not running:
            # idle until the next arrival
            time.sleep(max(0.0, min(pending[0].arrival - now(), 0.05)))
            continue
        peak_bytes = max(peak_bytes, eng.bytes_in_use())
        peak_running = max(peak_running, len(running))
        steps += eng.run(steps_per_call)[1]
        t = now()
        for jid in list(running):
            info = eng.info(jid)
            if info["state"] != swm.JOB_RUNNING:
                r = running.pop(jid)
                cells -= r.problem.n * r.problem.rows
                r.done = t
                r.state = info["state"]
                if keep_results:
                    r.result = eng.fetch(jid)
                eng.release(jid)
        if log and steps // 50 != (steps - 1) // 50:
            log(f"serve: t={t:.2f}s step {steps} running {len(running)} "
                f"waiting {len(pending) + len(arrived)}")
    makespan = now()
    eng.close()
    lat = np.array([r.latency for r in requests])
    return {"jobs": len(requests), "makespan_s": makespan, "jobs_per_s": len(requests) / makespan,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "engine_steps": steps, "peak_running": peak_running, "peak_bytes": peak_bytes,
            "failed": int(sum(r.state == swm.JOB_FAILED for r in requests))}


def sequential(requests: list[Request], keep_results: bool = True) -> dict:
    """The baseline: one plain solver per request, in arrival order, each
    started when both it has arrived and the previous one has finished."""
    reqs = sorted(requests, key=lambda r: r.arrival)
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0  # noqa: E731
    lat = []
    busy = 0.0
    for r in reqs:
        while now() < r.arrival:
            time.sleep(min(r.arrival - now(), 0.05))
        ts = time.perf_counter()
        s = _solver(r)
        state = swm.JOB_DONE
        try:
            s.run_steps(r.nsteps)
        except swm.SwmError:
            state = swm.JOB_FAILED
        if keep_results:
            r.result = s.download()
        s.close()
        busy += time.perf_counter() - ts
        r.done = now()
        r.state = state
        lat.append(r.latency)
    makespan = now()
    lat = np.array(lat)
    return {"jobs": len(reqs), "makespan_s": makespan, "jobs_per_s": len(reqs) / makespan,
            "busy_s": busy, "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99))}
