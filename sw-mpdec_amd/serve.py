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

Run as a program it serves (or solves sequentially) a prefix of the seeded
2000-request stream and writes one JSON record per request (arrival,
admission, completion, service time, state, SHA-256 of the final state), so
that runs on different boxes can be compared result by result:

    python sw-mpdec_amd/serve.py --mode fifo|sjf|sequential --out FILE [--first K --count M]
    python sw-mpdec_amd/serve.py --compare SERVED.json SEQUENTIAL.json [...]

The solver itself is the C ABI of include/swm.h (swm.Engine / swm.Solver);
this module only schedules.
"""
from __future__ import annotations

import argparse
import collections
import hashlib
import json
import math
import sys
import time
from dataclasses import dataclass, field

import numpy as np

import problems as P
import swm


@dataclass
class Request:
    rid: int
    arrival: float           # seconds after the stream starts
    problem: P.Problem | P.FamilySpec | P.ReferenceSetup  # host data, or data made on the device
    nsteps: int
    admitted: float = math.nan
    done: float = math.nan
    result: np.ndarray | None = None
    state: int = -1
    bytes: int = 0
    digest: str = ""         # SHA-256 of the final state's bytes (keep="hash")
    service: float = math.nan  # sequential: seconds the request's own solve took

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


def digest(U: np.ndarray) -> str:
    """SHA-256 of a state's bytes (a result compared across runs)."""
    return hashlib.sha256(np.ascontiguousarray(U).tobytes()).hexdigest()


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
    last_log = 0.0
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
                if keep_results == "hash":
                    r.digest = digest(eng.fetch(jid))
                elif keep_results:
                    r.result = eng.fetch(jid)
                eng.release(jid)
        if log and t - last_log > 20.0:
            last_log = t
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


def sequential(requests: list[Request], keep_results: bool = True, log=None) -> dict:
    """The baseline: one plain solver per request, in arrival order, each
    started when both it has arrived and the previous one has finished."""
    reqs = sorted(requests, key=lambda r: r.arrival)
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0  # noqa: E731
    lat = []
    busy = 0.0
    last_log = 0.0
    for k, r in enumerate(reqs):
        if log and now() - last_log > 20.0:
            last_log = now()
            log(f"sequential: t={last_log:.1f}s request {k} of {len(reqs)}")
        while now() < r.arrival:
            time.sleep(min(r.arrival - now(), 0.05))
        ts = time.perf_counter()
        s = _solver(r)
        state = swm.JOB_DONE
        try:
            s.run_steps(r.nsteps)
        except swm.SwmError:
            state = swm.JOB_FAILED
        if keep_results == "hash":
            r.digest = digest(s.download())
        elif keep_results:
            r.result = s.download()
        s.close()
        r.service = time.perf_counter() - ts
        busy += r.service
        r.done = now()
        r.state = state
        lat.append(r.latency)
    makespan = now()
    lat = np.array(lat)
    return {"jobs": len(reqs), "makespan_s": makespan, "jobs_per_s": len(reqs) / makespan,
            "busy_s": busy, "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99))}


# ---------------------------------------------------------------------------
# The full served stream, run by parts and compared result by result
# ---------------------------------------------------------------------------
STREAM = dict(count=2000, seed=2024, rate=20.0, nmin=64, nmax=1024, smin=50, smax=500)
BUDGET = 240 * 10**9  # bytes admitted at once: 288 GB of HBM less headroom


def stream(first: int = 0, count: int = 0) -> list[Request]:
    """Requests first .. first+count-1 (count 0: to the end) of the stream.
    Each request's draws follow the previous ones', so a prefix of the
    stream is the stream of that length."""
    total = min(first + count, STREAM["count"]) if count else STREAM["count"]
    reqs = request_mix(total, STREAM["seed"], STREAM["rate"], nmin=STREAM["nmin"],
                       nmax=STREAM["nmax"], smin=STREAM["smin"], smax=STREAM["smax"])
    return reqs[first:]


def record(r: Request) -> dict:
    p = r.problem
    return {"rid": r.rid, "n": p.n, "rows": p.rows, "kind": getattr(p, "kind", 0),
            "steps": r.nsteps, "arrival": r.arrival, "admitted": r.admitted, "done": r.done,
            "service": r.service, "state": r.state, "sha256": r.digest}


def queue_latencies(recs: list[dict]) -> list[float]:
    """Latencies of a sequential solver serving the requests in arrival order,
    from their measured service times: start = max(arrival, previous finish)."""
    t, out = 0.0, []
    for r in sorted(recs, key=lambda r: r["arrival"]):
        t = max(t, r["arrival"]) + r["service"]
        out.append(t - r["arrival"])
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--mode", choices=["fifo", "sjf", "sequential"])
    ap.add_argument("--out")
    ap.add_argument("--first", type=int, default=0)
    ap.add_argument("--count", type=int, default=0, help="requests from --first (0: all)")
    ap.add_argument("--sjf-cells", type=int, default=1 << 22)
    ap.add_argument("--compare", nargs="+", metavar="JSON")
    ap.add_argument("--profile-batch", type=int, metavar="JOBS",
                    help="kernel times of the seeded batched mix of JOBS jobs")
    a = ap.parse_args(argv)
    if a.compare:
        return compare(a.compare)
    if a.profile_batch:
        return profile_batch(a.profile_batch)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    reqs = stream(a.first, a.count)
    arrival = {r.rid: r.arrival for r in reqs}
    if a.mode == "sequential":
        # back to back: each request's service time is what is measured; the
        # queue over the stream's arrivals is computed from them (compare)
        for r in reqs:
            r.arrival = 0.0
    log(f"serve: {len(reqs)} requests from {a.first}, "
        f"{sum(work(r) for r in reqs) / 1e9:.1f} G cell-steps, mode {a.mode}")
    if a.mode == "sequential":
        out = sequential(reqs, keep_results="hash", log=log)
    else:
        out = serve(reqs, BUDGET, keep_results="hash", log=log, policy=a.mode,
                    max_running_cells=a.sjf_cells if a.mode == "sjf" else 0)
    for r in reqs:
        r.arrival = arrival[r.rid]
    out.update(mode=a.mode, first=a.first, stream=STREAM, budget_bytes=BUDGET,
               requests=[record(r) for r in reqs])
    with open(a.out, "w") as f:
        json.dump(out, f)
    log(f"serve: {a.mode} done, {out['jobs_per_s']:.3f} jobs/s, makespan {out['makespan_s']:.1f} s")
    return 0


def profile_batch(jobs: int, seed: int = 2024) -> int:
    """Prints the batched mix's kernel time per category (HIP events around
    every launch), per 10^9 cell-steps, and the mix's rate with and without
    the events."""
    specs = P.mixed_batch(jobs, seed)
    W = sum(p.n * p.rows * p.steps for p in specs)
    _, secs, esteps, prof, _ = run_batch(specs, [p.steps for p in specs], profile=True)
    _, secs_plain, _, _, _ = run_batch(specs, [p.steps for p in specs])
    print(json.dumps({
        "workload": f"{jobs} jobs, seed {seed} (the batched mix)", "cell_steps": W,
        "engine_steps": esteps,
        "rate_with_events": W / secs, "rate_without_events": W / secs_plain,
        "ms_per_1e9_cell_steps": {k: v[0] * 1e9 / W for k, v in prof.items() if v[1]},
        "launches": {k: v[1] for k, v in prof.items() if v[1]}}, indent=1), flush=True)
    return 0


def compare(files: list[str]) -> int:
    """Prints a summary of served runs against sequential parts of the same
    stream: jobs/s, p50/p99 latency, and whether every result's hash agrees."""
    runs = [json.load(open(f)) for f in files]
    seq = {r["rid"]: r for run in runs if run["mode"] == "sequential" for r in run["requests"]}
    summary = {}
    if seq:
        recs = sorted(seq.values(), key=lambda r: r["rid"])
        lat = queue_latencies(recs)
        busy = sum(r["service"] for r in recs)
        last = max(r["arrival"] for r in recs)
        summary["sequential"] = {
            "requests": len(recs), "busy_s": busy,
            "jobs_per_s": len(recs) / busy,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "note": "service times measured per request; latencies by the FIFO queue "
                    "recursion over the stream's arrivals (last arrival "
                    f"{last:.1f} s)"}
    for run in runs:
        if run["mode"] == "sequential":
            continue
        same = [seq[r["rid"]]["sha256"] == r["sha256"] for r in run["requests"] if r["rid"] in seq]
        summary[run["mode"]] = {
            "requests": run["jobs"], "jobs_per_s": run["jobs_per_s"],
            "makespan_s": run["makespan_s"], "p50_latency_s": run["p50_latency_s"],
            "p99_latency_s": run["p99_latency_s"], "peak_running": run["peak_running"],
            "peak_bytes": run["peak_bytes"], "failed": run["failed"],
            "results_compared": len(same), "bitwise_identical_to_sequential": all(same),
            "speedup_vs_sequential": (run["jobs_per_s"] / summary["sequential"]["jobs_per_s"]
                                      if seq else None)}
    print(json.dumps(summary, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
