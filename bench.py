#!/usr/bin/env python3
"""Benchmark of the mPDeC5 / WENO5 PATANKAR time step on one MI355X.

Workload (the configuration the headline metric is quoted on): one 2048x2048
periodic grid, FP64, PATANKAR + WELLBALANCED, lake at rest over 8 seeded
Gaussian bumps plus a seeded 1e-3 Gaussian surface perturbation, CFL 0.5.
A step is one full time step: the CFL reduction (TimeStep) and one mPDeC5
step (13 FVTimeDerivative evaluations, 13 Patankar assemblies and Jacobi
solves), with the state resident in HBM.

    python bench.py [--gpus 1] [--steps 100] [--warmup 2] [--size 2048]

Prints one JSON line.  value = cell-steps/s = size^2 * steps / seconds.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "sw-mpdec_amd"))

# untimed steps in which the engine's autotuner times its candidates (tile
# shape, graph replay, speculative margin), before the warmup steps
TUNING_STEPS = 20
# the dominant kernels, timed by HIP events in the timed region
FACE_CATEGORIES = ("fvtd_x", "fvtd_y")

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# Algorithmic bytes per unit of work of each kernel (8-byte words the kernel
# must move at least once), and the units of one step (cells = nx*ny):
#   fvtd_x / fvtd_y   per face of one node: read h, hu, hv (24), write F1..F3 (24)
#   fvtd_update       per cell of one node: read one x and one y face flux (48),
#                     write Ut(2:3) (16) and the east/north mass flux (16)
#   patankar_assemble per row-system: face mass fluxes of 4 nodes (64), h (8),
#                     write 4 off-diagonals, the diagonal and x1 (48); and the
#                     DeC momentum update of the same node, which runs in the
#                     same kernel: read Ua(1,2:3) and FUp(1:4,2:3) (80), write 16
#   jacobi_sweep      per row-sweep: 5 matrix planes, x_old, b (56), write x_new (8)
#   timestep          per cell: read h, hu, hv (24)
# The 13 FVTDs / momentum updates / Patankar systems of a step cover 13
# node-evaluations of every cell; Jacobi covers J_step sweeps of every row.
ALGO_BYTES = {"fvtd_x": 48.0, "fvtd_y": 48.0, "fvtd_update": 80.0,
              "patankar_assemble": 120.0 + 96.0, "jacobi_sweep": 64.0, "timestep": 24.0}
# FP64 operations per face of one node in fvtd_x / fvtd_y, counted from the
# source (each +, -, *, /, sqrt = 1): ConsToPrim 4, three first sweeps 348,
# six second sweeps 1320, limiter 70, Rusanov at 4 Gauss points 264.
FP64_OPS_PER_FACE = 2006.0


def units_per_step(kernel: str, cells: int, sweeps_per_step: float) -> float:
    if kernel == "jacobi_sweep":
        return cells * sweeps_per_step
    if kernel == "timestep":
        return float(cells)
    return 13.0 * cells


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_per_cell_node(kernel: str, size: int):
    """Measured HBM bytes per cell-node of an FVTD kernel (rocprofv3 --pmc
    FETCH_SIZE + WRITE_SIZE, scaled by a copy-kernel calibration), read from
    the committed summary profiles/traffic_<size>.md."""
    f = ROOT / "profiles" / f"traffic_{size}.md"
    if not f.exists():
        return None
    m = re.search(rf"traffic_bytes_per_cell_node\[{kernel}\]:\s*([0-9.eE+]+)", f.read_text())
    return float(m.group(1)) if m else None


def profiled(key: str, kernel: str, size: int):
    """A per-kernel figure (key[kernel]: value) of profiles/traffic_<size>.md."""
    f = ROOT / "profiles" / f"traffic_{size}.md"
    if not f.exists():
        return None
    m = re.search(rf"{key}\[{kernel}\]:\s*([0-9.eE+]+)", f.read_text())
    return float(m.group(1)) if m else None


def cpu_baseline(budget_s: float = 20.0, scheme: int = 5, recfix: int = 0):
    """The reference Fortran itself (oracle/_ref, built from its own sources)
    on the host: one core, the reference's own lake-at-rest bathymetry
    (BathymetryFlag=1) with the well-balanced build, 128x128, the number of
    CFL steps that fits the time budget.  Release-style build if present.
    scheme: the timescheme (1-4 run the well-balanced build without
    PATANKAR, which has no release-style variant here).  recfix: the
    ReconstructionFix the reference runs with (0: its own, 4)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import problems as P
    from oracle_api import big_stack

    if scheme >= 5:
        exe = ROOT / "oracle" / "_ref" / "swref_pwb_fast"
        kind_note = "amdflang -O3 -march=x86-64-v3"
        if not exe.exists():
            exe = ROOT / "oracle" / "_ref" / "swref_pwb"
            kind_note = "amdflang -O2 -ffp-contract=off"
    else:
        exe = ROOT / "oracle" / "_ref" / "swref_ewb"
        kind_note = "amdflang -O2 -ffp-contract=off, built without PATANKAR"
    if not exe.exists():
        return None
    n = 128
    p = P.lake_flag1(n, True)

    def run(steps):
        hdr = np.array([5, n, steps, 1, 1, 0, scheme, recfix], dtype="<i4")
        fl = np.array([p.x0, p.x1, p.gravity, p.cfl], dtype="<f8")
        with tempfile.TemporaryDirectory() as td:
            fin, fout = Path(td) / "in.bin", Path(td) / "out.bin"
            with open(fin, "wb") as f:
                f.write(hdr.tobytes())
                f.write(fl.tobytes())
                f.write(np.ascontiguousarray(p.U0, dtype="<f8").tobytes())
                f.write(np.ascontiguousarray(p.Uwb, dtype="<f8").tobytes())
            cmd = [str(exe), str(fin), str(fout)]
            if os.path.exists("/usr/bin/taskset"):
                cmd = ["taskset", "-c", "1"] + cmd
            subprocess.run(cmd, check=True, cwd=td, stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL, env=dict(os.environ, OMP_NUM_THREADS="1"),
                           preexec_fn=big_stack, timeout=600)
            raw = np.fromfile(fout, dtype="<f8")
        return float(raw[0]), int(raw[1])

    t1, k1 = run(1)
    steps = max(1, min(50, int(budget_s / max(t1, 1e-3))))
    t, k = run(steps)
    rate = n * n * k / t
    out = {"value": rate, "unit": "cell-steps/s", "cores": 1, "kind": "reference",
           "cpu_model": cpu_model(),
           "sample": (f"reference Fortran ({kind_note}), timescheme {scheme}, "
                      f"{f'ReconstructionFix {recfix}, ' if recfix else ''}WELLBALANCED, "
                      f"BathymetryFlag=1 lake at rest + 1e-3 bump, {n}x{n}, {k} CFL steps, "
                      f"{t:.1f} s, time loop only")}
    if scheme >= 5:
        # the run's Jacobi sweeps: the same steps on the GPU, whose sweep
        # counts are the reference's solve by solve
        import swm
        g = swm.Solver.from_problem(p)
        g.set_timescheme(scheme)
        _, its, _ = g.run_steps(k)
        g.close()
        out["jacobi_sweeps_per_step"] = float(np.sum(its)) / k
    if scheme == 5 and not recfix:
        out["port_on_headline_family"] = port_baseline(budget_s=min(budget_s, 15.0))
    return out


PORT_RUN = """
import json, sys, time
sys.path[:0] = [sys.argv[1] + '/sw-mpdec_amd', sys.argv[1] + '/tests']
import problems as P
from oracle_api import Oracle
n, budget = int(sys.argv[2]), float(sys.argv[3])
o = Oracle(P.lake_bumps_spec(n, 2024).build())
t0, k, sweeps = time.perf_counter(), 0, 0
while k == 0 or time.perf_counter() - t0 < budget:
    dt, _, _ = o.timestep()
    its, _ = o.step(dt)
    k, sweeps = k + 1, sweeps + sum(its)
el = time.perf_counter() - t0
print(json.dumps({"steps": k, "seconds": el, "sweeps": sweeps}))
"""


def port_baseline(budget_s: float = 15.0, n: int = 256):
    """The C restatement of the reference (oracle/, FP contraction off, one
    core) on the headline's own family -- the lake over 8 Gaussian bumps,
    seed 2024 -- at n x n, which the reference's Fortran cannot take (its
    bathymetry is built in): cell-steps/s over as many steps as fit the
    budget."""
    cmd = [sys.executable, "-c", PORT_RUN, str(ROOT), str(n), str(budget_s)]
    if os.path.exists("/usr/bin/taskset"):
        cmd = ["taskset", "-c", "1"] + cmd
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
    if r.returncode != 0:  # the oracle library not built here: no port figure
        log(f"bench: port baseline skipped ({r.stderr.strip().splitlines()[-1:]})")
        return None
    res = json.loads(r.stdout.strip().splitlines()[-1])
    return {"value": n * n * res["steps"] / res["seconds"], "unit": "cell-steps/s", "cores": 1,
            "kind": "port",
            "sample": f"the C oracle (gcc, -ffp-contract=off) on lake_bumps {n}x{n} seed 2024, "
                      f"{res['steps']} CFL steps, {res['seconds']:.1f} s",
            "jacobi_sweeps_per_step": res["sweeps"] / res["steps"]}


def cpu_model() -> str:
    """The host CPU's model name (with its family, model and stepping numbers,
    which identify the core generation where a virtual machine shortens the
    name) and the cores this process may use."""
    info = {}
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            key, _, val = line.partition(":")
            key = key.strip()
            if key in ("model name", "cpu family", "model", "stepping", "cpu MHz") \
                    and key not in info:
                info[key] = val.strip()
            if not line.strip() and info:
                break  # the first processor's block
    except OSError:
        pass
    name = info.get("model name", "unknown")
    ids = ", ".join(f"{k} {info[k]}" for k in ("cpu family", "model", "stepping", "cpu MHz")
                    if k in info)
    if ids:
        name = f"{name} ({ids})"
    return f"{name}; {len(os.sched_getaffinity(0))} cores available, 1 used (taskset -c 1)"


def schemes_section(n: int, seed: int, steps: int, cpu: bool) -> dict:
    """The other time schemes on the headline workload, `steps` CFL steps each
    after one warmup step, through swm_run_steps: the explicit ones of the
    build without PATANKAR (timescheme 1-4) and mPEuler / mPDeC2 (6, 7) with
    the by-node depth view (the reference's own view turns the state NaN
    within a few steps, so it has no meaningful rate)."""
    import problems as P
    import swm

    spec = P.lake_bumps_spec(n, seed)
    res = {}
    for ts, name in ((1, "forward_euler"), (2, "ssprk4"), (3, "rk65"), (4, "dec5"),
                     (6, "mpeuler"), (7, "mpdec2")):
        s = swm.Solver.from_family(spec)
        s.set_timescheme(ts)
        s.set_hp_association(swm.HP_BY_NODE)
        s.run_steps(1)
        s.sync()
        t0 = time.perf_counter()
        _, iters, _ = s.run_steps(steps)
        s.sync()
        el = time.perf_counter() - t0
        res[name] = {"timescheme": ts, "steps": steps,
                     "cell_steps_per_s": n * n * steps / el, "ms_per_step": 1e3 * el / steps,
                     "jacobi_sweeps_per_step": float(np.sum(iters)) / steps}
        if ts >= 6:
            res[name]["hp_association"] = "by_node"
        log(f"bench: {name} {steps} steps {1e3 * el / steps:.2f} ms/step")
        del s
        if cpu and ts <= 4:
            # the reference's own rate for the scheme (its mPEuler / mPDeC2
            # states turn NaN, so they have none to compare)
            cb = cpu_baseline(budget_s=4.0, scheme=ts)
            if cb:
                res[name]["cpu_reference"] = cb
                res[name]["gpu_vs_cpu"] = res[name]["cell_steps_per_s"] / cb["value"]
    return res


def flagged_fraction(h: np.ndarray) -> float:
    """Fraction of the cells of the periodic depths h (rows, cols) that the
    Jameson indicator flags along x (src/shocksindicator.f90:93-129)."""
    hm, hp = np.roll(h, 1, 1), np.roll(h, -1, 1)
    ind = np.abs(hm + hp - 2.0 * h)
    with np.errstate(invalid="ignore", divide="ignore"):
        ind = ind / (np.abs(hm) + np.abs(hp) + np.abs(2.0 * h))
    return float(np.mean(ind > 0.5e-3))


def time_loop_section(steps: int, sizes=(64, 128, 256, 512)) -> dict:
    """The device time loop on small grids, where host round trips weigh:
    ms per mPDeC5 step of swm_run_steps called once per step (one host read
    of the device status per step) against one call for all `steps` steps
    (one host read per batch of up to 16), both after the autotuner's steps,
    on a seeded vortex (data made on the device)."""
    import problems as P
    import swm

    res = {"workload": f"vortex N x N, mPDeC5, {steps} steps each way"}
    for n in sizes:
        s = swm.Solver.from_family(P.vortex_spec(n, 7))
        s.run_steps(TUNING_STEPS)
        s.sync()
        r0 = s.status_reads()
        t0 = time.perf_counter()
        for _ in range(steps):
            s.run_steps(1)
        s.sync()
        one = (time.perf_counter() - t0) / steps
        r1 = s.status_reads()
        t0 = time.perf_counter()
        s.run_steps(steps)
        s.sync()
        batched = (time.perf_counter() - t0) / steps
        r2 = s.status_reads()
        res[str(n)] = {"ms_per_step_one_read_per_step": 1e3 * one,
                       "ms_per_step_device_loop": 1e3 * batched,
                       "status_reads": [r1 - r0, r2 - r1],
                       "speedup": one / batched}
        s.close()
    return res


def sync_cost_section() -> dict:
    """What a step boundary costs (swm_debug_sync_cost): microseconds per
    round of a chain through every block, with one kernel launch per round,
    the launches replayed from a HIP graph, or a grid barrier inside one
    cooperative launch -- the alternative a persistent kernel would use."""
    import swm

    res = {}
    for nblk in (256, 1024, 2048):
        launch, graph, barrier = swm.sync_cost(nblk, 2000)
        res[str(nblk)] = {"launch_us": launch, "graph_us": graph, "grid_barrier_us": barrier}
    return res


def recfix_section(n: int, steps: int, cpu: bool) -> dict:
    """ReconstructionFix 1 (first order) and 2 (MUSCL) on the shock indicator's
    flagged cells, against 4 (WENO5 everywhere, the reference's setting), on a
    shallow dam break of n x n cells (h = 1 inside radius 0.25 of the centre,
    0.1 outside: the indicator flags the front), mPDeC5, `steps` CFL steps
    each after one warmup step.  The face kernels compute WENO5 for every cell
    and replace the flagged cells' values, so the fix adds work only where
    cells are flagged."""
    import problems as P
    import swm

    prob = P.dam_break(n, dry=0.1)
    res = {"workload": f"shallow dam break {n}x{n}, mPDeC5",
           "flagged_fraction_x_initial": flagged_fraction(prob.U0[..., 0])}
    for rf in (4, 1, 2):
        s = swm.Solver.from_problem(prob, recfix=rf)
        s.run_steps(1)
        s.sync()
        t0 = time.perf_counter()
        _, iters, _ = s.run_steps(steps)
        s.sync()
        el = time.perf_counter() - t0
        res[f"fix{rf}"] = {"steps": steps, "ms_per_step": 1e3 * el / steps,
                           "cell_steps_per_s": n * n * steps / el,
                           "jacobi_sweeps_per_step": float(np.sum(iters)) / steps}
        log(f"bench: reconstruction fix {rf}: {1e3 * el / steps:.2f} ms/step")
        del s
    if cpu:
        cb = cpu_baseline(budget_s=4.0, scheme=5, recfix=2)
        if cb:
            res["fix2"]["cpu_reference"] = cb
            res["fix2"]["gpu_vs_cpu"] = res["fix2"]["cell_steps_per_s"] / cb["value"]
    return res


def initial_data_section(n: int, cpu: bool) -> dict:
    """The reference's initial data made on the device (swm_fill_initial_data)
    for n x n cells of each initial condition's own domain: the fill of the
    state, the bathymetry tables and, for the well-balanced lakes 2 and 3,
    the well-balanced state and its FVTD; and PostProcessing's L1 error
    (swm_l1_error, the sum in the reference's sequential order included).  CPU: the reference's own
    FillInitialConditions + InitializeWBVariables (its WELLBALANCED build,
    driver mode 6) for the lake at rest at 256 x 256, wall time of the
    process."""
    import problems as P
    import swm

    res = {"cells": n * n}
    for ic in (1, 2, 3, 4):
        c = P.ReferenceSetup(ic, n, wellbalanced=ic in (2, 3))
        s = swm.Solver(c.nx, c.ny, c.mesh.dx, c.mesh_y.dx, c.gravity, c.cfl, c.wellbalanced,
                       c.mesh.wbnd)
        s.fill_initial_data(c.ic, c.bathy_flag, c.x0, c.x1)  # first call: allocations
        s.sync()
        t0 = time.perf_counter()
        s.fill_initial_data(c.ic, c.bathy_flag, c.x0, c.x1)
        s.sync()
        el = time.perf_counter() - t0
        res[f"ic{ic}"] = {"ms": 1e3 * el, "cells_per_s": n * n / el,
                          "wellbalanced": bool(c.wellbalanced)}
        log(f"bench: initial data ic {ic} {n}x{n}: {1e3 * el:.2f} ms")
        if ic == 1:  # PostProcessing's L1 error of this state at t = 0.1
            s.l1_error(1, 0, c.x0, c.x1, 0.1)
            t0 = time.perf_counter()
            s.l1_error(1, 0, c.x0, c.x1, 0.1)
            res["l1_error_ms"] = 1e3 * (time.perf_counter() - t0)
            log(f"bench: L1 error {n}x{n}: {res['l1_error_ms']:.2f} ms")
        del s
    exe = ROOT / "oracle" / "_ref" / "swref_pwb_fast"
    if cpu and exe.exists():
        sys.path.insert(0, str(ROOT / "tests"))
        from oracle_api import big_stack

        m = 256
        x0, x1, g, bf = P.REFERENCE_SETUPS[2]
        hdr = np.array([6, m, 2, bf, 1, m, 5, 4], dtype="<i4")
        fl = np.array([x0[0], x1[0], g, 0.5, x0[1], x1[1]], dtype="<f8")
        with tempfile.TemporaryDirectory() as td:
            fin, fout = Path(td) / "in.bin", Path(td) / "out.bin"
            with open(fin, "wb") as f:
                f.write(hdr.tobytes())
                f.write(fl.tobytes())
            t0 = time.perf_counter()
            subprocess.run([str(exe), str(fin), str(fout)], check=True, cwd=td,
                           stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                           preexec_fn=big_stack, timeout=600)
            el = time.perf_counter() - t0
        res["cpu_reference_ic2"] = {"cells": m * m, "s": el, "cells_per_s": m * m / el,
                                    "cores": 1, "kind": "reference",
                                    "sample": "amdflang -O3 build, WELLBALANCED, "
                                              "process wall time incl. start-up"}
        res["ic2_gpu_vs_cpu"] = res["ic2"]["cells_per_s"] / res["cpu_reference_ic2"]["cells_per_s"]
    return res


def served_ic_section(count: int, seed: int, rate: float) -> dict:
    """Continuous batching of `count` requests that name one of the
    reference's initial conditions (1-6) and carry no host data: each job's
    initial data are made on the device at admission (swm_fill_initial_data),
    against the sequential plain solver doing the same per request.  N ~
    logU[64, 512], 20..200 CFL steps, mPDeC5 (FIFO admission)."""
    import serve

    reqs = serve.reference_mix(count, seed, rate, nmin=64, nmax=512, smin=20, smax=200)
    work = sum(r.problem.nx * r.problem.ny * r.nsteps for r in reqs)
    sv = serve.serve(reqs, budget_bytes=0, keep_results=True, log=log)
    served = [r.result for r in reqs]
    for r in reqs:
        r.result, r.done, r.admitted = None, float("nan"), float("nan")
    sq = serve.sequential(reqs, keep_results=True)
    same = all(np.array_equal(a.view(np.uint64), r.result.view(np.uint64))
               for a, r in zip(served, reqs))
    log(f"bench: served by initial condition: {sv['jobs_per_s']:.2f} jobs/s, "
        f"sequential {sq['jobs_per_s']:.2f} jobs/s, identical={same}")
    return {"workload": f"{count} requests by InitialCondition 1-6, Poisson {rate}/s, "
                        f"N~logU[64,512], steps~U{{20..200}}, seed {seed}",
            "cell_steps": work, "jobs_per_s": sv["jobs_per_s"],
            "p50_latency_s": sv["p50_latency_s"], "p99_latency_s": sv["p99_latency_s"],
            "failed": sv["failed"], "sequential_jobs_per_s": sq["jobs_per_s"],
            "speedup_vs_sequential": sv["jobs_per_s"] / sq["jobs_per_s"],
            "bitwise_identical_to_sequential": bool(same)}


def batched_section(count: int, seed: int) -> dict:
    """`count` uneven solves (N = round(exp(U[ln 32, ln 512])), 20..200 CFL
    steps, half vortices, half well-balanced lakes over bumps, their data made
    on the device) run together in one engine, against the same solves run
    one after another with the plain solver.  Both timings exclude the
    initial data.  Every job's final state must be bitwise the same both
    ways."""
    import problems as P
    import serve
    import swm

    t0 = time.time()
    probs = P.mixed_batch(count, seed)
    work = sum(p.n * p.n * p.steps for p in probs)
    log(f"bench: batch of {count} problems built in {time.time() - t0:.1f} s "
        f"({work / 1e9:.2f} G cell-steps)")
    res, secs, esteps, _, states = serve.run_batch(probs, [p.steps for p in probs])
    log(f"bench: batched run {secs:.2f} s, {esteps} engine steps")
    seq = 0.0
    same = True
    for p, r in zip(probs, res):
        sv = serve.solver_for(p)
        sv.sync()
        ts = time.perf_counter()
        sv.run_steps(p.steps)
        sv.sync()
        seq += time.perf_counter() - ts
        same = same and bool(np.array_equal(sv.download().view(np.uint64), r.view(np.uint64)))
        sv.close()
    log(f"bench: sequential plain solver {seq:.2f} s, identical={same}")
    return {"workload": f"{count} jobs, N~logU[32,512], steps~U{{20..200}}, 50% vortex "
                        f"(PATANKAR) / 50% WB lake over bumps, seed {seed}",
            "jobs": count, "cell_steps": work,
            "jobs_per_s": count / secs, "cell_steps_per_s": work / secs, "seconds": secs,
            "sequential_jobs_per_s": count / seq, "sequential_seconds": seq,
            "speedup_vs_sequential": seq / secs, "engine_steps": esteps,
            "failed": int(sum(st == swm.JOB_FAILED for st in states)),
            "bitwise_identical_to_sequential": same}


def served_section(count: int, seed: int, rate: float) -> dict:
    """Continuous batching of `count` requests arriving as a seeded Poisson
    stream at `rate` per second (N ~ logU[64, 1024], 50..500 CFL steps,
    vortex / lake over bumps / shallow dam break), against the sequential
    plain solver on the same stream.  Latency = completion - arrival."""
    import serve

    t0 = time.time()
    reqs = serve.request_mix(count, seed, rate, nmin=64, nmax=1024, smin=50, smax=500)
    work = sum(r.problem.n ** 2 * r.nsteps for r in reqs)
    log(f"bench: {count} requests built in {time.time() - t0:.1f} s "
        f"({work / 1e9:.2f} G cell-steps)")
    sv = serve.serve(reqs, budget_bytes=0, keep_results=True, log=log)
    served = [r.result for r in reqs]
    log(f"bench: served (FIFO) makespan {sv['makespan_s']:.2f} s, p50 {sv['p50_latency_s']:.2f} s, "
        f"p99 {sv['p99_latency_s']:.2f} s")
    for r in reqs:
        r.result, r.done, r.admitted = None, float("nan"), float("nan")
    cap = 1 << 22
    sj = serve.serve(reqs, budget_bytes=0, keep_results=True, log=log, policy="sjf",
                     max_running_cells=cap)
    same_sjf = all(np.array_equal(a.view(np.uint64), r.result.view(np.uint64))
                   for a, r in zip(served, reqs))
    log(f"bench: served (SJF) makespan {sj['makespan_s']:.2f} s, p50 {sj['p50_latency_s']:.2f} s, "
        f"p99 {sj['p99_latency_s']:.2f} s")
    for r in reqs:
        r.result, r.done, r.admitted = None, float("nan"), float("nan")
    sq = serve.sequential(reqs, keep_results=True)
    same = all(np.array_equal(a.view(np.uint64), r.result.view(np.uint64))
               for a, r in zip(served, reqs))
    log(f"bench: sequential makespan {sq['makespan_s']:.2f} s, p50 {sq['p50_latency_s']:.2f} s")
    return {"workload": f"{count} requests, Poisson {rate}/s, N~logU[64,1024], "
                        f"steps~U{{50..500}}, seed {seed}",
            "requests": count, "cell_steps": work, "arrival_rate_per_s": rate,
            "jobs_per_s": sv["jobs_per_s"], "p50_latency_s": sv["p50_latency_s"],
            "p99_latency_s": sv["p99_latency_s"], "makespan_s": sv["makespan_s"],
            "peak_running": sv["peak_running"], "peak_bytes": sv["peak_bytes"],
            "failed": sv["failed"],
            "sequential_jobs_per_s": sq["jobs_per_s"],
            "sequential_p50_latency_s": sq["p50_latency_s"],
            "sequential_p99_latency_s": sq["p99_latency_s"],
            "speedup_vs_sequential": sv["jobs_per_s"] / sq["jobs_per_s"],
            "bitwise_identical_to_sequential": same and same_sjf,
            "sjf": {"policy": f"shortest work first, at most {cap} cells running",
                    "jobs_per_s": sj["jobs_per_s"], "p50_latency_s": sj["p50_latency_s"],
                    "p99_latency_s": sj["p99_latency_s"], "makespan_s": sj["makespan_s"],
                    "speedup_vs_sequential": sj["jobs_per_s"] / sq["jobs_per_s"]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--scheme-steps", type=int, default=20,
                    help="steps of each other time scheme at the headline size (0: skip)")
    ap.add_argument("--initial-data", type=int, default=1,
                    help="1: time the initial data made on the device at the headline size")
    ap.add_argument("--sync-cost", type=int, default=1,
                    help="1: measure a step boundary's cost (launch, graph, grid barrier)")
    ap.add_argument("--time-loop-steps", type=int, default=48,
                    help="steps per grid of the device time loop section (0: skip)")
    ap.add_argument("--fix-steps", type=int, default=10,
                    help="steps of each ReconstructionFix run on a dam break at the headline "
                         "size (0: skip)")
    ap.add_argument("--batch-jobs", type=int, default=256,
                    help="jobs of the batched run (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=48,
                    help="requests of the served run (0: skip)")
    ap.add_argument("--serve-ic-requests", type=int, default=24,
                    help="requests of the served run by initial condition (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=20.0,
                    help="Poisson arrival rate of the served run, per second")
    ap.add_argument("--serve-rates", type=str, default="",
                    help="comma-separated further arrival rates, each a served run of its own "
                         "(a rate sweep, reported under served_sweep)")
    ap.add_argument("--graph", type=int, default=-1,
                    help="step execution: 1 replay a captured HIP graph, 0 no graph, -1 autotuned")
    ap.add_argument("--margin", type=int, default=-1,
                    help="extra speculative Jacobi sweeps per correction (-1 autotuned, -2 none)")
    ap.add_argument("--tile-shape", type=int, default=-1,
                    help="face-kernel tile shape: 0 16x16, 1 32x16/16x32, -1 autotuned")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this benchmark measures one GPU only (asked for --gpus {args.gpus}); "
              "batched and served solves would shard across GPUs by job, one engine per GPU, "
              "with no communication",
              file=sys.stderr)
        sys.exit(2)

    import problems as P
    import swm

    n = args.size
    t0 = time.time()
    # the state, its well-balanced reference and the bathymetry tables made on
    # the device, bitwise the host builder's problems.lake_bumps(n, seed)
    s = swm.Solver.from_family(P.lake_bumps_spec(n, args.seed, steps=args.steps))
    s.set_execution(args.graph, args.margin)
    s.set_tile_shape(args.tile_shape)
    s.sync()
    log(f"bench: problem {n}x{n} lake_bumps seed {args.seed} made in {time.time() - t0:.1f} s")
    # the autotuner's candidate steps (execution mode and tile shape), then
    # the warmup steps; TimeStep and the step run on the device
    # (swm_run_steps: the device time loop, one host read per batch)
    s.run_steps(TUNING_STEPS)
    if args.warmup:
        s.run_steps(args.warmup)
    s.sync()
    log(f"bench: {TUNING_STEPS} tuning + {args.warmup} warmup steps done")

    # the timed region: HIP events around the face-kernel launches only (the
    # dominant kernels, for the roofline), which keeps graph replay off; every
    # other launch as in production
    s.profile(True, categories=FACE_CATEGORIES)
    s.sync()
    t0 = time.perf_counter()
    _, iters_t, _ = s.run_steps(args.steps)
    s.sync()
    elapsed = time.perf_counter() - t0
    prof_t = s.profile_read()
    s.profile(False)
    log(f"bench: {args.steps} timed steps in {elapsed:.3f} s")
    # as many steps again in the production mode (the tuned graph replay and
    # margin, no events), reported beside the timed figure
    s.sync()
    t1 = time.perf_counter()
    _, iters_p, _ = s.run_steps(args.steps)
    s.sync()
    el_prod = time.perf_counter() - t1
    log(f"bench: {args.steps} production-mode steps in {el_prod:.3f} s")
    # and again with an event pair around every launch: the per-kernel table
    s.profile(True)
    s.sync()
    t2 = time.perf_counter()
    _, iters, _ = s.run_steps(args.steps)
    s.sync()
    el_prof = time.perf_counter() - t2
    prof = s.profile_read()
    s.profile(False)
    log(f"bench: {args.steps} profiled steps in {el_prof:.3f} s")

    value = n * n * args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    cells = n * n

    def kernel_table(rec, sweeps):
        table = {k: {"ms_total": v[0], "launches": v[1],
                     "avg_us": (1e3 * v[0] / v[1]) if v[1] else 0.0}
                 for k, v in rec.items() if v[1]}
        for k, v in table.items():
            if k in ALGO_BYTES:
                u = units_per_step(k, cells, sweeps / args.steps) * args.steps / v["launches"]
                v["algo_bytes_per_launch"] = ALGO_BYTES[k] * u
                v["algo_GBps"] = ALGO_BYTES[k] * u / (v["avg_us"] * 1e-6) / 1e9
        return table

    jsum = int(np.sum(iters))
    kern = kernel_table(prof, jsum)
    kern_t = kernel_table(prof_t, int(np.sum(iters_t)))
    # the dominant kernel of the step (by the full table), timed in the timed region
    dom = max(kern, key=lambda k: kern[k]["ms_total"])
    src = kern_t if dom in kern_t else kern
    avg_s = src[dom]["avg_us"] * 1e-6
    algo = src[dom].get("algo_bytes_per_launch", 0.0)
    achieved = algo / avg_s / 1e9 if avg_s > 0 else 0.0
    roof = {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "algo_bytes_per_launch": algo, "avg_launch_us": src[dom]["avg_us"],
            "timed_in": "timed region" if src is kern_t else "profiled pass"}
    tpc = traffic_per_cell_node(dom, n)
    if tpc is not None:
        roof["traffic"] = tpc * units_per_step(dom, cells, 0) * args.steps / src[dom]["launches"]
    if dom in ("fvtd_x", "fvtd_y"):
        ops = FP64_OPS_PER_FACE * units_per_step(dom, cells, 0) * args.steps / src[dom]["launches"]
        tf = ops / avg_s / 1e12
        roof["fp64_valu"] = {"achieved": tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                             "frac": tf / FP64_PEAK_TFLOPS, "ops_per_launch": ops,
                             "note": "each +,-,*,/,sqrt counted once; the kernel is bound by "
                                     "FP64 VALU issue, see issue_frac_profiled"}
        roof["binding_roof"] = "fp64_valu_issue"
        roof["issue_frac_profiled"] = profiled("valu_issue_fraction", dom, n)
        roof["valu_instr_per_thread_node_profiled"] = profiled("valu_instr_per_thread_node", dom, n)
    out = {
        "metric": "mPDeC5 cell-steps/s FP64 at 2048\u00b2 vs CPU; batched jobs/s + p50/p99 latency",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lake at rest over Gaussian bumps + perturbation)",
        "config": {"workload": f"mPDeC5/WENO5 PATANKAR+WELLBALANCED {n}x{n} periodic, "
                               f"8 Gaussian bumps, CFL 0.5",
                   "grid": [n, n], "cfl": 0.5, "wellbalanced": True, "seed": args.seed},
        "jacobi_sweeps_per_step": int(np.sum(iters_t)) / args.steps,
        "tuning_steps": TUNING_STEPS,
        "timed_region": "HIP events around the face-kernel launches (the roofline's); "
                        "graph replay off; every other launch as in production",
        "production": {"cell_steps_per_s": n * n * args.steps / el_prod,
                       "ms_per_step": 1e3 * el_prod / args.steps,
                       "jacobi_sweeps_per_step": int(np.sum(iters_p)) / args.steps,
                       "note": "the next steps after the timed region, as many, in the "
                               "production mode (the tuned graph replay and margin, no events)"},
        "profiled": {"cell_steps_per_s": n * n * args.steps / el_prof,
                     "ms_per_step": 1e3 * el_prof / args.steps,
                     "jacobi_sweeps_per_step": jsum / args.steps,
                     "note": "the steps after those, as many, with a HIP event pair around "
                             "every kernel launch; 'kernels' comes from this pass"},
        "execution": dict(s.execution_stats(), tile_shape=s.tile_shape()[1],
                          status_reads=s.status_reads()),
        "kernels": kern,
        "roofline": roof,
    }
    # the boundary's host transfers: U in the reference's own layout through
    # swm_upload / swm_download (pageable host memory, as a Fortran caller
    # hands it over); not part of `value`, whose state stays on the device
    U = s.download()
    t0 = time.perf_counter()
    for _ in range(3):
        U = s.download()
    el_d = (time.perf_counter() - t0) / 3
    t0 = time.perf_counter()
    for _ in range(3):
        s.upload(U)
    s.sync()
    el_u = (time.perf_counter() - t0) / 3
    host_ms = 1e3 * (el_d + el_u)
    out["boundary_transfer"] = {
        "bytes_each_way": int(U.nbytes), "download_ms": 1e3 * el_d, "upload_ms": 1e3 * el_u,
        "cell_steps_per_s_with_upload_and_download_every_step":
            n * n / (1e-3 * (ms_per_step + host_ms)),
        "note": "U (ny, nx, 3) float64 through swm_download / swm_upload from pageable host "
                "memory; the drop-in moves U only at output times"}
    del s
    if args.scheme_steps > 0:
        out["schemes"] = schemes_section(n, args.seed, args.scheme_steps, not args.no_cpu)
    if args.initial_data:
        out["initial_data"] = initial_data_section(n, not args.no_cpu)
    if args.time_loop_steps > 0:
        out["time_loop"] = time_loop_section(args.time_loop_steps)
    if args.sync_cost:
        out["sync_cost"] = sync_cost_section()
    if args.fix_steps > 0:
        out["reconstruction_fix"] = recfix_section(n, args.fix_steps, not args.no_cpu)
    if args.batch_jobs > 0:
        out["batched"] = batched_section(args.batch_jobs, args.seed)
    if args.serve_requests > 0:
        out["served"] = served_section(args.serve_requests, args.seed, args.serve_rate)
        if args.serve_ic_requests > 0:
            out["served_by_initial_condition"] = served_ic_section(args.serve_ic_requests,
                                                                   args.seed, args.serve_rate)
        rates = [float(r) for r in args.serve_rates.split(",") if r.strip()]
        if rates:
            out["served_sweep"] = [served_section(args.serve_requests, args.seed, r)
                                   for r in rates]
    if not args.no_cpu:
        cb = cpu_baseline()
        if cb:
            out["cpu_baseline"] = cb
            out["gpu_vs_cpu"] = value / cb["value"]
            port = cb.get("port_on_headline_family")
            if port:
                out["gpu_vs_port_on_headline_family"] = value / port["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
