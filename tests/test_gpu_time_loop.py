"""The device time loop: batches of up to 16 steps between host reads of the
device status, with the commit, the records and the stops decided on the
device (k_gate).  Results are bitwise those of one step per host read (and
of the oracle), and the host waits once per batch."""
from __future__ import annotations

import math

import numpy as np
import pytest

import problems as P
import swm
from oracle_api import Oracle
from parity_util import describe_diff, same_bits

pytestmark = pytest.mark.gpu

BATCH = 16  # kMaxBatch


def oracle_run(p, steps):
    o = Oracle(p)
    dts, its = [], []
    for _ in range(steps):
        dt, _, _ = o.timestep()
        it, _ = o.step(dt)
        dts.append(dt)
        its.append(list(it))
    return o.get_U(), np.array(dts), np.array(its)


@pytest.mark.parametrize("graph", [1, 0], ids=["graph", "nograph"])
def test_run_steps_one_status_read_per_batch(graph):
    """40 steps of a 24x24 vortex: after the first (polling) step, 39 steps
    take ceil(39/16) = 3 host reads when no step is redone; dt sequence,
    sweep counts and state are the oracle's."""
    p = P.vortex(24, 6, steps=40)
    s = swm.Solver.from_problem(p)
    s.set_execution(graph, 2)
    s.set_tile_shape(0)  # every setting explicit: no autotuner steps
    dts0, its0, _ = s.run_steps(1)
    r0 = s.status_reads()
    dts1, its1, _ = s.run_steps(39)
    reads = s.status_reads() - r0
    st = s.execution_stats()
    U_r, dts_r, its_r = oracle_run(p, 40)
    assert same_bits(np.concatenate([dts0, dts1]), dts_r)
    assert (np.concatenate([its0, its1]) == its_r).all()
    assert same_bits(s.download(), U_r), describe_diff(s.download(), U_r)
    if st["redone_steps"] == 0:
        assert reads == math.ceil(39 / BATCH), reads
    else:
        assert reads < 39


def test_advance_batches_stop_at_the_output_time():
    """swm_advance over output intervals of ~20 steps: the device clips dt to
    the output time and stops there; dts, t and the state match a loop of
    single steps on the oracle, and the batches need fewer reads than steps."""
    p = P.lake_flag1(20, True)
    o = Oracle(p)
    dt0, _, _ = o.timestep()
    t_end = 41.5 * dt0
    dt_an = t_end / 2.0
    # the reference's loop on the oracle (src/timediscretization.f90:46-95)
    t, t_an, ref_dts, ref_t = 0.0, dt_an, [], []
    while True:
        dt_min, _, _ = o.timestep()
        dt = min(min(dt_min, t_an - t), min(dt_min, t_end - t))
        o.step(dt)
        ref_dts.append(dt)
        t = t + dt
        if abs(t_an - t) < 1.0e-10:
            ref_t.append(t)
            t_an = min(t_an + dt_an, t_end)
        if abs(t - t_end) < 1.0e-10:
            break
    s = swm.Solver.from_problem(p)
    s.set_execution(1, 2)
    s.set_tile_shape(0)
    t, t_an, got_dts, got_t, reads = 0.0, dt_an, [], [], []
    while True:
        r0 = s.status_reads()
        t, dts, _ = s.advance(t, t_an, t_end)
        reads.append((s.status_reads() - r0, len(dts)))
        got_dts.extend(dts)
        if abs(t_an - t) < 1.0e-10:
            got_t.append(t)
            t_an = min(t_an + dt_an, t_end)
        if abs(t - t_end) < 1.0e-10:
            break
    assert got_t == ref_t
    assert same_bits(np.array(got_dts), np.array(ref_dts))
    assert same_bits(s.download(), o.get_U())
    # the second interval (~21 steps, predictions known): a few batches
    assert len(reads) == 2 and reads[1][1] >= 15 and reads[1][0] <= 4, reads


def test_engine_run_matches_engine_step():
    """Jobs of 3..33 steps, some finishing inside a batch: swm_engine_run gives
    every job swm_engine_step's state, step count, t and sweep total, with
    fewer host reads."""
    jobs = [(P.vortex(20, 1), 3), (P.lake_bumps(33, 2), 7), (P.vortex(47, 3), 20),
            (P.lake_flag1(18, True), 33), (P.vortex(16, 9), 17)]

    def run(batched):
        eng = swm.Engine()
        eng.set_execution(1, 2)
        eng.set_tile_shape(0)
        ids = [eng.submit(p, k) for p, k in jobs]
        eng.step()  # the first step polls (no sweep-count prediction yet)
        r0 = eng.status_reads()
        if batched:
            while eng.run(16)[0]:
                pass
        else:
            while eng.step():
                pass
        return eng, ids, eng.status_reads() - r0

    a, ida, reads_a = run(False)
    b, idb, reads_b = run(True)
    # 32 steps after the first: one read each, against a few batches
    assert reads_a >= 32 and reads_b <= 8, (reads_a, reads_b)
    for (p, k), ja, jb in zip(jobs, ida, idb):
        ia, ib = a.info(ja), b.info(jb)
        assert ia["steps_done"] == ib["steps_done"] == k
        assert ia["state"] == ib["state"] == swm.JOB_DONE
        assert same_bits([ia["t"]], [ib["t"]])
        assert ia["jacobi_sweeps"] == ib["jacobi_sweeps"]
        assert same_bits(a.fetch(ja), b.fetch(jb)), describe_diff(a.fetch(ja), b.fetch(jb))


def test_non_finite_dt_inside_a_batch_keeps_the_last_state():
    """A dry-bed dam break whose dt turns non-finite after its first step:
    run_steps(5) stops there with an error and the state of step 1, as a
    step-by-step run does; in an engine the job fails while a vortex in the
    same batch runs to the end, bitwise the oracle's."""
    p = P.dam_break(48)
    a = swm.Solver.from_problem(p)
    done = 0
    with pytest.raises(swm.SwmError):
        for _ in range(5):
            a.run_steps(1)
            done += 1
    assert 1 <= done < 5
    b = swm.Solver.from_problem(p)
    b.set_execution(1, 2)
    with pytest.raises(swm.SwmError):
        b.run_steps(5)
    assert same_bits(a.download(), b.download())
    v = P.vortex(24, 4)
    eng = swm.Engine()
    ja, jv = eng.submit(p, 5), eng.submit(v, 12)
    while eng.run(16)[0]:
        pass
    assert eng.info(ja)["state"] == swm.JOB_FAILED
    assert eng.info(ja)["steps_done"] == done
    assert same_bits(eng.fetch(ja), a.download())
    U_r, _, _ = oracle_run(v, 12)
    assert eng.info(jv)["steps_done"] == 12
    assert same_bits(eng.fetch(jv), U_r), describe_diff(eng.fetch(jv), U_r)


@pytest.mark.parametrize("ts", [2, 4], ids=["ssprk4", "dec5"])
def test_explicit_schemes_run_whole_batches(ts):
    """The explicit schemes solve nothing, so every step of a batch runs
    without host reads: 20 steps take 2 reads, bitwise one step per call."""
    p = P.lake_flag1(16, True)
    a = swm.Solver.from_problem(p)
    a.set_timescheme(ts)
    for _ in range(20):
        a.run_steps(1)
    b = swm.Solver.from_problem(p)
    b.set_timescheme(ts)
    r0 = b.status_reads()
    dts, _, _ = b.run_steps(20)
    assert b.status_reads() - r0 == 2
    assert same_bits(a.download(), b.download())
