"""Step execution modes: speculative steps (predicted Jacobi sweeps launched
without read-backs, one check per step), their HIP-graph replay, the polling
fallback and the redo of an under-predicted step all give the oracle's
results bitwise."""
from __future__ import annotations

import numpy as np
import pytest

import problems as P
import swm
from oracle_api import Oracle
from parity_util import describe_diff, same_bits

pytestmark = pytest.mark.gpu

MODES = {"graph": (1, 1), "nograph": (0, 1), "graph_margin0": (1, 0), "polling": (-1, -2),
         "auto": (-1, -1)}


def oracle_run(p, steps, limits=None):
    o = Oracle(p)
    dts = []
    for k in range(steps):
        if limits:
            o.set_jacobi_limits(*limits(k))
        dt, _, _ = o.timestep()
        o.step(dt)
        dts.append(dt)
    return o.get_U(), np.array(dts)


@pytest.mark.parametrize("mode", sorted(MODES))
@pytest.mark.parametrize("make", [lambda: P.vortex(37, 2, steps=6),
                                  lambda: P.lake_bumps(40, 3, steps=5)], ids=["vortex37", "lake40"])
def test_modes_match_oracle(mode, make):
    p = make()
    s = swm.Solver.from_problem(p)
    s.set_execution(*MODES[mode])
    dts, iters, _ = s.run_steps(p.steps)
    U = s.download()
    st = s.execution_stats()
    U_r, dts_r = oracle_run(p, p.steps)
    assert same_bits(dts, dts_r)
    assert same_bits(U, U_r), f"{mode}: {describe_diff(U, U_r)}"
    if mode == "polling":
        assert st["speculative_steps"] == 0
    else:
        assert st["speculative_steps"] >= p.steps - 1 - st["redone_steps"]
    if mode in ("graph", "graph_margin0"):
        assert st["graph_launches"] >= 1 and st["graph_captures"] >= 1


def test_underpredicted_step_is_redone():
    """Step 1 runs with a 2-sweep limit, so its counts predict 2 sweeps; the
    limit is then lifted and, with no margin, step 2's solves need more
    sweeps than were launched: the step is redone and still exact."""
    p = P.vortex(32, 4, steps=3)
    s = swm.Solver.from_problem(p)
    s.set_execution(1, 0)
    s.set_jacobi_limits(2, 1e-16)
    s.run_steps(1)
    s.set_jacobi_limits(1000, 1e-16)
    s.run_steps(2)
    U = s.download()
    st = s.execution_stats()
    assert st["redone_steps"] >= 1
    U_r, _ = oracle_run(p, 3, limits=lambda k: (2, 1e-16) if k == 0 else (1000, 1e-16))
    assert same_bits(U, U_r), describe_diff(U, U_r)


def test_redone_step_keeps_only_the_accepted_stage_minima():
    """The stage-depth record of a step that is redone after an overrun holds
    the accepted attempt's minimum only: the same as with polling alone."""
    p = P.vortex(32, 4, steps=3)

    def run(graph, margin):
        s = swm.Solver.from_problem(p)
        s.set_execution(graph, margin)
        s.set_jacobi_limits(2, 1e-16)
        s.run_steps(1)
        s.set_jacobi_limits(1000, 1e-16)
        s.track_stage_min(True)
        s.run_steps(2)
        return s.stage_min_h(), s.execution_stats(), s.download()

    h_spec, st, U_spec = run(1, 0)
    h_poll, st_poll, U_poll = run(-1, -2)
    assert st["redone_steps"] >= 1 and st_poll["speculative_steps"] == 0
    assert same_bits([h_spec], [h_poll]), (h_spec, h_poll)
    assert same_bits(U_spec, U_poll)


def test_engine_graph_mode_matches_oracle():
    jobs = [(P.vortex(20, 1), 4), (P.lake_bumps(33, 2), 3), (P.vortex(47, 3), 5),
            (P.lake_flag1(18, True), 4)]
    eng = swm.Engine()
    eng.set_execution(1, 1)
    ids = [eng.submit(p, k) for p, k in jobs]
    while eng.step():
        pass
    st = eng.execution_stats()
    assert st["graph_launches"] >= 1
    for (p, k), jid in zip(jobs, ids):
        U_r, _ = oracle_run(p, k)
        assert same_bits(eng.fetch(jid), U_r), p.kind
