"""The two other modified-Patankar schemes of the PATANKAR build, mPEuler
(timescheme 6) and mPDeC2 (timescheme 7), on the mPDeC5 kernels: bitwise
against the oracle (itself pinned to the reference by
test_oracle_vs_reference.py and the golden fixtures *_mpeuler / *_mpdec2),
with both depth views -- the reference's sequence-associated hp, whose states
turn NaN within a few steps, and the by-node view -- on the plain solver, the
multi-step entry points, the engine (schemes mixed in one engine) and every
execution mode."""
from __future__ import annotations

import numpy as np
import pytest

import problems as P
import swm
from oracle_api import Oracle
from parity_util import describe_diff, same_bits

pytestmark = pytest.mark.gpu

VIEWS = {"ref": False, "by_node": True}


def random_state(n, ny, seed, amp=0.3):
    rng = np.random.default_rng(seed)
    U = np.empty((ny, n, 3))
    U[..., 0] = 1.0 + amp * rng.uniform(-1, 1, (ny, n))
    U[..., 1] = U[..., 0] * rng.normal(0, 0.3, (ny, n))
    U[..., 2] = U[..., 0] * rng.normal(0, 0.3, (ny, n))
    return P.Problem(n, U, kind="random", ny=ny)


CASES = {
    "vortex24": lambda: P.vortex(24, 3),
    "lake_bumps20": lambda: P.lake_bumps(20, 4),
    "flag1wb16": lambda: P.lake_flag1(16, True),
    "shallowdam18": lambda: P.dam_break(18, dry=0.1),
    "random_rect17x11": lambda: random_state(17, 11, 5),
    "random_3x29": lambda: random_state(3, 29, 6),
}


def solver(p, scheme, by_node):
    s = swm.Solver.from_problem(p)
    s.set_timescheme(scheme)
    s.set_hp_association(swm.HP_BY_NODE if by_node else swm.HP_AS_REFERENCE)
    return s


@pytest.mark.parametrize("view", sorted(VIEWS))
@pytest.mark.parametrize("scheme", [6, 7])
@pytest.mark.parametrize("case", sorted(CASES))
def test_steps_match_oracle(case, scheme, view):
    """Host-driven steps (swm_step_mpeuler / swm_step_mpdec2 with the
    TimeStep dt), sweep counts included, up to 4 steps or the first
    non-finite dt."""
    p = CASES[case]()
    by_node = VIEWS[view]
    s = solver(p, scheme, by_node)
    o = Oracle(p, scheme, hp_by_node=by_node)
    for k in range(4):
        dt, lx, ly = s.timestep()
        dt_r, lx_r, ly_r = o.timestep()
        assert same_bits([dt, lx, ly], [dt_r, lx_r, ly_r]), (case, k)
        if not np.isfinite(dt):
            assert not by_node, "the by-node view keeps states finite here"
            break
        it, _ = s.step(dt)
        it_r, _ = o.step(dt)
        assert it == it_r, (case, k)
        U, U_r = s.download(), o.get_U()
        assert same_bits(U, U_r), f"{case} step {k}: {describe_diff(U, U_r)}"


@pytest.mark.parametrize("scheme", [6, 7])
def test_run_steps_and_advance_by_node(scheme):
    p = P.lake_bumps(28, 7)
    s = solver(p, scheme, True)
    dts, iters, _ = s.run_steps(3)
    assert iters.shape == (3, 1 if scheme == 6 else 2)
    o = Oracle(p, scheme, hp_by_node=True)
    for k in range(3):
        dt, _, _ = o.timestep()
        it, _ = o.step(dt)
        assert dt == dts[k] and list(iters[k]) == it
    assert same_bits(s.download(), o.get_U())
    # the output loop to t_analyze = 2.5 dt0 on top of that
    t0 = 0.0
    dt0, _, _ = o.timestep()
    t_an = t0 + 2.5 * dt0
    t, dts, _ = s.advance(t0, t_an, 10 * dt0)
    tr = t0
    while True:
        dt_min, _, _ = o.timestep()
        dt = min(min(dt_min, t_an - tr), min(dt_min, 10 * dt0 - tr))
        o.step(dt)
        tr = tr + dt
        if abs(t_an - tr) < 1e-10:
            break
    assert t == tr
    assert same_bits(s.download(), o.get_U())


@pytest.mark.parametrize("mode", [(1, 1), (0, 1), (1, 0), (-1, -2)],
                         ids=["graph", "nograph", "graph_margin0", "polling"])
@pytest.mark.parametrize("scheme", [6, 7])
def test_execution_modes_by_node(scheme, mode):
    p = P.vortex(37, 2)
    s = solver(p, scheme, True)
    s.set_execution(*mode)
    dts, _, _ = s.run_steps(5)
    o = Oracle(p, scheme, hp_by_node=True)
    for k in range(5):
        dt, _, _ = o.timestep()
        o.step(dt)
        assert dt == dts[k]
    assert same_bits(s.download(), o.get_U())
    st = s.execution_stats()
    if mode[1] != -2:
        assert st["speculative_steps"] >= 4 - st["redone_steps"]
    if mode[0] == 1:
        assert st["graph_launches"] >= 1


def test_engine_mixes_schemes_and_views():
    """One engine holding mPDeC5, mPEuler and mPDeC2 jobs with both views:
    each job equals the oracle; a reference-view job fails (non-finite dt) at
    the step where the oracle's dt turns NaN, without disturbing the others."""
    jobs = [(P.vortex(20, 1), 3, 5, False), (P.lake_bumps(33, 2), 3, 6, True),
            (P.vortex(24, 3), 4, 7, True), (P.lake_flag1(18, True), 3, 7, True),
            (P.vortex(16, 4), 4, 6, False), (P.vortex(21, 6), 5, 7, False),
            (P.dam_break(26, dry=0.1), 3, 5, False)]
    eng = swm.Engine()
    eng.set_execution(1, 1)
    ids = [eng.submit(p, k, timescheme=sc,
                      hp_association=swm.HP_BY_NODE if bn else swm.HP_AS_REFERENCE)
           for p, k, sc, bn in jobs]
    while eng.step():
        pass
    for (p, k, sc, bn), jid in zip(jobs, ids):
        o = Oracle(p, sc, hp_by_node=bn)
        done = 0
        for _ in range(k):
            dt, _, _ = o.timestep()
            if not np.isfinite(dt):
                break
            o.step(dt)
            done += 1
        info = eng.info(jid)
        assert info["steps_done"] == done, (p.kind, sc, bn)
        assert info["state"] == (swm.JOB_DONE if done == k else swm.JOB_FAILED)
        assert same_bits(eng.fetch(jid), o.get_U()), (p.kind, sc, bn)


def test_same_job_alone_and_in_mix():
    target = P.lake_bumps(40, 9)
    eng1 = swm.Engine()
    a = eng1.submit(target, 3, timescheme=7, hp_association=swm.HP_BY_NODE)
    while eng1.step():
        pass
    eng2 = swm.Engine()
    for p, k, sc in [(P.vortex(20, 1), 2, 5), (P.vortex(30, 2), 3, 6), (P.lake_bumps(25, 3), 2, 7)]:
        eng2.submit(p, k, timescheme=sc, hp_association=swm.HP_BY_NODE)
    b = eng2.submit(target, 3, timescheme=7, hp_association=swm.HP_BY_NODE)
    while eng2.step():
        pass
    assert same_bits(eng2.fetch(b), eng1.fetch(a))


def test_bad_scheme_is_rejected():
    p = P.vortex(12, 1)
    s = swm.Solver.from_problem(p)
    for bad in (0, 8, -1):
        with pytest.raises(swm.SwmError):
            s.set_timescheme(bad)
    with pytest.raises(swm.SwmError):
        s.set_hp_association(2)
    eng = swm.Engine()
    with pytest.raises(swm.SwmError):
        eng.submit(p, 1, timescheme=9)


@pytest.mark.parametrize("scheme", [6, 7])
def test_by_node_large_grid_conserves_mass_and_positivity(scheme):
    """A 1024^2 well-balanced lake over Gaussian bumps: the by-node schemes
    keep every stage depth non-negative and the total depth to rounding over
    3 steps."""
    p = P.lake_bumps(1024, 11)
    s = solver(p, scheme, True)
    s.track_stage_min(True)
    m0 = p.U0[..., 0].sum()
    s.run_steps(3)
    U = s.download()
    assert s.stage_min_h() >= 0.0
    assert (U[..., 0] >= 0).all()
    assert abs(U[..., 0].sum() - m0) <= 1e-11 * m0
