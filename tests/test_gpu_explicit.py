"""The explicit schemes of the reference's build without PATANKAR
(timescheme 1 ForwardEuler, 2 SSPRK4, 3 RK65, 4 DeC5) on the FVTD kernels with
Ut(1) and the stage-combination kernel: bitwise against the oracle (itself
pinned to that build of the reference by test_oracle_vs_reference.py and the
golden fixtures *_euler, *_ssprk4, *_rk65, *_dec5), through the plain solver,
the multi-step entry points and the engine (mixed with Patankar jobs)."""
from __future__ import annotations

import numpy as np
import pytest

import problems as P
import swm
from oracle_api import Oracle
from parity_util import describe_diff, same_bits

pytestmark = pytest.mark.gpu

SCHEMES = [1, 2, 3, 4]


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


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("case", sorted(CASES))
def test_steps_match_oracle(case, scheme):
    p = CASES[case]()
    s = swm.Solver.from_problem(p)
    s.set_timescheme(scheme)
    o = Oracle(p, scheme)
    for k in range(3):
        dt, lx, ly = s.timestep()
        dt_r, lx_r, ly_r = o.timestep()
        assert same_bits([dt, lx, ly], [dt_r, lx_r, ly_r]), (case, k)
        s.step(dt)
        o.step(dt)
        U, U_r = s.download(), o.get_U()
        assert same_bits(U, U_r), f"{case} step {k}: {describe_diff(U, U_r)}"


@pytest.mark.parametrize("scheme", SCHEMES)
def test_run_steps_and_advance(scheme):
    p = P.lake_bumps(28, 7)
    s = swm.Solver.from_problem(p)
    s.set_timescheme(scheme)
    dts, iters, nonconv = s.run_steps(3)
    assert iters.shape == (3, 0) and nonconv == 0
    o = Oracle(p, scheme)
    for k in range(3):
        dt, _, _ = o.timestep()
        o.step(dt)
        assert dt == dts[k]
    assert same_bits(s.download(), o.get_U())
    dt0, _, _ = o.timestep()
    t_an = 2.5 * dt0
    t, _, _ = s.advance(0.0, t_an, 10 * dt0)
    tr = 0.0
    while True:
        dt_min, _, _ = o.timestep()
        dt = min(min(dt_min, t_an - tr), min(dt_min, 10 * dt0 - tr))
        o.step(dt)
        tr = tr + dt
        if abs(t_an - tr) < 1e-10:
            break
    assert t == tr
    assert same_bits(s.download(), o.get_U())


def test_engine_mixes_explicit_and_patankar_jobs():
    jobs = [(P.vortex(20, 1), 3, 1), (P.lake_bumps(33, 2), 2, 2), (P.vortex(24, 3), 3, 3),
            (P.lake_flag1(18, True), 2, 4), (P.vortex(16, 4), 2, 5), (P.vortex(21, 6), 3, 2),
            (P.dam_break(26, dry=0.1), 2, 4), (P.vortex(19, 8), 2, 7)]
    eng = swm.Engine()
    ids = [eng.submit(p, k, timescheme=sc, hp_association=swm.HP_BY_NODE) for p, k, sc in jobs]
    while eng.step():
        pass
    for (p, k, sc), jid in zip(jobs, ids):
        o = Oracle(p, sc, hp_by_node=True)
        for _ in range(k):
            dt, _, _ = o.timestep()
            o.step(dt)
        info = eng.info(jid)
        assert info["state"] == swm.JOB_DONE and info["steps_done"] == k
        assert same_bits(eng.fetch(jid), o.get_U()), (p.kind, p.n, sc)


def test_scheme_switch_on_one_handle():
    """A handle stepped with mPDeC5, then SSPRK4, then mPDeC2 (by node), then
    RK65: the explicit steps reuse the Patankar buffers, so each switch must
    leave nothing stale behind."""
    p = P.lake_bumps(30, 12)
    s = swm.Solver.from_problem(p)
    s.set_hp_association(swm.HP_BY_NODE)
    o = Oracle(p, 5, hp_by_node=True)
    for sc in [5, 2, 7, 3, 5, 4, 6]:
        dt, _, _ = s.timestep()
        dt_r, _, _ = o.timestep()
        assert dt == dt_r
        s.step(dt, timescheme=sc)
        o.scheme = sc
        o.step(dt)
        assert same_bits(s.download(), o.get_U()), sc


@pytest.mark.parametrize("scheme", SCHEMES)
def test_large_grid_conserves_mass_and_keeps_lake(scheme):
    """1024^2: the explicit schemes conserve the total depth to rounding; the
    well-balanced lake at rest stays bitwise, except under SSPRK4: its convex
    stage combinations a*K0 + b*K1 of equal states round h by an ulp, which
    the fluxes then turn into O(ulp/dx) velocities (a property of the
    reference's scheme: the oracle shows the same at small sizes)."""
    p = P.lake_bumps(1024, 11)
    s = swm.Solver.from_problem(p)
    s.set_timescheme(scheme)
    m0 = p.U0[..., 0].sum()
    s.run_steps(2)
    U = s.download()
    assert abs(U[..., 0].sum() - m0) <= 1e-11 * m0
    lake = P.lake_bumps(1024, 11, amp=0.0)
    s = swm.Solver.from_problem(lake)
    s.set_timescheme(scheme)
    s.run_steps(2)
    U = s.download()
    if scheme == 2:
        assert np.abs(U - lake.U0).max() < 1e-7
    else:
        assert same_bits(U, lake.U0)


@pytest.mark.parametrize("scheme", [1, 2, 3, 4, 6, 7])
def test_dry_dam_break_nan_pattern(scheme):
    """A dam break over a dry bed (h = 0 outside the column): ConsToPrim
    divides by zero depths, the states turn NaN after one step and the next
    dt is NaN, Inf or absurd.  The dt sequence and the NaN pattern of every
    step agree with the oracle (NaNs compared by position)."""
    p = P.dam_break(20)
    s = swm.Solver.from_problem(p)
    s.set_timescheme(scheme)
    s.set_hp_association(swm.HP_BY_NODE)
    o = Oracle(p, scheme, hp_by_node=True)
    for k in range(3):
        dt, lx, ly = s.timestep()
        ref = o.timestep()
        assert same_bits([dt, lx, ly], list(ref)), (scheme, k)
        if not np.isfinite(dt):
            break
        s.step(dt)
        o.step(dt)
        U, U_r = s.download(), o.get_U()
        assert same_bits(U, U_r), f"scheme {scheme} step {k}: {describe_diff(U, U_r)}"
