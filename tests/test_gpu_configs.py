"""GPU tests at the sizes of the benchmark configurations.

* 64x64 vortex, CFL 0.5, 50 steps (the CPU-runnable case): dt sequence and
  final state bitwise equal to the oracle, hence relative L1 = 0 <= 1e-12.
* 2048x2048 well-balanced lake over 8 bumps + 1e-3 perturbation (the headline
  case), its 100 steps: depth >= 0 at every DeC stage, finite state, total
  mass conserved to roundoff; without the perturbation the lake at rest stays
  at rest to 1e-13.  (Bitwise parity at this size: test_gpu_config_pins.py.)
* 1024x1024 circular dam break: over a dry bed (h = 0) the first step keeps
  h >= 0 at every stage and TimeStep then turns non-finite, where the
  reference's time loop never ends; over a 1e-3 film, 200 steps with h >= 0
  at every stage, bounded Jacobi counts and conserved mass.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import problems as P
import swm
from oracle_api import Oracle
from parity_util import describe_diff, rel_l1, same_bits

pytestmark = pytest.mark.gpu


def mass(U):
    return math.fsum(np.asarray(U)[..., 0].ravel().tolist())


def test_vortex_64_50_steps_matches_oracle():
    p = P.vortex(64, 1, steps=50)
    s = swm.Solver.from_problem(p)
    dts, iters, nonconv = s.run_steps(50)
    U = s.download()
    o = Oracle(p)
    dts_r = []
    for _ in range(50):
        dt, _, _ = o.timestep()
        o.step(dt)
        dts_r.append(dt)
    U_r = o.get_U()
    assert same_bits(dts, np.array(dts_r))
    assert same_bits(U, U_r), describe_diff(U, U_r)
    for v in range(3):
        assert rel_l1(U, U_r, v) <= 1e-12
    assert nonconv == 0


def test_lake_bumps_2048_stages_positive_and_mass_conserved():
    """The headline configuration over its full 100 steps (data made on the
    device, bitwise the host builder's)."""
    s = swm.Solver.from_family(P.lake_bumps_spec(2048, 2024, steps=100))
    m0 = mass(s.download())
    s.track_stage_min(True)
    dts, iters, nonconv = s.run_steps(100)
    hmin = s.stage_min_h()
    U = s.download()
    assert np.isfinite(U).all()
    assert hmin >= 0.0, hmin
    assert (U[..., 0] >= 0).all()
    rel = abs(mass(U) - m0) / m0
    assert rel <= 1e-12, rel
    assert nonconv == 0 and (iters >= 1).all() and (iters < 1000).all()


def test_lake_at_rest_2048_preserved():
    p = P.lake_bumps(2048, 7, steps=3, amp=0.0)
    U0 = p.U0.copy()
    s = swm.Solver.from_problem(p)
    del p
    s.run_steps(3)
    U = s.download()
    assert np.max(np.abs(U[..., 0] - U0[..., 0])) <= 1e-13
    assert np.max(np.abs(U[..., 1:])) <= 1e-13


def test_dam_break_1024_dry_bed():
    p = P.dam_break(1024)
    s = swm.Solver.from_problem(p)
    s.track_stage_min(True)
    s.run_steps(1)
    hmin = s.stage_min_h()
    assert hmin >= 0.0, hmin
    assert (s.download()[..., 0] >= 0).all()
    with pytest.raises(swm.SwmError):
        s.run_steps(1)  # TimeStep is no longer finite


def test_dam_break_1024_thin_film_200_steps():
    p = P.dam_break(1024, dry=1e-3)
    m0 = mass(p.U0)
    s = swm.Solver.from_problem(p)
    s.track_stage_min(True)
    dts, iters, nonconv = s.run_steps(200)
    hmin = s.stage_min_h()
    U = s.download()
    assert np.isfinite(U).all()
    assert hmin >= 0.0, hmin
    assert nonconv == 0 and iters.max() < 1000
    rel = abs(mass(U) - m0) / m0
    assert rel <= 1e-11, rel


def _min_r(a, b):
    return a if a < b else b


@pytest.mark.parametrize("make", [lambda: P.vortex(40, 11), lambda: P.lake_flag1(24, True)],
                         ids=["vortex40", "lake_flag1_wb24"])
def test_advance_reproduces_the_time_loop(make):
    """swm_advance against the reference's loop (src/timediscretization.f90:
    46-95) run on the oracle: two output intervals, dt clipped to the output
    times and to tEnd, t accumulated in the same double arithmetic."""
    p = make()
    o = Oracle(p)
    dt0, _, _ = o.timestep()
    t_end = 9.5 * dt0
    n_out = 2
    dt_an = t_end / float(n_out)
    # the oracle, as the reference's loop
    t, t_an, ref_t, ref_dts = 0.0, dt_an, [], []
    o = Oracle(p)
    while True:
        dt_min, _, _ = o.timestep()
        dt = _min_r(_min_r(dt_min, t_an - t), _min_r(dt_min, t_end - t))
        o.step(dt)
        ref_dts.append(dt)
        t = t + dt
        if abs(t_an - t) < 1.0e-10:
            ref_t.append(t)
            t_an = t_an + dt_an
            if t_an > t_end:
                t_an = t_end
        if abs(t - t_end) < 1.0e-10:
            break
    U_r = o.get_U()
    # the device, one swm_advance per output interval
    s = swm.Solver.from_problem(p)
    t, t_an, got_t, got_dts = 0.0, dt_an, [], []
    while True:
        t, dts, _ = s.advance(t, t_an, t_end)
        got_dts.extend(dts)
        if abs(t_an - t) < 1.0e-10:
            got_t.append(t)
            t_an = t_an + dt_an
            if t_an > t_end:
                t_an = t_end
        if abs(t - t_end) < 1.0e-10:
            break
    assert got_t == ref_t
    assert same_bits(np.array(got_dts), np.array(ref_dts))
    U = s.download()
    assert same_bits(U, U_r), describe_diff(U, U_r)


@pytest.mark.parametrize("n,ny", [(4096, 0), (3072, 1536)], ids=["4096x4096", "3072x1536"])
def test_large_grid_mass_and_positivity(n, ny):
    """Grids beyond the benchmark size (16.8M and 4.7M cells): tile tables,
    partial sums and 64-bit indexing at scale; mass conserved to roundoff,
    depths positive at every stage, state finite."""
    p = P.vortex(n, 21, ny=ny)
    m0 = mass(p.U0)
    s = swm.Solver.from_problem(p)
    del p
    s.track_stage_min(True)
    dts, iters, nonconv = s.run_steps(2)
    U = s.download()
    assert np.isfinite(U).all()
    assert s.stage_min_h() > 0.0
    assert abs(mass(U) - m0) / m0 <= 1e-12
    assert nonconv == 0
