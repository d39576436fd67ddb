"""The face kernels' tile shapes (swm_set_tile_shape): every shape computes
each face with the same arithmetic, so results are bitwise the same whatever
the shape -- checked against the oracle for each shape on square and
rectangular grids down to 3 cells, under several schemes -- and the
autotuner settles on one of them."""
from __future__ import annotations

import numpy as np
import pytest

import problems as P
import swm
from oracle_api import Oracle
from parity_util import describe_diff, same_bits

pytestmark = pytest.mark.gpu


def random_state(n, ny, seed, amp=0.3):
    rng = np.random.default_rng(seed)
    U = np.empty((ny, n, 3))
    U[..., 0] = 1.0 + amp * rng.uniform(-1, 1, (ny, n))
    U[..., 1] = U[..., 0] * rng.normal(0, 0.3, (ny, n))
    U[..., 2] = U[..., 0] * rng.normal(0, 0.3, (ny, n))
    return P.Problem(n, U, kind="random", ny=ny)


CASES = {
    "vortex40x33": lambda: P.vortex(40, 2, ny=33),
    "lake_bumps70": lambda: P.lake_bumps(70, 3),
    "random_3x41": lambda: random_state(3, 41, 4),
    "random_37x5": lambda: random_state(37, 5, 5),
    "flag1wb31x62": lambda: P.lake_flag1(31, True, ny=62),
}


@pytest.mark.parametrize("shape", [0, 1])
@pytest.mark.parametrize("scheme", [5, 2, 7])
@pytest.mark.parametrize("case", sorted(CASES))
def test_shapes_match_oracle(case, scheme, shape):
    p = CASES[case]()
    s = swm.Solver.from_problem(p)
    s.set_tile_shape(shape)
    s.set_timescheme(scheme)
    s.set_hp_association(swm.HP_BY_NODE)
    assert s.tile_shape()[0] == shape
    dts, _, _ = s.run_steps(3)
    o = Oracle(p, scheme, hp_by_node=True)
    for k in range(3):
        dt, _, _ = o.timestep()
        o.step(dt)
        assert dt == dts[k]
    U, U_r = s.download(), o.get_U()
    assert same_bits(U, U_r), f"{case} scheme {scheme} shape {shape}: {describe_diff(U, U_r)}"


def test_fvtd_outputs_are_shape_independent():
    p = P.lake_bumps(45, 8)
    outs = []
    for shape in (0, 1):
        s = swm.Solver.from_problem(p)
        s.set_tile_shape(shape)
        outs.append(s.fvtd())
    for k in outs[0]:
        assert same_bits(outs[0][k], outs[1][k]), k


def test_autotuner_settles_on_a_shape():
    p = P.vortex(96, 7)
    s = swm.Solver.from_problem(p)
    # one polling step, then two rounds over the 2 shapes and over the 4
    # (graph, margin) settings, plus the steps that capture a graph
    s.run_steps(20)
    shape, dims = s.tile_shape()
    assert shape in (0, 1) and len(dims) == 4
    st = s.execution_stats()
    assert st["graph"] in (0, 1) and st["margin"] in (1, 2)
    U = s.download()
    o = Oracle(p)
    o.run(20)
    assert same_bits(U, o.get_U())
    eng = swm.Engine()
    with pytest.raises(swm.SwmError):
        eng.set_tile_shape(2)


def test_tuner_leaves_unread_profile_records():
    """The tuner times the tile shapes by HIP events of its own; records a
    caller took with profiling on and has not read yet stay as they were."""
    p = P.vortex(48, 5)
    s = swm.Solver.from_problem(p)
    s.profile(True)
    s.run_steps(1)  # the polling step: 5 launches of each face kernel
    s.profile(False)
    s.run_steps(20)  # the tuning steps
    rec = s.profile_read()
    assert rec["fvtd_x"][1] == 5 and rec["fvtd_y"][1] == 5
    assert rec["fvtd_x"][0] > 0.
    U = s.download()
    o = Oracle(p)
    o.run(21)
    assert same_bits(U, o.get_U())


def test_engine_with_wide_tiles_matches_oracle():
    """An engine forced to the wide shape, graph replay on, with tiny,
    rectangular and larger grids under several schemes."""
    jobs = [(random_state(3, 29, 9), 3, 5), (random_state(29, 3, 10), 2, 1),
            (P.vortex(47, 3), 3, 5), (P.lake_bumps(33, 2), 2, 7), (P.vortex(20, 1), 4, 3),
            (P.lake_flag1(18, True), 3, 4)]
    eng = swm.Engine()
    eng.set_tile_shape(1)
    eng.set_execution(1, 1)
    ids = [eng.submit(p, k, timescheme=sc, hp_association=swm.HP_BY_NODE) for p, k, sc in jobs]
    while eng.step():
        pass
    assert eng.tile_shape()[0] == 1
    for (p, k, sc), jid in zip(jobs, ids):
        o = Oracle(p, sc, hp_by_node=True)
        for _ in range(k):
            dt, _, _ = o.timestep()
            o.step(dt)
        assert same_bits(eng.fetch(jid), o.get_U()), (p.kind, p.n, p.rows, sc)
