"""Rectangular grids (nx != ny, so dx != dy on the unit square): the x and y
sweeps, seams and tilings are exercised with different extents, bitwise
against the oracle (itself pinned to the reference on rectangles by
test_oracle_vs_reference.py and the golden fixtures *_rect*)."""
from __future__ import annotations

import numpy as np
import pytest

import problems as P
import swm
from oracle_api import Oracle
from parity_util import FVTD_KEYS, describe_diff, fvtd_view, same_bits

pytestmark = pytest.mark.gpu


def random_state(n, ny, seed, amp=0.3):
    rng = np.random.default_rng(seed)
    U = np.empty((ny, n, 3))
    U[..., 0] = 1.0 + amp * rng.uniform(-1, 1, (ny, n))
    U[..., 1] = U[..., 0] * rng.normal(0, 0.3, (ny, n))
    U[..., 2] = U[..., 0] * rng.normal(0, 0.3, (ny, n))
    return P.Problem(n, U, kind="random", ny=ny)


CASES = {
    "vortex_21x34": lambda: P.vortex(21, 5, ny=34),
    "vortex_40x17": lambda: P.vortex(40, 6, ny=17),
    "random_17x11": lambda: random_state(17, 11, 3),
    "random_3x29": lambda: random_state(3, 29, 4),
    "random_29x3": lambda: random_state(29, 3, 7),
    "random_3x3": lambda: random_state(3, 3, 8),
    "flag1wb_19x26": lambda: P.lake_flag1(19, True, ny=26),
    "flag1_33x16": lambda: P.lake_flag1(33, False, ny=16),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_fvtd_rect(case):
    p = CASES[case]()
    ref = Oracle(p).fvtd()
    got = swm.Solver.from_problem(p).fvtd()
    for k in FVTD_KEYS:
        a, b = fvtd_view(got, k), fvtd_view(ref, k)
        assert same_bits(a, b), f"{case} {k}: {describe_diff(a, b)}"


@pytest.mark.parametrize("case", sorted(CASES))
def test_steps_rect(case):
    p = CASES[case]()
    s = swm.Solver.from_problem(p)
    dts, iters, _ = s.run_steps(3)
    o = Oracle(p)
    for k in range(3):
        dt, _, _ = o.timestep()
        it, _ = o.step(dt)
        assert dt == dts[k], (case, k)
        assert list(iters[k]) == it, (case, k)
    U, U_r = s.download(), o.get_U()
    assert same_bits(U, U_r), f"{case}: {describe_diff(U, U_r)}"


def test_engine_mixes_rectangles():
    jobs = [(CASES["vortex_21x34"](), 3), (CASES["random_17x11"](), 2),
            (CASES["flag1wb_19x26"](), 2), (P.vortex(24, 9), 3)]
    eng = swm.Engine()
    ids = [eng.submit(p, k) for p, k in jobs]
    while eng.step():
        pass
    for (p, k), jid in zip(jobs, ids):
        o = Oracle(p)
        for _ in range(k):
            dt, _, _ = o.timestep()
            o.step(dt)
        U, U_r = eng.fetch(jid), o.get_U()
        assert same_bits(U, U_r), f"{p.kind} {p.n}x{p.rows}: {describe_diff(U, U_r)}"
