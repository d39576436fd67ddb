"""ReconstructionFix != Reconstruction on the device: the Jameson shock
indicator (with the reference's overwrite reading of its flags) and the first
order / MUSCL values of the flagged cells inside the face kernels, bitwise
against the oracle (itself pinned to the reference built with the same
ReconstructionFix by test_oracle_vs_reference.py and the golden fixtures
*_fix1, *_fix2), through FVTD, the time schemes, both tile shapes, the engine
and the well-balanced state."""
from __future__ import annotations

import numpy as np
import pytest

import problems as P
import swm
from oracle_api import Oracle
from parity_util import FVTD_KEYS, describe_diff, same_bits

pytestmark = pytest.mark.gpu


def random_state(n, ny, seed, amp=0.3):
    rng = np.random.default_rng(seed)
    U = np.empty((ny, n, 3))
    U[..., 0] = 1.0 + amp * rng.uniform(-1, 1, (ny, n))
    U[..., 1] = U[..., 0] * rng.normal(0, 0.3, (ny, n))
    U[..., 2] = U[..., 0] * rng.normal(0, 0.3, (ny, n))
    return P.Problem(n, U, kind="random", ny=ny)


CASES = {
    "vortex24": lambda: P.vortex(24, 3),
    "vortex_rect21x34": lambda: P.vortex(21, 5, ny=34),
    "lake_bumps20": lambda: P.lake_bumps(20, 4, amp=0.05),
    "flag1wb16": lambda: P.lake_flag1(16, True, pert=0.05),
    "shallowdam18": lambda: P.dam_break(18, dry=0.1),
    "random_rect17x11": lambda: random_state(17, 11, 5),
    "random_3x29": lambda: random_state(3, 29, 6, amp=0.01),
    "vortex40x19": lambda: P.vortex(40, 7, ny=19),
}


@pytest.mark.parametrize("recfix", [1, 2])
@pytest.mark.parametrize("case", sorted(CASES))
def test_fvtd_matches_oracle(case, recfix):
    p = CASES[case]()
    s = swm.Solver.from_problem(p, recfix=recfix)
    got = s.fvtd()
    ref = Oracle(p, recfix=recfix).fvtd()
    for k in FVTD_KEYS:
        a, b = got[k], ref[k]
        if k in ("Ut", "S"):
            a, b = a[..., 1:], b[..., 1:]
        assert same_bits(a, b), f"{case} fix {recfix} {k}: {describe_diff(a, b)}"


@pytest.mark.parametrize("scheme,by_node", [(5, False), (2, False), (4, False), (7, True)])
@pytest.mark.parametrize("recfix", [1, 2])
@pytest.mark.parametrize("case", sorted(CASES))
def test_steps_match_oracle(case, recfix, scheme, by_node):
    p = CASES[case]()
    s = swm.Solver.from_problem(p, recfix=recfix)
    s.set_timescheme(scheme)
    s.set_hp_association(swm.HP_BY_NODE if by_node else swm.HP_AS_REFERENCE)
    o = Oracle(p, scheme, hp_by_node=by_node, recfix=recfix)
    for k in range(2):
        dt, lx, ly = s.timestep()
        assert same_bits([dt, lx, ly], list(o.timestep())), (case, k)
        if not np.isfinite(dt):
            break
        s.step(dt)
        o.step(dt)
        U, U_r = s.download(), o.get_U()
        assert same_bits(U, U_r), f"{case} fix {recfix} scheme {scheme} step {k}: " \
                                  f"{describe_diff(U, U_r)}"


@pytest.mark.parametrize("recfix", [1, 2])
def test_wide_tiles_give_the_same_bits(recfix):
    """Tile shape 1 (32x16 x-face tiles, 16x32 y-face tiles) moves the tile
    seams; the flags near them read the halo.  Same bits as the oracle."""
    for case in ("vortex40x19", "vortex_rect21x34", "random_rect17x11"):
        p = CASES[case]()
        s = swm.Solver.from_problem(p, recfix=recfix)
        s.set_tile_shape(1)
        o = Oracle(p, recfix=recfix)
        for _ in range(2):
            dt, _, _ = s.timestep()
            s.step(dt)
            o.step(dt)
        assert same_bits(s.download(), o.get_U()), case


def test_engine_mixes_reconstruction_fixes():
    """Jobs of one scheme with ReconstructionFix 4, 1 and 2 share launches;
    each job's result is its own."""
    jobs = [(P.vortex(20, 1), 4), (P.vortex(20, 1), 1), (P.vortex(20, 1), 2),
            (P.lake_flag1(18, True, pert=0.05), 1), (P.lake_flag1(18, True, pert=0.05), 2),
            (P.dam_break(22, dry=0.1), 2), (random_state(13, 9, 3), 1)]
    eng = swm.Engine()
    ids = [eng.submit(p, 2, recfix=rf) for p, rf in jobs]
    while eng.step():
        pass
    for (p, rf), jid in zip(jobs, ids):
        o = Oracle(p, recfix=rf)
        for _ in range(2):
            dt, _, _ = o.timestep()
            o.step(dt)
        assert eng.info(jid)["steps_done"] == 2
        assert same_bits(eng.fetch(jid), o.get_U()), (p.kind, rf)
    assert not same_bits(eng.fetch(ids[0]), eng.fetch(ids[1]))


def test_setting_discards_the_well_balanced_state():
    """The well-balanced reference is an FVTD under the setting in force: a
    change after swm_init_wb discards it (stepping fails until init_wb runs
    again), and the rebuilt state gives the oracle's bits."""
    p = P.lake_flag1(16, True, pert=0.05)
    s = swm.Solver.from_problem(p)
    s.set_reconstruction_fix(2)
    dt, _, _ = s.timestep()
    with pytest.raises(swm.SwmError, match="well-balanced"):
        s.step(dt)
    s.init_wb(p.Uwb)
    o = Oracle(p, recfix=2)
    dt, _, _ = s.timestep()
    s.step(dt)
    o.step(dt)
    assert same_bits(s.download(), o.get_U())
    s.set_reconstruction_fix(2)  # unchanged: the state stays
    dt, _, _ = s.timestep()
    s.step(dt)


@pytest.mark.parametrize("bad", [0, 3, 5, -1])
def test_bad_values_are_rejected(bad):
    s = swm.Solver.from_problem(P.vortex(12, 1))
    with pytest.raises(swm.SwmError, match="reconstruction fix"):
        s.set_reconstruction_fix(bad)
    eng = swm.Engine()
    if bad == 0:
        assert eng.submit(P.vortex(12, 1), 1, recfix=0) is not None  # 0: the default
    else:
        with pytest.raises(swm.SwmError, match="reconstruction_fix"):
            eng.submit(P.vortex(12, 1), 1, recfix=bad)


def test_dry_bed_first_order_nan_pattern():
    """Dry-bed dam break under fix 1 (first order has no sign-of-NaN
    dependence): dt sequence and NaN positions as the oracle's."""
    p = P.dam_break(20)
    s = swm.Solver.from_problem(p, recfix=1)
    o = Oracle(p, recfix=1)
    for k in range(3):
        dt, lx, ly = s.timestep()
        assert same_bits([dt, lx, ly], list(o.timestep())), k
        if not np.isfinite(dt):
            break
        s.step(dt)
        o.step(dt)
        assert same_bits(s.download(), o.get_U()), k


@pytest.mark.parametrize("recfix", [1, 2])
def test_large_grid_positivity_and_lake(recfix):
    """1024^2: depths stay non-negative over two mPDeC5 steps of a dam break
    (mass is not a property here: under the fix the reference leaks mass
    through the periodic seam, whose two face copies see different flags);
    the well-balanced lake at rest stays bitwise at rest (the flags of U and
    of the well-balanced state agree, so FVTD(U) - FVTD(Uwb) is exactly 0)."""
    p = P.dam_break(1024, dry=0.1)
    s = swm.Solver.from_problem(p, recfix=recfix)
    s.run_steps(2)
    U = s.download()
    assert np.isfinite(U).all() and U[..., 0].min() >= 0.0
    lake = P.lake_bumps(1024, 11, amp=0.0)
    s = swm.Solver.from_problem(lake, recfix=recfix)
    s.run_steps(2)
    assert same_bits(s.download(), lake.U0)
