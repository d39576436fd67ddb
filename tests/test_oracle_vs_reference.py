"""The C oracle against the reference Fortran itself, built unchanged from its
own sources into oracle/_ref/ (skipped where that build is absent)."""
from __future__ import annotations

import numpy as np
import pytest

import problems as P
from oracle_api import (Oracle, have_swref, have_swref_explicit, oracle_initial_data,
                        oracle_l1_error, run_swref, run_swref_initial_data, run_swref_l1_error)
from parity_util import FVTD_KEYS, describe_diff, same_bits

pytestmark = pytest.mark.skipif(not have_swref(), reason="oracle/_ref not built")


def random_state(n, seed, amp=0.2, ny=0):
    """A rough but positive random state: exercises the nonlinear WENO weights
    and the limiter far from the smooth regime (n x ny cells)."""
    rng = np.random.default_rng(seed)
    r = ny or n
    U = np.empty((r, n, 3))
    U[..., 0] = 1.0 + amp * rng.uniform(-1, 1, (r, n))
    U[..., 1] = U[..., 0] * rng.normal(0, 0.3, (r, n))
    U[..., 2] = U[..., 0] * rng.normal(0, 0.3, (r, n))
    return P.Problem(n, U, kind="random", ny=ny)


CASES = {
    "vortex21": lambda: P.vortex(21, 31),
    "random13": lambda: random_state(13, 1),
    "random16_rough": lambda: random_state(16, 2, amp=0.9),
    "lakeflag1_wb14": lambda: P.lake_flag1(14, True),
    "dam15": lambda: P.dam_break(15),
    "dam15_mindepth": lambda: P.dam_break(15, dry=1e-6),
    "vortex_rect21x34": lambda: P.vortex(21, 5, ny=34),
    "random_rect17x11": lambda: random_state(17, 3, ny=11),
    "lakeflag1_wb_rect19x26": lambda: P.lake_flag1(19, True, ny=26),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_fvtd_bitwise(case):
    p = CASES[case]()
    ref = run_swref(p, 1)
    got = Oracle(p).fvtd()
    for k in FVTD_KEYS:
        a, b = got[k], ref[k]
        if k in ("Ut", "S"):
            a, b = a[..., 1:], b[..., 1:]
        assert same_bits(a, b), f"{case} {k}: {describe_diff(a, b)}"


@pytest.mark.parametrize("case,steps", [("vortex21", 5), ("random13", 3), ("lakeflag1_wb14", 4),
                                        ("dam15_mindepth", 3), ("dam15", 2),
                                        ("vortex_rect21x34", 3), ("random_rect17x11", 2),
                                        ("lakeflag1_wb_rect19x26", 2)])
def test_steps_bitwise(case, steps):
    p = CASES[case]()
    hist_r, U_r = run_swref(p, 2, steps=steps)
    o = Oracle(p)
    U = None
    for k, row in enumerate(hist_r):
        dt, lx, ly = o.timestep()
        assert same_bits([dt, lx, ly], row), f"step {k}"
        if not np.isfinite(dt):
            break
        o.step(dt)
        U = o.get_U()
    assert same_bits(U, U_r), describe_diff(U, U_r)


def test_jacobi_bitwise_random_matrices():
    rng = np.random.default_rng(3)
    for n in (5, 9):
        N = n * n
        mass = -rng.uniform(0, 0.24, 5 * N)
        mass[2::5] = 0.0
        diag = 1.0 + rng.uniform(0, 0.1, N)
        rhs = rng.uniform(0, 2, N)
        p = P.Problem(n, np.zeros((n, n, 3)))
        ref = run_swref(p, 4, jac=(mass, diag, rhs))
        sol, k = Oracle(p).jacobi(mass, diag, rhs)
        assert k > 5
        assert same_bits(sol, ref)


# The two-node modified-Patankar schemes (timescheme 6 mPEuler, 7 mPDeC2).  In
# the reference their depths come through the sequence association of
# hp(MSteps,nx,ny) with Up(:,1,:,:) (first extent 4), so states turn NaN
# within a few steps; the history is compared up to the first non-finite dt.
@pytest.mark.parametrize("scheme", [6, 7])
@pytest.mark.parametrize("case,steps", [("vortex21", 4), ("random13", 3), ("lakeflag1_wb14", 4),
                                        ("dam15_mindepth", 3), ("vortex_rect21x34", 3),
                                        ("random_rect17x11", 3)])
def test_two_node_schemes_bitwise(case, steps, scheme):
    p = CASES[case]()
    hist_r, U_r = run_swref(p, 2, steps=steps, scheme=scheme)
    o = Oracle(p, scheme)
    U = None
    for k, row in enumerate(hist_r):
        dt, lx, ly = o.timestep()
        assert same_bits([dt, lx, ly], row), f"step {k}"
        if not np.isfinite(dt):
            break
        o.step(dt)
        U = o.get_U()
    assert same_bits(U, U_r), describe_diff(U, U_r)


@pytest.mark.parametrize("scheme", [6, 7])
def test_reference_two_node_schemes_lose_mass(scheme):
    """What the sequence association does in the reference: the first step's
    right-hand side is h^n of cell k/2 at even cells k and 0 at odd ones; the
    matrix conserves mass, so the new total depth is the sum of that
    right-hand side -- the first half of the cells, about half the mass."""
    p = CASES["vortex21"]()
    _, U_r = run_swref(p, 2, steps=1, scheme=scheme)
    h0 = p.U0[..., 0].ravel()
    n = h0.size
    expect = sum(h0[k // 2] for k in range(0, n, 2))
    assert abs(U_r[..., 0].sum() - expect) < 1e-11 * n
    assert abs(U_r[..., 0].sum() - h0.sum()) > 0.4 * h0.sum()


def test_by_node_mpeuler_momentum_matches_reference():
    """The momentum update of mPEuler's single correction does not read hp,
    so the by-node variant's first-step momenta are the reference's."""
    p = CASES["vortex21"]()
    _, U_r = run_swref(p, 2, steps=1, scheme=6)
    o = Oracle(p, 6, hp_by_node=True)
    dt, _, _ = o.timestep()
    o.step(dt)
    U = o.get_U()
    assert same_bits(U[..., 1:], U_r[..., 1:])
    assert not same_bits(U[..., 0], U_r[..., 0])


# The explicit schemes of the build without PATANKAR (timescheme 1 ForwardEuler,
# 2 SSPRK4, 3 RK65, 4 DeC5), against that build of the reference.
explicit = pytest.mark.skipif(not have_swref_explicit(), reason="oracle/_ref/swref_e not built")


@explicit
@pytest.mark.parametrize("case", ["vortex21", "random16_rough", "lakeflag1_wb14", "dam15_mindepth",
                                  "vortex_rect21x34", "lakeflag1_wb_rect19x26"])
def test_fvtd_explicit_build_bitwise(case):
    """That build's FVTimeDerivative: Ut(1:3), with Ut(1) from the mass
    fluxes (src/finitevolume.f90:529-533), and the same FX, FY, S."""
    p = CASES[case]()
    ref = run_swref(p, 1, scheme=1)
    got = Oracle(p).fvtd()
    for k in ("Ut", "FX", "FY", "S"):
        assert same_bits(got[k], ref[k]), f"{case} {k}: {describe_diff(got[k], ref[k])}"


@explicit
@pytest.mark.parametrize("scheme", [1, 2, 3, 4])
@pytest.mark.parametrize("case,steps", [("vortex21", 3), ("random13", 3), ("lakeflag1_wb14", 3),
                                        ("dam15_mindepth", 2), ("vortex_rect21x34", 2),
                                        ("random_rect17x11", 3)])
def test_explicit_schemes_bitwise(case, steps, scheme):
    p = CASES[case]()
    hist_r, U_r = run_swref(p, 2, steps=steps, scheme=scheme)
    o = Oracle(p, scheme)
    U = None
    for k, row in enumerate(hist_r):
        dt, lx, ly = o.timestep()
        assert same_bits([dt, lx, ly], row), f"step {k}"
        if not np.isfinite(dt):
            break
        o.step(dt)
        U = o.get_U()
    assert same_bits(U, U_r), describe_diff(U, U_r)


@pytest.mark.parametrize("scheme", [1, 3, 4])
def test_patankar_build_explicit_schemes_freeze_the_depth(scheme):
    """Why the explicit schemes follow the build without PATANKAR: in the
    PATANKAR build nothing writes Ut(1) (src/finitevolume.f90:525-527 set
    Ut(2:3) only), so these schemes leave h bitwise unchanged there."""
    p = CASES["vortex21"]()
    _, U = run_swref(p, 2, steps=1, scheme=scheme, patankar=True)
    assert same_bits(U[..., 0], p.U0[..., 0])
    assert not same_bits(U[..., 1], p.U0[..., 1])


# ---- ReconstructionFix != Reconstruction: the shock-indicator path ----------
# The reference fixes ReconstructionFix = Reconstruction (src/parameters.f90:142);
# the driver of oracle/_ref sets it from its input header instead.

FIX_CASES = ["vortex21", "random13", "random16_rough", "lakeflag1_wb14", "dam15_mindepth",
             "vortex_rect21x34", "random_rect17x11", "lakeflag1_wb_rect19x26"]


def jameson_np(hm, h0, hp):
    ind = np.abs(hm + hp - 2.0 * h0)
    with np.errstate(invalid="ignore", divide="ignore"):
        ind = ind / (np.abs(hm) + np.abs(hp) + np.abs(2.0 * h0))
    return ind > 0.5e-3


def indicator_readings(h):
    """Flags of cells 0..n+1 along the last axis of the periodic depths h:
    the reference's (each cell's verdict overwrites Ind(ii-1:ii+1) in loop
    order, src/shocksindicator.f90:50-55) and an OR over the three writers."""
    n = h.shape[-1]
    shock = jameson_np(np.roll(h, 1, -1), h, np.roll(h, -1, -1))  # cells 1..n
    idx = np.arange(n + 2)
    last = shock[..., np.minimum(idx + 1, n) - 1]
    ors = np.zeros(h.shape[:-1] + (n + 2,), bool)
    for w in range(1, n + 1):  # writer w sets cells w-1..w+1
        ors[..., w - 1:w + 2] |= shock[..., w - 1:w]
    return last, ors


@pytest.mark.parametrize("recfix", [1, 2])
@pytest.mark.parametrize("case", FIX_CASES)
def test_reconstruction_fix_fvtd_bitwise(case, recfix):
    """One FVTimeDerivative with ReconstructionFix 1 (first order) or 2
    (MUSCL) on the flagged cells: bitwise, and different from the WENO5-only
    result (the indicator flags cells in every case)."""
    p = CASES[case]()
    ref = run_swref(p, 1, recfix=recfix)
    assert not same_bits(ref["FX"], run_swref(p, 1)["FX"]), "no cell flagged"
    got = Oracle(p, recfix=recfix).fvtd()
    for k in FVTD_KEYS:
        a, b = got[k], ref[k]
        if k in ("Ut", "S"):
            a, b = a[..., 1:], b[..., 1:]
        assert same_bits(a, b), f"{case} fix {recfix} {k}: {describe_diff(a, b)}"


def test_fix_cases_separate_the_indicator_readings():
    """On the smooth cases the reference's overwrite reading of the indicator
    and an OR of the three writers flag different cells along x and along y,
    so the bitwise comparisons above and the device tests would catch an
    implementation that ORs.  (The random states are rough enough that every
    cell is flagged: they exercise the fix arithmetic everywhere.)"""
    for case in FIX_CASES:
        h = CASES[case]().U0[..., 0]
        for hh in (h, h.T):
            last, ors = indicator_readings(hh)
            if not case.startswith("random"):
                assert (last != ors).any(), case
                assert 0.05 < last.mean() < 0.95, case


@pytest.mark.parametrize("recfix", [1, 2])
@pytest.mark.parametrize("case,steps", [("vortex21", 3), ("random13", 2), ("lakeflag1_wb14", 3),
                                        ("dam15_mindepth", 2), ("vortex_rect21x34", 2),
                                        ("random_rect17x11", 2)])
def test_reconstruction_fix_steps_bitwise(case, steps, recfix):
    """mPDeC5 steps with the fix (the well-balanced state built under it too)."""
    p = CASES[case]()
    hist_r, U_r = run_swref(p, 2, steps=steps, recfix=recfix)
    o = Oracle(p, recfix=recfix)
    for k, row in enumerate(hist_r):
        dt, lx, ly = o.timestep()
        assert same_bits([dt, lx, ly], row), f"step {k}"
        if not np.isfinite(dt):
            break
        o.step(dt)
    U = o.get_U()
    assert same_bits(U, U_r), f"{case} fix {recfix}: {describe_diff(U, U_r)}"


@explicit
@pytest.mark.parametrize("scheme", [2, 4])
@pytest.mark.parametrize("recfix", [1, 2])
@pytest.mark.parametrize("case", ["vortex21", "random16_rough", "lakeflag1_wb14"])
def test_reconstruction_fix_explicit_schemes_bitwise(case, recfix, scheme):
    p = CASES[case]()
    hist_r, U_r = run_swref(p, 2, steps=2, scheme=scheme, recfix=recfix)
    o = Oracle(p, scheme, recfix=recfix)
    for k, row in enumerate(hist_r):
        dt, lx, ly = o.timestep()
        assert same_bits([dt, lx, ly], row), f"step {k}"
        o.step(dt)
    U = o.get_U()
    assert same_bits(U, U_r), f"{case} fix {recfix} scheme {scheme}: {describe_diff(U, U_r)}"


def test_reference_fix_does_not_conserve_mass():
    """The reference itself, built with ReconstructionFix 1 or 2, loses mass
    on a dam break (through the periodic seam, whose two face copies see
    different flags; test_oracle_closed_forms pins the mechanism), and the
    oracle reproduces the loss bitwise (test_reconstruction_fix_steps_bitwise)."""
    p = P.dam_break(20, dry=0.1)
    m0 = p.U0[..., 0].sum()
    _, U4 = run_swref(p, 2, steps=3)
    assert abs(U4[..., 0].sum() - m0) <= 1e-12 * m0
    for recfix in (1, 2):
        _, U = run_swref(p, 2, steps=3, recfix=recfix)
        assert abs(U[..., 0].sum() - m0) > 1e-10 * m0


# ---- initial data: ExactFunction, ExactFunctionWB, Bathymetry_X/_Y ----------
# The reference's own setups (src/parameters.f90:62-125): domain, gravity and
# BathymetryFlag per InitialCondition.
IC_SETUPS = {1: ((0., 0.), (3., 3.), 9.81, 0), 2: ((0., 0.), (1., 1.), 9.81, 1),
             3: ((-5., -2.), (5., 2.), 9.8, 2), 4: ((0., 0.), (40., 40.), 9.8, 0),
             5: ((0., 0.), (50., 50.), 9.8, 0), 6: ((-5., -2.), (5., 2.), 9.8, 2)}


@pytest.mark.parametrize("ic", sorted(IC_SETUPS))
@pytest.mark.parametrize("nx,ny", [(16, 16), (23, 11), (64, 40)])
def test_initial_data_bitwise(ic, nx, ny):
    """FillInitialConditions (and InitializeWBVariables' averages for the
    lakes 2 and 3) of the reference against the oracle's restatement."""
    x0, x1, g, bf = IC_SETUPS[ic]
    wb = ic in (2, 3)
    U, Uwb = run_swref_initial_data(ic, nx, ny, x0, x1, g, bf, wb)
    Uo, Uwbo, _, _ = oracle_initial_data(ic, nx, ny, x0, x1, g, bf, wb)
    assert same_bits(Uo, U), describe_diff(Uo, U)
    if wb:
        assert same_bits(Uwbo, Uwb), describe_diff(Uwbo, Uwb)


@pytest.mark.parametrize("bflag,x0,x1", [(1, 0.0, 1.0), (2, -2.0, 2.0), (2, -1.5, 0.5)])
def test_bathymetry_tables_give_the_reference_source(bflag, x0, x1):
    """The oracle's Bathymetry_X/_Y tables at the Gauss points, fed to the
    FVTD, give the source term S the reference computes with its own
    Bathymetry_X/_Y (bitwise), on a rough state over the domain."""
    n = 20
    _, _, bx, by = oracle_initial_data(2, n, n, (x0, x0), (x1, x1), 9.81, bflag, False)
    rng = np.random.default_rng(bflag)
    U = np.empty((n, n, 3))
    U[..., 0] = 1.0 + 0.3 * rng.uniform(-1, 1, (n, n))
    U[..., 1] = U[..., 0] * rng.normal(0, 0.3, (n, n))
    U[..., 2] = U[..., 0] * rng.normal(0, 0.3, (n, n))
    p = P.Problem(n, U, kind="random", dbdx=bx, dbdy=by, x0=x0, x1=x1)
    ref = run_swref(p, 1, bathy_flag=bflag)
    got = Oracle(p).fvtd()
    assert np.abs(ref["S"][..., 1:]).max() > 0
    assert same_bits(got["S"][..., 1:], ref["S"][..., 1:]), describe_diff(got["S"], ref["S"])
    assert same_bits(got["Ut"][..., 1:], ref["Ut"][..., 1:])



@pytest.mark.parametrize("ic", sorted(IC_SETUPS))
def test_l1_error_bitwise(ic):
    """PostProcessing's L1 error (read back from the ErrorL1 file it writes)
    against the oracle's, of a perturbed initial state at t = 0, 0.05 and
    1.7 (the vortex wraps around the periodic domain)."""
    x0, x1, g, bf = IC_SETUPS[ic]
    U = oracle_initial_data(ic, 24, 14, x0, x1, g, bf, False)[0]
    U = U * (1 + 1e-3 * np.random.default_rng(ic).standard_normal(U.shape))
    for t in (0.0, 0.05, 1.7):
        ref = run_swref_l1_error(U, ic, x0, x1, g, bf, t)
        got = oracle_l1_error(U, ic, x0, x1, g, bf, t)
        assert same_bits(got, ref), (t, got, ref)



@pytest.mark.parametrize("scheme", [5, 6, 7])
@pytest.mark.parametrize("case", ["vortex21", "lakeflag1_wb14", "vortex_rect21x34"])
def test_writer_state_bitwise(case, scheme):
    """The module array V after two steps -- what WriteSolutionToDisk prints,
    the primitives of the last FVTD input -- against the oracle's."""
    p = CASES[case]()
    V_ref = run_swref(p, 8, steps=2, scheme=scheme)
    o = Oracle(p, scheme)
    for _ in range(2):
        dt, _, _ = o.timestep()
        if not np.isfinite(dt):
            break  # the reference's loop stops there too
        o.step(dt)
    assert same_bits(o.get_V(), V_ref)



@explicit
@pytest.mark.parametrize("scheme", [1, 2, 3, 4])
@pytest.mark.parametrize("case", ["vortex21", "lakeflag1_wb14"])
def test_writer_state_explicit_bitwise(case, scheme):
    """The same for the explicit schemes of the build without PATANKAR: V is
    the last stage state's primitives."""
    p = CASES[case]()
    V_ref = run_swref(p, 8, steps=2, scheme=scheme)
    o = Oracle(p, scheme)
    for _ in range(2):
        dt, _, _ = o.timestep()
        o.step(dt)
    assert same_bits(o.get_V(), V_ref)


def test_reference_dam_break_reaches_a_non_finite_time_step():
    """The reference's own circular dam break 1 (InitialCondition 4, its
    domain [0,40]^2, 64^2 cells) under mPDeC5 reaches a non-finite TimeStep
    at step 40 (index 39); the oracle follows it dt for dt.  Served requests
    of this initial condition therefore end FAILED, as in the reference."""
    x0, x1, g, bf = IC_SETUPS[4]
    U = oracle_initial_data(4, 64, 64, x0, x1, g, bf, False)[0]
    p = P.Problem(64, U, gravity=g, x0=x0[0], x1=x1[0])
    hist, _ = run_swref(p, 2, steps=45)
    o = Oracle(p)
    for k, row in enumerate(hist):
        dt, lx, ly = o.timestep()
        assert same_bits([dt, lx, ly], row), k
        if not np.isfinite(dt):
            break
        o.step(dt)
    assert k == 39 and not np.isfinite(hist[-1][0])
