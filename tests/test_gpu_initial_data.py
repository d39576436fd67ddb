"""The reference's initial data made on the device (swm_fill_initial_data and
the engine's initial_condition jobs): ExactFunction cell averages, the
well-balanced state and the bathymetry tables, against the oracle's
restatement (itself bitwise equal to the reference's FillInitialConditions,
InitializeWBVariables and source term, test_oracle_vs_reference.py).

Tolerance: sqrt, + and * are IEEE, so the dam breaks (4, 5) are bitwise; the
other initial conditions go through the device's exp, sin, cos and atan,
which may differ from the host libm in the last bit, so their cell averages
must agree to 1e-14 relative (a few ulps).  Everything after the initial
data is bitwise: steps from the device's data equal the oracle's steps from
the same data."""
from __future__ import annotations

import numpy as np
import pytest

import problems as P
import serve
import swm
from oracle_api import Oracle, oracle_initial_data, oracle_l1_error
from parity_util import describe_diff, same_bits

pytestmark = pytest.mark.gpu

SIZES = [(16, 16), (23, 11), (64, 40)]
RTOL = 1e-14


def oracle_data(c):
    return oracle_initial_data(c.ic, c.nx, c.ny, c.x0, c.x1, c.gravity, c.bathy_flag,
                               c.wellbalanced)


def problem_from_oracle(c):
    U, Uwb, bx, by = oracle_data(c)
    p = P.Problem(c.nx, U, wellbalanced=c.wellbalanced, Uwb=Uwb,
                  dbdx=bx if c.bathy_flag else None, dbdy=by if c.bathy_flag else None,
                  gravity=c.gravity, cfl=c.cfl, ny=c.ny if c.ny != c.nx else 0,
                  x0=c.x0[0], x1=c.x1[0], y0=c.x0[1], y1=c.x1[1])
    assert p.mesh.dx == c.mesh.dx and p.mesh_y.dx == c.mesh_y.dx
    return p


def close(a, b, rtol=RTOL):
    return np.all(np.abs(a - b) <= rtol * np.maximum(1.0, np.abs(b)))


@pytest.mark.parametrize("nx,ny", SIZES)
@pytest.mark.parametrize("ic", [1, 2, 3, 4, 5, 6])
def test_initial_state_matches_oracle(ic, nx, ny):
    c = P.ReferenceSetup(ic, nx, ny, wellbalanced=ic in (2, 3))
    s = swm.Solver.from_setup(c)
    U = s.download()
    U_r = oracle_data(c)[0]
    if ic in (4, 5):
        assert same_bits(U, U_r), describe_diff(U, U_r)
    else:
        assert close(U, U_r), describe_diff(U, U_r)
        print(f"ic {ic} {nx}x{ny}: {np.mean(U == U_r):.3f} of the values bitwise")


@pytest.mark.parametrize("ic", [2, 3, 6])
def test_bathymetry_tables_give_the_oracle_source(ic):
    """The device's Bathymetry_X/_Y tables (flag 1 for ic 2, flag 2 for 3 and
    6) enter the source term S: the device FVTD of the oracle's state with
    device tables against the oracle's FVTD with its own tables."""
    c = P.ReferenceSetup(ic, 40, 24)
    s = swm.Solver.from_setup(c)
    p = problem_from_oracle(c)
    s.upload(p.U0)
    S = s.fvtd()["S"][..., 1:]
    S_r = Oracle(p).fvtd()["S"][..., 1:]
    assert np.abs(S_r).max() > 0
    scale = np.abs(S_r).max()
    assert np.all(np.abs(S - S_r) <= 1e-13 * scale), describe_diff(S, S_r)


def test_well_balanced_state_keeps_the_lake_at_rest():
    """For the lake at rest (ic 2) ExactFunction and ExactFunctionWB are the
    same function, so the device's state equals its well-balanced state:
    Ut(2:3) = 0 exactly, and steps leave the lake bitwise at rest.  (The
    well-balanced state of ic 3 enters the steps test below.)"""
    c = P.ReferenceSetup(2, 32, 20, wellbalanced=True)
    s = swm.Solver.from_setup(c)
    U0 = s.download()
    assert (s.fvtd()["Ut"][..., 1:] == 0).all()
    s.run_steps(3)
    assert same_bits(s.download(), U0)


@pytest.mark.parametrize("ic,wb", [(1, False), (2, True), (3, False), (4, False), (6, False)])
def test_steps_from_device_data(ic, wb):
    """Three mPDeC5 steps from the device's initial data are bitwise the
    oracle's three steps from the same data (downloaded; for the lake at rest
    the well-balanced state equals the state; with bathymetry both use the
    oracle's tables, whose device counterparts are checked above).  Against
    the steps from the
    reference's own data the difference is what the scheme makes of the
    few-ulp differences of the initial data: WENO5 with eps = 1e-24 is
    ill-conditioned next to flat regions, and the reference itself turns
    one-ulp changes of a few initial values into ~1e-9 relative L1 after
    three steps (test_oracle_closed_forms.py); bitwise for the dam break."""
    c = P.ReferenceSetup(ic, 40, 24, wellbalanced=wb)
    s = swm.Solver.from_setup(c)
    U0 = s.download()
    p_dev = problem_from_oracle(c)
    p_dev.U0 = U0
    if wb:
        p_dev.Uwb = U0
    if c.bathy_flag:
        s.set_bathy_grad(p_dev.dbdx, p_dev.dbdy)
        if wb:
            s.init_wb(U0)
    o = Oracle(p_dev)
    o_ref = Oracle(problem_from_oracle(c))
    for _ in range(3):
        dt, _, _ = s.timestep()
        assert dt == o.timestep()[0]
        s.step(dt)
        o.step(dt)
        dt_r, _, _ = o_ref.timestep()
        o_ref.step(dt_r)
    U = s.download()
    assert same_bits(U, o.get_U()), describe_diff(U, o.get_U())
    U_r = o_ref.get_U()
    if ic == 4:
        assert same_bits(U, U_r)
    else:
        assert np.abs(U - U_r).sum() <= 1e-7 * np.abs(U_r).sum(), describe_diff(U, U_r)
        print(f"ic {ic}: relative L1 against the reference's data "
              f"{np.abs(U - U_r).sum() / np.abs(U_r).sum():.2e}")


def test_engine_jobs_from_initial_conditions():
    """Engine jobs submitted as initial conditions (no host data) run as the
    plain solver does from the same device-made data: bitwise."""
    setups = [P.ReferenceSetup(1, 30), P.ReferenceSetup(2, 24, 16, wellbalanced=True),
              P.ReferenceSetup(3, 40, 16, wellbalanced=True), P.ReferenceSetup(4, 32),
              P.ReferenceSetup(5, 20, 28), P.ReferenceSetup(6, 36, 12)]
    eng = swm.Engine()
    ids = [eng.submit_setup(c, 2) for c in setups]
    while eng.step():
        pass
    for c, jid in zip(setups, ids):
        s = swm.Solver.from_setup(c)
        s.run_steps(2)
        info = eng.info(jid)
        assert info["steps_done"] == 2, c.ic
        assert same_bits(eng.fetch(jid), s.download()), c.ic
        # PostProcessing's error of the job at its own time, without a fetch
        err = eng.l1_error(jid, c.ic, c.bathy_flag, c.x0, c.x1, info["t"])
        assert same_bits(err, s.l1_error(c.ic, c.bathy_flag, c.x0, c.x1, info["t"])), c.ic


def test_bad_arguments_are_rejected():
    c = P.ReferenceSetup(1, 16)
    s = swm.Solver.from_setup(c)
    for ic, bf, x0, x1, what in [(0, 0, c.x0, c.x1, "initial condition"),
                                 (7, 0, c.x0, c.x1, "initial condition"),
                                 (1, 3, c.x0, c.x1, "bathymetry flag"),
                                 (1, 0, (0.0, 0.0), (2.0, 3.0), "Mesh_DX")]:
        with pytest.raises(swm.SwmError, match=what):
            s.fill_initial_data(ic, bf, x0, x1)
    w = swm.Solver.from_setup(P.ReferenceSetup(2, 16, wellbalanced=True))
    with pytest.raises(swm.SwmError, match="well-balanced"):
        w.fill_initial_data(1, 0, (0.0, 0.0), (1.0, 1.0))


def test_large_grids():
    """2048^2 lake at rest (ic 2, well-balanced): at rest bitwise after two
    steps; 1024^2 dam break (ic 4) bitwise the oracle's initial data."""
    c = P.ReferenceSetup(2, 2048, wellbalanced=True)
    s = swm.Solver.from_setup(c)
    U0 = s.download()
    s.run_steps(2)
    assert same_bits(s.download(), U0)
    del s
    c = P.ReferenceSetup(4, 1024)
    s = swm.Solver.from_setup(c)
    assert same_bits(s.download(), oracle_data(c)[0])


def test_served_requests_by_initial_condition():
    """The continuous-batching loop with requests that carry no host data
    (the reference's initial conditions, made on the device at admission):
    every result bitwise the sequential plain solver's."""
    reqs = serve.reference_mix(10, 5, rate=0.0, nmin=24, nmax=96, smin=2, smax=6)
    res = serve.serve(reqs, budget_bytes=0)
    got = [r.result for r in reqs]
    states = [r.state for r in reqs]
    seq_reqs = serve.reference_mix(10, 5, rate=0.0, nmin=24, nmax=96, smin=2, smax=6)
    serve.sequential(seq_reqs)
    assert res["jobs"] == 10
    for a, r, st in zip(got, seq_reqs, states):
        assert st == r.state, r.problem.ic
        assert same_bits(a, r.result), (r.problem.ic, r.problem.nx)



@pytest.mark.parametrize("ic", [1, 2, 3, 4, 5, 6])
def test_l1_error_matches_oracle(ic):
    """PostProcessing's L1 error on the device, of a perturbed state at
    t = 0, 0.05 and 1.7: bitwise the oracle's for the dam breaks, within
    1e-13 relative for the others (ExactFunction through the device libm)."""
    c = P.ReferenceSetup(ic, 40, 24)
    U = oracle_data(c)[0] * (1 + 1e-3 * np.random.default_rng(ic).standard_normal((24, 40, 3)))
    s = swm.Solver.from_setup(c)
    s.upload(U)
    for t in (0.0, 0.05, 1.7):
        err = s.l1_error(c.ic, c.bathy_flag, c.x0, c.x1, t)
        ref = oracle_l1_error(U, c.ic, c.x0, c.x1, c.gravity, c.bathy_flag, t)
        if ic in (4, 5):
            assert same_bits(err, ref), (t, err, ref)
        else:
            assert np.all(np.abs(err - ref) <= 1e-13 * np.abs(ref)), (t, err, ref)


def test_vortex_error_after_steps():
    """The reference's convergence measurement: the vortex stepped to time t,
    then its L1 error against the exact (translated) vortex, on the device,
    against the oracle's error of the same state."""
    c = P.ReferenceSetup(1, 48, 48)
    s = swm.Solver.from_setup(c)
    t = 0.0
    for _ in range(4):
        dt, _, _ = s.timestep()
        s.step(dt)
        t = t + dt
    U = s.download()
    err = s.l1_error(1, 0, c.x0, c.x1, t)
    ref = oracle_l1_error(U, 1, c.x0, c.x1, c.gravity, 0, t)
    assert np.all(np.abs(err - ref) <= 1e-13 * np.abs(ref)), (err, ref)
    assert 0 < err[0] < 1e-2


def test_primitives_are_cons_to_prim():
    c = P.ReferenceSetup(3, 40, 24, wellbalanced=True)
    s = swm.Solver.from_setup(c)
    s.run_steps(2)
    U, V = s.download(), s.download_primitive()
    ht = U[..., 0] + 1e-6 / U[..., 0]
    want = np.stack([U[..., 0], U[..., 1] / ht, U[..., 2] / ht], -1)
    assert same_bits(V, want)



def test_async_primitive_download_is_a_snapshot():
    """swm_download_primitive_async copies the primitives of the state at the
    call while the following steps run; a second call reuses the snapshot
    buffer only after the first copy.  Into pinned (hipHostMalloc) and
    pageable host memory."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipHostMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_uint]
    hip.hipHostFree.argtypes = [ctypes.c_void_p]
    pinned = []

    def pinned_array():
        ptr = ctypes.c_void_p()
        assert hip.hipHostMalloc(ctypes.byref(ptr), 40 * 64 * 3 * 8, 0) == 0
        pinned.append(ptr)
        buf = (ctypes.c_double * (40 * 64 * 3)).from_address(ptr.value)
        return np.ctypeslib.as_array(buf).reshape(40, 64, 3)

    c = P.ReferenceSetup(1, 64, 40)
    s = swm.Solver.from_setup(c)
    s.run_steps(1)
    want1 = s.download_primitive()
    V1 = pinned_array()
    s.download_primitive_async(V1)
    s.run_steps(2)
    s.output_wait()
    assert same_bits(V1, want1)
    want2 = s.download_primitive()
    assert not same_bits(want1, want2)
    V2 = pinned_array()
    V3 = np.empty((40, 64, 3))
    s.download_primitive_async(V2)
    s.run_steps(1)
    s.download_primitive_async(V3)  # pageable memory works too
    s.output_wait()
    assert same_bits(V2, want2)
    assert same_bits(V3, s.download_primitive())
    del V1, V2
    for ptr in pinned:
        assert hip.hipHostFree(ptr) == 0



@pytest.mark.parametrize("scheme,by_node", [(5, False), (6, True), (7, True), (7, False),
                                            (1, False), (2, False), (3, False), (4, False)])
@pytest.mark.parametrize("case", ["vortex", "lake", "rect"])
def test_output_state_is_the_writers_v(case, scheme, by_node):
    """swm_keep_output_state / swm_download_output_primitive: the primitives
    of the last FVTD input of the step, bitwise the oracle's V (itself
    bitwise the reference's, test_oracle_vs_reference.py), under every
    execution mode's step bodies (graph replay included), and not the
    primitives of the new state."""
    p = {"vortex": lambda: P.vortex(24, 3), "lake": lambda: P.lake_flag1(20, True, pert=0.05),
         "rect": lambda: P.vortex(17, 5, ny=26)}[case]()
    for graph, margin in ((0, 1), (1, 1), (0, -2)):
        s = swm.Solver.from_problem(p)
        s.set_execution(graph, margin)
        s.set_timescheme(scheme)
        s.set_hp_association(swm.HP_BY_NODE if by_node else swm.HP_AS_REFERENCE)
        with pytest.raises(swm.SwmError, match="kept output state"):
            s.download_output_primitive()
        s.keep_output_state()
        o = Oracle(p, scheme, hp_by_node=by_node)
        for _ in range(2):
            dt, _, _ = s.timestep()
            if not np.isfinite(dt):
                break  # the reference view of mPDeC2 can turn NaN
            s.step(dt)
            o.step(dt)
        V = s.download_output_primitive()
        assert same_bits(V, o.get_V()), (graph, margin)
        if scheme in (2, 3, 4, 5):
            assert not same_bits(V, s.download_primitive())



def test_engine_jobs_keep_the_writers_v():
    """Per-job writer state in a mixed engine: jobs that keep it (of several
    schemes, sharing launches with jobs that do not) give the oracle's V."""
    jobs = [(P.vortex(20, 1), 5, True), (P.lake_flag1(18, True, pert=0.05), 5, False),
            (P.vortex(22, 2), 2, True), (P.vortex(16, 3, ny=24), 7, True),
            (P.lake_bumps(20, 4), 5, True)]
    eng = swm.Engine()
    ids = []
    for p, sc, keep in jobs:
        jid = eng.submit(p, 2, timescheme=sc, hp_association=swm.HP_BY_NODE)
        if keep:
            eng.keep_output_state(jid)
        ids.append(jid)
    while eng.step():
        pass
    for (p, sc, keep), jid in zip(jobs, ids):
        if not keep:
            with pytest.raises(swm.SwmError, match="kept output state"):
                eng.fetch_output_primitive(jid)
            continue
        o = Oracle(p, sc, hp_by_node=True)
        for _ in range(2):
            dt, _, _ = o.timestep()
            o.step(dt)
        assert same_bits(eng.fetch_output_primitive(jid), o.get_V()), (p.kind, sc)


def test_job_bytes_count_every_buffer():
    """A job's reported bytes and the engine's use cover the state slab, the
    bathymetry tables and a kept output state; the budget covers the last."""
    p = P.lake_bumps(40, 1)  # well-balanced, with bathymetry
    fp = swm.job_footprint_bytes(40, 40, True, True)
    eng = swm.Engine()
    jid = eng.submit(p, 1)
    assert eng.info(jid)["bytes"] == fp == eng.bytes_in_use()
    eng.keep_output_state(jid)
    extra = 3 * 40 * 40 * 8
    assert eng.info(jid)["bytes"] == fp + extra == eng.bytes_in_use()
    eng.release(jid)
    assert eng.bytes_in_use() == 0
    tight = swm.Engine(budget_bytes=fp)
    j2 = tight.submit(p, 1)
    with pytest.raises(swm.SwmError, match="does not fit"):
        tight.keep_output_state(j2)
