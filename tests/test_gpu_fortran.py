"""The Fortran side of the boundary: sw-mpdec_amd/fortran/swm_fdriver (built
by build.py from swm_bridge.f90 + swm_fdriver.f90 against libswm.so) runs the
reference-shaped time loop (dt = TimeStep(), one mPDeC5 step) through
ISO_C_BINDING; its dt sequence and final state must equal the oracle's
bitwise."""
from __future__ import annotations

import subprocess

import numpy as np
import pytest

import problems as P
import swm
from conftest import ROOT
from oracle_api import Oracle
from parity_util import describe_diff, same_bits

pytestmark = pytest.mark.gpu

FDRIVER = ROOT / "sw-mpdec_amd" / "lib" / "swm_fdriver"


def with_steps(p, k):
    p.steps = k
    return p


def write_input(path, p, loop=0, scheme=5, hp_by_node=False, recfix=0):
    m = p.mesh
    with open(path, "wb") as f:
        np.array([p.n, p.rows, p.steps, int(p.wellbalanced), int(p.dbdx is not None), loop,
                  scheme, int(hp_by_node), recfix, 0, 0], dtype=np.int32).tofile(f)
        np.array([m.dx, p.mesh_y.dx, p.gravity, p.cfl, *m.wbnd, 0., 0., 0., 0.],
                 dtype=np.float64).tofile(f)
        np.ascontiguousarray(p.U0, dtype=np.float64).tofile(f)
        if p.wellbalanced:
            np.ascontiguousarray(p.Uwb, dtype=np.float64).tofile(f)
        if p.dbdx is not None:
            np.ascontiguousarray(p.dbdx, dtype=np.float64).tofile(f)
            np.ascontiguousarray(p.dbdy, dtype=np.float64).tofile(f)


def read_output(path, n, rows):
    raw = open(path, "rb").read()
    nsteps, nonconv = np.frombuffer(raw[:8], dtype=np.int32)
    dts = np.frombuffer(raw[8:8 + 8 * nsteps], dtype=np.float64)
    U = np.frombuffer(raw[8 + 8 * nsteps:], dtype=np.float64).reshape(rows, n, 3)
    return dts, U, int(nonconv)


@pytest.mark.parametrize("make", [lambda: P.vortex(24, 3, steps=3),
                                  lambda: P.lake_bumps(20, 4, steps=2),
                                  lambda: with_steps(P.lake_flag1(16, True), 2),
                                  lambda: P.vortex(19, 8, steps=2, ny=30)],
                         ids=["vortex24", "lake_bumps20", "lake_flag1_wb16", "vortex_rect19x30"])
def test_fortran_driver_matches_oracle(tmp_path, make):
    if not FDRIVER.exists():
        pytest.skip("Fortran driver not built (no Fortran compiler at build time)")
    p = make()
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(fin, p)
    r = subprocess.run([str(FDRIVER), str(fin), str(fout)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    dts, U, nonconv = read_output(fout, p.n, p.rows)
    o = Oracle(p)
    dts_r = []
    for _ in range(p.steps):
        dt, _, _ = o.timestep()
        o.step(dt)
        dts_r.append(dt)
    assert same_bits(dts, np.array(dts_r))
    U_r = o.get_U()
    assert same_bits(U, U_r), describe_diff(U, U_r)
    assert nonconv == 0


def _min_r(a, b):
    return a if a < b else b


@pytest.mark.parametrize("make", [lambda: P.vortex(28, 6, steps=5),
                                  lambda: with_steps(P.lake_flag1(20, True), 4)],
                         ids=["vortex28", "lake_flag1_wb20"])
def test_fortran_output_loop_with_advance(tmp_path, make):
    """The driver's loop = 1: the reference's output loop (tEnd = (nsteps -
    1/2) dt_min(0), two output intervals) with one swm_advance call per
    interval, against the same loop on the oracle."""
    if not FDRIVER.exists():
        pytest.skip("Fortran driver not built (no Fortran compiler at build time)")
    p = make()
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(fin, p, loop=1)
    r = subprocess.run([str(FDRIVER), str(fin), str(fout)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    dts, U, nonconv = read_output(fout, p.n, p.rows)
    o = Oracle(p)
    dt0, _, _ = o.timestep()
    t_end = (float(p.steps) - 0.5) * dt0
    dt_an = t_end / 2.0
    t, t_an, ref = 0.0, dt_an, []
    while True:
        dt_min, _, _ = o.timestep()
        dt = _min_r(_min_r(dt_min, t_an - t), _min_r(dt_min, t_end - t))
        o.step(dt)
        ref.append(dt)
        t = t + dt
        if abs(t_an - t) < 1.0e-10:
            t_an = t_an + dt_an
            if t_an > t_end:
                t_an = t_end
        if abs(t - t_end) < 1.0e-10:
            break
    assert same_bits(dts, np.array(ref))
    U_r = o.get_U()
    assert same_bits(U, U_r), describe_diff(U, U_r)
    assert nonconv == 0


@pytest.mark.parametrize("scheme,by_node,make",
                         [(1, False, lambda: P.vortex(20, 3, steps=3)),
                          (2, False, lambda: P.lake_bumps(18, 4, steps=2)),
                          (3, False, lambda: with_steps(P.lake_flag1(16, True), 2)),
                          (4, False, lambda: P.vortex(17, 5, steps=2, ny=23)),
                          (6, False, lambda: P.vortex(24, 3, steps=2)),
                          (7, False, lambda: P.vortex(20, 5, steps=1)),
                          (6, True, lambda: P.vortex(24, 3, steps=3)),
                          (7, True, lambda: with_steps(P.lake_flag1(16, True), 3))],
                         ids=["euler", "ssprk4", "rk65", "dec5", "mpeuler_ref", "mpdec2_ref",
                              "mpeuler_by_node", "mpdec2_by_node"])
def test_fortran_driver_other_schemes(tmp_path, scheme, by_node, make):
    """The driver's SELECT CASE(timescheme) with the explicit schemes 1-4
    (swm_step_forward_euler ... swm_step_dec5), 6 (swm_step_mpeuler) and 7
    (swm_step_mpdec2) in both depth views, against the oracle (the
    reference's view only for the steps before its states turn NaN)."""
    if not FDRIVER.exists():
        pytest.skip("Fortran driver not built (no Fortran compiler at build time)")
    p = make()
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(fin, p, scheme=scheme, hp_by_node=by_node)
    r = subprocess.run([str(FDRIVER), str(fin), str(fout)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    dts, U, _ = read_output(fout, p.n, p.rows)
    o = Oracle(p, scheme, hp_by_node=by_node)
    dts_r = []
    for _ in range(p.steps):
        dt, _, _ = o.timestep()
        o.step(dt)
        dts_r.append(dt)
    assert same_bits(dts, np.array(dts_r))
    U_r = o.get_U()
    assert same_bits(U, U_r), describe_diff(U, U_r)


@pytest.mark.parametrize("recfix,scheme,make",
                         [(1, 5, lambda: with_steps(P.lake_flag1(16, True, pert=0.05), 2)),
                          (2, 5, lambda: P.vortex(20, 3, steps=2)),
                          (2, 2, lambda: P.dam_break(18, steps=2, dry=0.1))],
                         ids=["first_order_wb", "muscl", "muscl_ssprk4"])
def test_fortran_driver_reconstruction_fix(tmp_path, recfix, scheme, make):
    """The driver's header reconstruction_fix: swm_set_reconstruction_fix
    before the well-balanced state is built, then the time loop, against the
    oracle with the same ReconstructionFix."""
    if not FDRIVER.exists():
        pytest.skip("Fortran driver not built (no Fortran compiler at build time)")
    p = make()
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(fin, p, scheme=scheme, recfix=recfix)
    r = subprocess.run([str(FDRIVER), str(fin), str(fout)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    dts, U, _ = read_output(fout, p.n, p.rows)
    o = Oracle(p, scheme, recfix=recfix)
    dts_r = []
    for _ in range(p.steps):
        dt, _, _ = o.timestep()
        o.step(dt)
        dts_r.append(dt)
    assert same_bits(dts, np.array(dts_r))
    U_r = o.get_U()
    assert same_bits(U, U_r), describe_diff(U, U_r)



@pytest.mark.parametrize("ic,wb,scheme", [(1, False, 5), (2, True, 2), (3, True, 5), (4, False, 5)])
def test_fortran_driver_initial_data_and_l1_error(tmp_path, ic, wb, scheme):
    """The driver's initial_condition header: swm_fill_initial_data instead of
    host data, the time loop, then PostProcessing's L1 error by swm_l1_error
    at the final time and the writer's V -- all through ISO_C_BINDING --
    against the same calls from Python (bitwise).  (The wet-dry lake 3 runs mPDeC5: the
    explicit schemes are not positivity preserving, and SSPRK4 turns its dry
    cells negative within three steps, in the reference as here.)"""
    if not FDRIVER.exists():
        pytest.skip("Fortran driver not built (no Fortran compiler at build time)")
    c = P.ReferenceSetup(ic, 32, 20, wellbalanced=wb)
    steps = 3
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([c.nx, c.ny, steps, int(wb), int(c.bathy_flag > 0), 0, scheme, 0, 0, ic,
                  c.bathy_flag], dtype=np.int32).tofile(f)
        np.array([c.mesh.dx, c.mesh_y.dx, c.gravity, c.cfl, *c.mesh.wbnd, *c.x0, *c.x1],
                 dtype=np.float64).tofile(f)
    r = subprocess.run([str(FDRIVER), str(fin), str(fout)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(fout, "rb").read()
    nsteps = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
    body = np.frombuffer(raw[8:], dtype=np.float64)
    n3 = 3 * c.nx * c.ny
    dts, U = body[:nsteps], body[nsteps:nsteps + n3]
    l1, V = body[nsteps + n3:nsteps + n3 + 3], body[nsteps + n3 + 3:]
    s = swm.Solver.from_setup(c)
    s.keep_output_state()
    s.set_timescheme(scheme)
    t = 0.0
    for k in range(steps):
        dt, _, _ = s.timestep()
        assert dt == dts[k]
        s.step(dt)
        t = t + dt
    assert same_bits(U.reshape(c.ny, c.nx, 3), s.download())
    assert same_bits(l1, s.l1_error(ic, c.bathy_flag, c.x0, c.x1, t))
    assert same_bits(V.reshape(c.ny, c.nx, 3), s.download_output_primitive())
