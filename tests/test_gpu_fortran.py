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
from conftest import ROOT
from oracle_api import Oracle
from parity_util import describe_diff, same_bits

pytestmark = pytest.mark.gpu

FDRIVER = ROOT / "sw-mpdec_amd" / "lib" / "swm_fdriver"


def with_steps(p, k):
    p.steps = k
    return p


def write_input(path, p):
    m = p.mesh
    with open(path, "wb") as f:
        np.array([p.n, p.rows, p.steps, int(p.wellbalanced), int(p.dbdx is not None), 0],
                 dtype=np.int32).tofile(f)
        np.array([m.dx, p.mesh_y.dx, p.gravity, p.cfl, *m.wbnd], dtype=np.float64).tofile(f)
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
