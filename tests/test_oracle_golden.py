"""The C oracle against fixtures the reference Fortran wrote (tests/golden/,
made by tests/golden/make_golden.py from the reference's own sources).  This
pins the oracle bitwise wherever /root/reference is absent."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import problems as P
from oracle_api import Oracle, oracle_initial_data
from parity_util import FVTD_KEYS, describe_diff, same_bits

GOLDEN = Path(__file__).resolve().parent / "golden"
FILES = sorted(f for f in GOLDEN.glob("*.npz")
               if not f.name.startswith(("jacobi", "initial_")))
INITIAL = sorted(GOLDEN.glob("initial_*.npz"))


def recfix(z):
    """The ReconstructionFix the fixture was written with (4: the reference's)."""
    return int(z["recfix"]) if "recfix" in z else 4


def load_problem(z):
    return P.Problem(int(z["n"]), z["U0"], ny=int(z["ny"]) if "ny" in z else 0,
                     wellbalanced=bool(int(z["wellbalanced"])),
                     Uwb=z["Uwb"] if "Uwb" in z else None,
                     dbdx=z["dbdx"] if "dbdx" in z else None,
                     dbdy=z["dbdy"] if "dbdy" in z else None,
                     gravity=float(z["gravity"]), cfl=float(z["cfl"]),
                     x0=float(z["x0"]), x1=float(z["x1"]))


@pytest.mark.parametrize("f", FILES, ids=lambda f: f.stem)
def test_fvtd_matches_reference_fixture(f):
    z = np.load(f)
    p = load_problem(z)
    got = Oracle(p, recfix=recfix(z)).fvtd()
    for k in FVTD_KEYS:
        a, b = got[k], z["fvtd_" + k]
        if k in ("Ut", "S"):
            a, b = a[..., 1:], b[..., 1:]
        assert same_bits(a, b), f"{k}: {describe_diff(a, b)}"


@pytest.mark.parametrize("f", [f for f in FILES if "fvtd_explicit_Ut" in np.load(f)],
                         ids=lambda f: f.stem)
def test_explicit_build_ut_matches_reference_fixture(f):
    """Ut(1:3) of the build without PATANKAR, Ut(1) included."""
    z = np.load(f)
    got = Oracle(load_problem(z), recfix=recfix(z)).fvtd()["Ut"]
    assert same_bits(got, z["fvtd_explicit_Ut"]), describe_diff(got, z["fvtd_explicit_Ut"])


@pytest.mark.parametrize("f", FILES, ids=lambda f: f.stem)
def test_mesh_quadrature_matches_reference(f):
    z = np.load(f)
    p = load_problem(z)
    m = p.mesh
    assert same_bits(m.wbnd, z["fvtd_wbnd"])
    gp = z["fvtd_MeshGP"]  # (yg, xg, j, i, dim)
    assert same_bits(gp[0, :, 0, :, 0].T, m.gp)
    assert same_bits(gp[:, 0, :, 0, 1].T, p.mesh_y.gp)


@pytest.mark.parametrize("f", FILES, ids=lambda f: f.stem)
def test_timestep_matches_reference(f):
    z = np.load(f)
    o = Oracle(load_problem(z))
    assert same_bits(o.timestep(), z["ts"])


@pytest.mark.parametrize("f", [f for f in FILES if "steps_hist" in np.load(f)],
                         ids=lambda f: f.stem)
def test_steps_match_reference(f):
    z = np.load(f)
    o = Oracle(load_problem(z), int(z["scheme"]) if "scheme" in z else 5, recfix=recfix(z))
    hist = z["steps_hist"]
    U = None
    for k, row in enumerate(hist):
        dt, lx, ly = o.timestep()
        assert same_bits([dt, lx, ly], row), f"step {k}: {(dt, lx, ly)} vs {row}"
        if not np.isfinite(dt):
            break  # the reference's own time loop cannot continue from here
        o.step(dt)
        U = o.get_U()
    assert same_bits(U, z["steps_U"]), describe_diff(U, z["steps_U"])


def test_jacobi_matches_reference():
    z = np.load(GOLDEN / "jacobi8.npz")
    n = int(z["n"])
    o = Oracle(P.Problem(n, np.zeros((n, n, 3))))
    sol, k = o.jacobi(z["mass"], z["diag"], z["rhs"])
    assert k > 1
    assert same_bits(sol, z["sol"])


@pytest.mark.parametrize("f", INITIAL, ids=lambda f: f.stem)
def test_initial_data_matches_reference_fixture(f):
    """The reference's FillInitialConditions (and InitializeWBVariables'
    averages) as its own build wrote them."""
    z = np.load(f)
    wb = "Uwb" in z
    U, Uwb, _, _ = oracle_initial_data(int(z["ic"]), int(z["nx"]), int(z["ny"]), z["x0"], z["x1"],
                                       float(z["gravity"]), int(z["bathy_flag"]), wb)
    assert same_bits(U, z["U"]), describe_diff(U, z["U"])
    if wb:
        assert same_bits(Uwb, z["Uwb"]), describe_diff(Uwb, z["Uwb"])
