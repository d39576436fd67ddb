"""Writes the golden fixtures of tests/golden/ from the reference Fortran itself.

The reference (accdavlo/sw-mpdec) ships no tests or fixtures, so these are
produced by running the reference's own sources, built unchanged by
oracle/Makefile into oracle/_ref/ and driven by oracle/ref_driver.f90, on
small seeded inputs.  Each .npz holds the inputs and the reference outputs
(data only).  Run from the repository root after __graft_entry__.build():

    python tests/golden/make_golden.py [case ...]      # default: every case
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT / "sw-mpdec_amd"))
sys.path.insert(0, str(ROOT / "tests"))

import problems as P  # noqa: E402
from oracle_api import run_swref, run_swref_initial_data  # noqa: E402


def problem_arrays(p):
    d = dict(n=p.n, ny=p.ny, wellbalanced=int(p.wellbalanced), gravity=p.gravity, cfl=p.cfl,
             x0=p.x0, x1=p.x1, U0=p.U0)
    if p.Uwb is not None:
        d["Uwb"] = p.Uwb
    if p.dbdx is not None:
        d["dbdx"] = p.dbdx
        d["dbdy"] = p.dbdy
    return d


CASES = {
    "vortex12": (lambda: P.vortex(12, 101), 4),
    "lakeflag1_wb10": (lambda: P.lake_flag1(10, True), 3),
    "lakeflag1_nowb9": (lambda: P.lake_flag1(9, False), 2),
    # the state of a Gaussian-bump lake, run with a flat bottom (the reference
    # has no Gaussian-bump bathymetry)
    "bumpstate_flat11": (lambda: P.Problem(11, P.lake_bumps(11, 102).U0, kind="flat"), 2),
    "dambreak10": (lambda: P.dam_break(10), 2),
    "shallowdam12": (lambda: P.dam_break(12, dry=0.1), 3),
    # rectangular grids (nx != ny, hence dx != dy on the unit square)
    "vortex_rect13x20": (lambda: P.vortex(13, 103, ny=20), 3),
    "lakeflag1_wb_rect11x8": (lambda: P.lake_flag1(11, True, ny=8), 2),
    # the two-node Patankar schemes (timescheme 6 mPEuler, 7 mPDeC2), run by
    # the reference until its state turns NaN
    "vortex12_mpeuler": (lambda: P.vortex(12, 104), 3, 6),
    "lakeflag1_wb10_mpdec2": (lambda: P.lake_flag1(10, True), 2, 7),
    "vortex_rect13x9_mpdec2": (lambda: P.vortex(13, 105, ny=9), 2, 7),
    # the explicit schemes of the build without PATANKAR (timescheme 1-4)
    "vortex12_euler": (lambda: P.vortex(12, 106), 3, 1),
    "lakeflag1_wb10_ssprk4": (lambda: P.lake_flag1(10, True), 2, 2),
    "vortex_rect11x14_rk65": (lambda: P.vortex(11, 107, ny=14), 2, 3),
    "shallowdam12_dec5": (lambda: P.dam_break(12, dry=0.1), 2, 4),
    # ReconstructionFix 1 (first order) and 2 (MUSCL) on the cells the Jameson
    # indicator flags (the reference's own setting is ReconstructionFix = 4)
    "lakeflag1_wb10_fix1": (lambda: P.lake_flag1(10, True, pert=0.05), 2, 5, 1),
    "vortex_rect13x11_fix2": (lambda: P.vortex(13, 108, ny=11), 3, 5, 2),
    "shallowdam12_ssprk4_fix2": (lambda: P.dam_break(12, dry=0.1), 2, 2, 2),
}


# The reference's initial data (driver mode 6): InitialCondition, nx, ny on
# the reference's own domain for that initial condition.
INITIAL_DATA = [(1, 23, 11), (2, 16, 16), (3, 20, 8), (6, 24, 10)]


def main(only=None):
    for name, (mk, steps, *rest) in CASES.items():
        if only and name not in only:
            continue
        scheme = rest[0] if rest else 5
        recfix = rest[1] if len(rest) > 1 else 4
        p = mk()
        out = problem_arrays(p)
        out["scheme"] = scheme
        out["recfix"] = recfix
        f = run_swref(p, 1, recfix=recfix)
        for k in ("Ut", "FX", "FY", "S", "Prod", "Dest", "wbnd", "MeshGP"):
            out["fvtd_" + k] = f[k]
        if scheme < 5:  # that build's Ut(1:3)
            out["fvtd_explicit_Ut"] = run_swref(p, 1, scheme=scheme, recfix=recfix)["Ut"]
        if steps:
            hist, U = run_swref(p, 2, steps=steps, scheme=scheme, recfix=recfix)
            out["steps_hist"] = np.array(hist)
            out["steps_U"] = U if U is not None else np.zeros(0)
        out["ts"] = np.array(run_swref(p, 3, recfix=recfix))
        np.savez_compressed(HERE / f"{name}.npz", **out)
        print(f"wrote {name}.npz")
    for ic, nx, ny in INITIAL_DATA:
        name = f"initial_ic{ic}_{nx}x{ny}"
        if only and name not in only:
            continue
        x0, x1, g, bf = P.REFERENCE_SETUPS[ic]
        wb = ic in (2, 3)
        U, Uwb = run_swref_initial_data(ic, nx, ny, x0, x1, g, bf, wb)
        out = dict(ic=ic, nx=nx, ny=ny, x0=np.array(x0), x1=np.array(x1), gravity=g,
                   bathy_flag=bf, U=U)
        if wb:
            out["Uwb"] = Uwb
        np.savez_compressed(HERE / f"{name}.npz", **out)
        print(f"wrote {name}.npz")
    if only:
        return
    # Jacobi on a random diagonally dominant M-matrix in the reference CSR pattern
    rng = np.random.default_rng(7)
    n = 8
    N = n * n
    mass = -rng.uniform(0, 0.2, 5 * N)
    mass[2::5] = 0.0  # self slot
    diag = 1.0 + rng.uniform(0, 1, N)
    rhs = rng.uniform(0, 1, N)
    p = P.Problem(n, np.zeros((n, n, 3)))
    sol = run_swref(p, 4, jac=(mass, diag, rhs))
    np.savez_compressed(HERE / "jacobi8.npz", n=n, mass=mass, diag=diag, rhs=rhs, sol=sol)
    print("wrote jacobi8.npz")


if __name__ == "__main__":
    main(sys.argv[1:])
