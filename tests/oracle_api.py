"""Test-side access to the two oracles:

* ``Oracle``: ctypes binding of oracle/lib/liboracle.so, the C restatement of
  the reference time step (oracle/swm_oracle.c);
* ``run_swref``: the reference Fortran itself, built unchanged from its own
  sources into oracle/_ref/ (oracle/Makefile), driven through
  oracle/ref_driver.f90.

Only tests, smoke() and bench.py's cpu_baseline leg use this module.
"""
from __future__ import annotations

import ctypes
import os
import resource
import subprocess
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_LIB = ROOT / "oracle" / "lib" / "liboracle.so"
SWREF = {False: ROOT / "oracle" / "_ref" / "swref_p", True: ROOT / "oracle" / "_ref" / "swref_pwb"}
# builds without -DPATANKAR, for the explicit schemes (timescheme 1-4)
SWREF_EXPLICIT = {False: ROOT / "oracle" / "_ref" / "swref_e",
                  True: ROOT / "oracle" / "_ref" / "swref_ewb"}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


_lib = None


def oracle_lib():
    global _lib
    if _lib is None:
        if not ORACLE_LIB.exists():
            raise RuntimeError(f"{ORACLE_LIB} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(str(ORACLE_LIB))
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_bathy.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_init_wb.argtypes = [ctypes.c_void_p, _dp]
        L.orc_set_U.argtypes = [ctypes.c_void_p, _dp]
        L.orc_get_U.argtypes = [ctypes.c_void_p, _dp]
        L.orc_get_V.argtypes = [ctypes.c_void_p, _dp]
        L.orc_fvtd.argtypes = [ctypes.c_void_p] + [_dp] * 6
        L.orc_timestep.restype = ctypes.c_double
        L.orc_timestep.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_step_mpdec5.restype = ctypes.c_int
        L.orc_step_mpdec5.argtypes = [ctypes.c_void_p, ctypes.c_double, _ip]
        L.orc_step_patankar.restype = ctypes.c_int
        L.orc_step_patankar.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _ip]
        L.orc_set_hp_by_node.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_set_reconstruction_fix.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_set_reconstruction_fix.restype = ctypes.c_int
        L.orc_initial_data.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       _dp, _dp, ctypes.c_double, _dp, _dp, _dp, _dp]
        L.orc_initial_data.restype = ctypes.c_int
        L.orc_l1_error.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                   ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.orc_l1_error.restype = ctypes.c_int
        L.orc_step_explicit.restype = ctypes.c_int
        L.orc_step_explicit.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        L.orc_jacobi.restype = ctypes.c_int
        L.orc_jacobi.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_set_jacobi_limits.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        L.orc_patankar_div.restype = ctypes.c_double
        L.orc_patankar_div.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_weno5_first.argtypes = [_dp, _dp, _dp]
        L.orc_weno5_second.argtypes = [_dp, _dp]
        L.orc_theta.argtypes = [_dp]
        _lib = L
    return _lib


# Jacobi solves per step of each scheme (timescheme number): the explicit
# schemes 1-4 of the build without PATANKAR solve nothing
NSOLVES = {1: 0, 2: 0, 3: 0, 4: 0, 5: 13, 6: 1, 7: 2}


class Oracle:
    """One problem instance on the C oracle.  scheme: the reference's
    timescheme (1 ForwardEuler, 2 SSPRK4, 3 RK65, 4 DeC5 of the build without
    PATANKAR; 5 mPDeC5, 6 mPEuler, 7 mPDeC2); hp_by_node: the two-node
    schemes' by-node depth view instead of the reference's; recfix: the
    reference's ReconstructionFix (4, its own setting; 1 first order, 2 MUSCL
    for the cells the shock indicator flags)."""

    def __init__(self, problem, scheme: int = 5, hp_by_node: bool = False, recfix: int = 4):
        L = oracle_lib()
        self.p = problem
        self.scheme = scheme
        m = problem.mesh
        self.n = problem.n
        self.rows = problem.rows
        w = np.ascontiguousarray(m.wbnd, dtype=np.float64)
        self.h = L.orc_create(self.n, self.rows, m.dx, problem.mesh_y.dx, problem.gravity,
                              problem.cfl,
                              1 if problem.wellbalanced else 0, _ptr(w))
        if not self.h:
            raise ValueError("orc_create failed")
        L.orc_set_hp_by_node(self.h, 1 if hp_by_node else 0)
        if L.orc_set_reconstruction_fix(self.h, recfix) != 0:
            raise ValueError(f"reconstruction fix {recfix} is not 1, 2 or 4")
        if problem.dbdx is not None:
            L.orc_set_bathy(self.h, _ptr(np.ascontiguousarray(problem.dbdx)),
                            _ptr(np.ascontiguousarray(problem.dbdy)))
        if problem.wellbalanced:
            L.orc_init_wb(self.h, _ptr(np.ascontiguousarray(problem.Uwb)))
        self.set_U(problem.U0)

    def __del__(self):
        if getattr(self, "h", None):
            oracle_lib().orc_destroy(self.h)
            self.h = None

    def set_U(self, U):
        oracle_lib().orc_set_U(self.h, _ptr(np.ascontiguousarray(U, dtype=np.float64)))

    def get_U(self):
        U = np.empty((self.rows, self.n, 3))
        oracle_lib().orc_get_U(self.h, _ptr(U))
        return U

    def get_V(self):
        """The primitives of the last FVTD input (the writer's V)."""
        V = np.empty((self.rows, self.n, 3))
        oracle_lib().orc_get_V(self.h, _ptr(V))
        return V

    def fvtd(self):
        n, r = self.n, self.rows
        out = dict(Ut=np.empty((r, n, 3)), FX=np.empty((r, n + 1, 3)), FY=np.empty((r + 1, n, 3)),
                   S=np.empty((r, n, 3)), Prod=np.empty(5 * n * r), Dest=np.empty(5 * n * r))
        oracle_lib().orc_fvtd(self.h, *(_ptr(out[k]) for k in ("Ut", "FX", "FY", "S", "Prod", "Dest")))
        return out

    def timestep(self):
        lx, ly = ctypes.c_double(), ctypes.c_double()
        dt = oracle_lib().orc_timestep(self.h, ctypes.byref(lx), ctypes.byref(ly))
        return dt, lx.value, ly.value

    def step(self, dt):
        if self.scheme < 5:
            if oracle_lib().orc_step_explicit(self.h, self.scheme, dt) < 0:
                raise ValueError(f"timescheme {self.scheme} is not an explicit scheme")
            return [], 0
        it = (ctypes.c_int * 13)()
        nfail = oracle_lib().orc_step_patankar(self.h, self.scheme, dt, it)
        if nfail < 0:
            raise ValueError(f"timescheme {self.scheme} is not a Patankar scheme")
        return list(it)[:NSOLVES[self.scheme]], nfail

    def run(self, steps):
        """steps CFL steps; returns per-step (dt, lmx, lmy, iters) and final U."""
        hist = []
        for _ in range(steps):
            dt, lx, ly = self.timestep()
            it, _ = self.step(dt)
            hist.append((dt, lx, ly, it))
        return hist, self.get_U()

    def set_jacobi_limits(self, max_iter, tol):
        oracle_lib().orc_set_jacobi_limits(self.h, max_iter, tol)

    def jacobi(self, Mass, diag, rhs):
        sol = np.empty(self.n * self.rows)
        k = oracle_lib().orc_jacobi(self.h, _ptr(np.ascontiguousarray(Mass)),
                                    _ptr(np.ascontiguousarray(diag)),
                                    _ptr(np.ascontiguousarray(rhs)), _ptr(sol))
        return sol, k


def big_stack():
    """preexec_fn raising the stack limit: the reference keeps O(n) automatic
    arrays on the stack (src/equation.f90:313-315, src/timediscretization.f90:786-788)."""
    soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
    resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))


def have_swref():
    return SWREF[False].exists() and SWREF[True].exists()


def have_swref_explicit():
    return SWREF_EXPLICIT[False].exists() and SWREF_EXPLICIT[True].exists()


def run_swref(problem, mode: int, steps: int = 0, jac=None, timeout=600, scheme: int = 5,
              patankar=None, recfix: int = 4, bathy_flag=None):
    """Run the reference build on problem.  mode 1: one FVTimeDerivative,
    2: `steps` CFL steps through TimeDiscretization, 3: TimeStep only,
    4: jacobi on jac=(Mass, diag, rhs).  scheme: the reference's timescheme
    (5 mPDeC5, 6 mPEuler, 7 mPDeC2 with the PATANKAR build; 1-4 with the
    build without it, unless patankar=True forces the PATANKAR build).
    recfix: the ReconstructionFix the build runs with (4: its own setting).
    bathy_flag: BathymetryFlag (default: 1 when the problem has bathymetry
    tables, else 0); the reference evaluates its own Bathymetry_X/_Y."""
    if patankar is None:
        patankar = scheme >= 5
    exe = (SWREF if patankar else SWREF_EXPLICIT)[bool(problem.wellbalanced)]
    if not exe.exists():
        raise RuntimeError(f"{exe} not built")
    n, r = problem.n, problem.rows
    bflag = (1 if problem.dbdx is not None else 0) if bathy_flag is None else bathy_flag
    hdr = np.array([mode, n, steps, bflag, 1 if problem.wellbalanced else 0, r if r != n else 0,
                    scheme, recfix], dtype="<i4")
    fl = np.array([problem.x0, problem.x1, problem.gravity, problem.cfl], dtype="<f8")
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as td:
        fin, fout = Path(td) / "in.bin", Path(td) / "out.bin"
        with open(fin, "wb") as f:
            f.write(hdr.tobytes())
            f.write(fl.tobytes())
            if mode == 4:
                for a in jac:
                    f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
            else:
                f.write(np.ascontiguousarray(problem.U0, dtype="<f8").tobytes())
                if problem.wellbalanced:
                    f.write(np.ascontiguousarray(problem.Uwb, dtype="<f8").tobytes())
        subprocess.run([str(exe), str(fin), str(fout)], check=True, cwd=td, timeout=timeout,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, preexec_fn=big_stack)
        raw = np.fromfile(fout, dtype="<f8")
    if mode == 4:
        return raw
    if mode == 3:
        return tuple(raw[:3])
    if mode == 8:
        return raw.reshape(r, n, 3)
    if mode == 1:
        sizes = [("Ut", (r, n, 3)), ("FX", (r, n + 1, 3)), ("FY", (r + 1, n, 3)), ("S", (r, n, 3))]
        if patankar:
            sizes += [("Prod", (5 * n * r,)), ("Dest", (5 * n * r,))]
        sizes += [("wbnd", (4,)), ("dx", (2,)), ("MeshGP", (4, 4, r, n, 2))]
        out, off = {}, 0
        for k, shp in sizes:
            cnt = int(np.prod(shp))
            out[k] = raw[off:off + cnt].reshape(shp)
            off += cnt
        assert off == raw.size
        return out
    # mode 2
    # per step: dt, LambdaMaxX, LambdaMaxY, U; a non-finite dt ends the run
    rec = 3 + 3 * n * r
    hist = []
    U = None
    off = 0
    while off < raw.size:
        hist.append(tuple(raw[off:off + 3]))
        if off + 3 == raw.size:
            break
        U = raw[off + 3:off + rec].reshape(r, n, 3)
        off += rec
    return hist, U


def run_swref_initial_data(ic: int, nx: int, ny: int, x0, x1, gravity: float, bathy_flag: int,
                           wellbalanced: bool, timeout=600):
    """The reference's own initial data (driver mode 6): InitialCondition ic
    on [x0[0], x1[0]] x [x0[1], x1[1]] with nx x ny cells and BathymetryFlag
    bathy_flag.  Returns (U, Uwb), each (ny, nx, 3); Uwb is None unless
    wellbalanced (the WELLBALANCED build's InitializeWBVariables averages)."""
    exe = SWREF[bool(wellbalanced)]
    if not exe.exists():
        raise RuntimeError(f"{exe} not built")
    hdr = np.array([6, nx, ic, bathy_flag, 1 if wellbalanced else 0, ny, 5, 4], dtype="<i4")
    fl = np.array([x0[0], x1[0], gravity, 0.5, x0[1], x1[1]], dtype="<f8")
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as td:
        fin, fout = Path(td) / "in.bin", Path(td) / "out.bin"
        with open(fin, "wb") as f:
            f.write(hdr.tobytes())
            f.write(fl.tobytes())
        subprocess.run([str(exe), str(fin), str(fout)], check=True, cwd=td, timeout=timeout,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, preexec_fn=big_stack)
        raw = np.fromfile(fout, dtype="<f8")
    k = 3 * nx * ny
    arrs = [raw[i * k:(i + 1) * k].reshape(ny, nx, 3) for i in range(len(raw) // k)]
    return (arrs[-1], arrs[0] if wellbalanced else None)


def oracle_initial_data(ic: int, nx: int, ny: int, x0, x1, gravity: float, bathy_flag: int,
                        wellbalanced: bool):
    """The oracle's restatement of the reference's initial data: (U, Uwb,
    dbdx, dbdy) with U, Uwb (ny, nx, 3) (Uwb None unless wellbalanced) and the
    bathymetry tables (16, ny, nx) ([yg*4+xg][j][i])."""
    L = oracle_lib()
    U = np.zeros((ny, nx, 3))
    Uwb = np.zeros((ny, nx, 3)) if wellbalanced else None
    bx, by = np.zeros((16, ny, nx)), np.zeros((16, ny, nx))
    a0 = np.array(x0, dtype=np.float64)
    a1 = np.array(x1, dtype=np.float64)
    rc = L.orc_initial_data(ic, bathy_flag, nx, ny, _ptr(a0), _ptr(a1), gravity, _ptr(U),
                            _ptr(Uwb) if Uwb is not None else None, _ptr(bx), _ptr(by))
    if rc != 0:
        raise ValueError(f"orc_initial_data: initial condition {ic} not available")
    return U, Uwb, bx, by


def run_swref_l1_error(U, ic: int, x0, x1, gravity: float, bathy_flag: int, t: float,
                       timeout=600):
    """The reference's PostProcessing (driver mode 7): the L1 error of the
    state U (ny, nx, 3) against ExactFunction(ic, t) cell averages, as it
    writes it to ErrorL1_<nx>_<ny>.dat (list-directed, so to the digits
    printed)."""
    exe = SWREF[False]
    if not exe.exists():
        raise RuntimeError(f"{exe} not built")
    ny, nx = U.shape[:2]
    hdr = np.array([7, nx, ic, bathy_flag, 0, ny, 5, 4], dtype="<i4")
    fl = np.array([x0[0], x1[0], gravity, 0.5, x0[1], x1[1], t], dtype="<f8")
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as td:
        fin, fout = Path(td) / "in.bin", Path(td) / "out.bin"
        with open(fin, "wb") as f:
            f.write(hdr.tobytes())
            f.write(fl.tobytes())
            f.write(np.ascontiguousarray(U, dtype="<f8").tobytes())
        subprocess.run([str(exe), str(fin), str(fout)], check=True, cwd=td, timeout=timeout,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, preexec_fn=big_stack)
        txt = (Path(td) / f"ErrorL1_{nx:04d}_{ny:04d}.dat").read_text()
    return np.array([float(v) for v in txt.split()])


def oracle_l1_error(U, ic: int, x0, x1, gravity: float, bathy_flag: int, t: float):
    """The oracle's PostProcessing: L1 error of U (ny, nx, 3) against
    ExactFunction(ic, t)."""
    ny, nx = U.shape[:2]
    err = np.zeros(3)
    a0 = np.array(x0, dtype=np.float64)
    a1 = np.array(x1, dtype=np.float64)
    rc = oracle_lib().orc_l1_error(ic, bathy_flag, nx, ny, _ptr(a0), _ptr(a1), gravity, t,
                                   _ptr(np.ascontiguousarray(U, dtype=np.float64)), _ptr(err))
    if rc != 0:
        raise ValueError(f"orc_l1_error: initial condition {ic} not available")
    return err
