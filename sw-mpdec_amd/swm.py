"""ctypes binding of libswm.so, the C ABI of include/swm.h.

The product path: every call goes to the HIP library.  There is no CPU
fallback -- importing this module on a machine without the built library, or
creating a solver without a GPU, raises.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
# SWM_LIBRARY selects another build of the same library (e.g. one compiled
# with different -D options); the default is the in-tree build.
LIB_PATH = Path(os.environ.get("SWM_LIBRARY") or PKG / "lib" / "libswm.so")

_dp = ctypes.POINTER(ctypes.c_double)
EXEC_STATS = ("speculative_steps", "redone_steps", "graph_launches", "graph_captures", "graph",
              "margin")
_ip = ctypes.POINTER(ctypes.c_int)

SWM_OK = 0
PROF_NCAT = 8
# the event categories, by index (swm_profile_name)
PROF_NAMES = ("fvtd_x", "fvtd_y", "fvtd_update", "momentum", "patankar_assemble",
              "jacobi_sweep", "timestep", "other")


def _prof_mask(categories):
    mask = 0
    for c in categories:
        if c not in PROF_NAMES:
            raise ValueError(f"unknown profile category {c!r}; one of {PROF_NAMES}")
        mask |= 1 << PROF_NAMES.index(c)
    return mask


class SwmError(RuntimeError):
    pass


def _ptr(a):
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_dp)


_lib = None

FAMILY_MAX_BUMPS = 8


class Family(ctypes.Structure):
    """swm_family (include/swm.h)."""
    _fields_ = [("kind", ctypes.c_int), ("nbumps", ctypes.c_int), ("eps", ctypes.c_double),
                ("A", ctypes.c_double * FAMILY_MAX_BUMPS),
                ("cx", ctypes.c_double * FAMILY_MAX_BUMPS),
                ("cy", ctypes.c_double * FAMILY_MAX_BUMPS),
                ("s", ctypes.c_double * FAMILY_MAX_BUMPS),
                ("amp", ctypes.c_double), ("px", ctypes.c_double), ("py", ctypes.c_double),
                ("ps", ctypes.c_double), ("radius", ctypes.c_double), ("h_in", ctypes.c_double),
                ("h_out", ctypes.c_double)]

    @classmethod
    def of(cls, spec) -> "Family":
        """From a problems.FamilySpec."""
        f = cls()
        f.kind = int(spec.kind)
        q = spec.params
        for key in ("eps", "amp", "px", "py", "ps", "radius", "h_in", "h_out"):
            if key in q:
                setattr(f, key, float(q[key]))
        if "A" in q:
            if len(q["A"]) > FAMILY_MAX_BUMPS:
                raise ValueError(f"at most {FAMILY_MAX_BUMPS} bumps")
            f.nbumps = len(q["A"])
            for k in range(f.nbumps):
                f.A[k], f.cx[k], f.cy[k], f.s[k] = q["A"][k], q["cx"][k], q["cy"][k], q["s"][k]
        return f


def lib():
    """Load libswm.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise SwmError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = ctypes.CDLL(str(LIB_PATH))
    vp = ctypes.c_void_p
    L.swm_create.restype = ctypes.c_int
    L.swm_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                             ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp,
                             ctypes.POINTER(vp)]
    L.swm_destroy.argtypes = [vp]
    L.swm_destroy.restype = None
    L.swm_last_error.restype = ctypes.c_char_p
    L.swm_last_error.argtypes = [vp]
    L.swm_footprint_bytes.restype = ctypes.c_longlong
    L.swm_footprint_bytes.argtypes = [ctypes.c_int] * 4
    L.swm_set_bathy_grad.argtypes = [vp, _dp, _dp]
    L.swm_init_wb.argtypes = [vp, _dp, ctypes.c_int]
    L.swm_upload.argtypes = [vp, _dp, ctypes.c_int]
    L.swm_download.argtypes = [vp, _dp]
    L.swm_timestep.argtypes = [vp, _dp, _dp, _dp]
    L.swm_step_mpdec5.argtypes = [vp, ctypes.c_double, ctypes.c_double, _ip]
    L.swm_step_mpeuler.argtypes = [vp, ctypes.c_double, ctypes.c_double, _ip]
    L.swm_step_mpdec2.argtypes = [vp, ctypes.c_double, ctypes.c_double, _ip]
    L.swm_set_timescheme.argtypes = [vp, ctypes.c_int]
    for name in ("swm_step_forward_euler", "swm_step_ssprk4", "swm_step_rk65", "swm_step_dec5"):
        getattr(L, name).argtypes = [vp, ctypes.c_double, ctypes.c_double]
    L.swm_set_hp_association.argtypes = [vp, ctypes.c_int]
    L.swm_set_reconstruction_fix.argtypes = [vp, ctypes.c_int]
    L.swm_fill_initial_data.argtypes = [vp, ctypes.c_int, ctypes.c_int, _dp, _dp]
    L.swm_l1_error.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp]
    L.swm_download_primitive.argtypes = [vp, _dp]
    L.swm_download_primitive_async.argtypes = [vp, _dp]
    L.swm_keep_output_state.argtypes = [vp, ctypes.c_int]
    L.swm_download_output_primitive.argtypes = [vp, _dp]
    L.swm_output_wait.argtypes = [vp]
    L.swm_run_steps.argtypes = [vp, ctypes.c_int, _dp, _ip]
    L.swm_fvtd.argtypes = [vp] + [_dp] * 6
    L.swm_set_jacobi_limits.argtypes = [vp, ctypes.c_int, ctypes.c_double]
    L.swm_ambiguous_decisions.restype = ctypes.c_longlong
    L.swm_ambiguous_decisions.argtypes = [vp]
    L.swm_sync.argtypes = [vp]
    L.swm_debug_divsqrt.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
    L.swm_debug_shared_div.argtypes = [ctypes.c_int, _dp, _dp, _dp]
    L.swm_track_stage_min.argtypes = [vp, ctypes.c_int]
    L.swm_advance.argtypes = [vp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int, _ip, _dp]
    L.swm_set_execution.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.swm_set_tile_shape.argtypes = [vp, ctypes.c_int]
    L.swm_tile_shape.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.swm_execution_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong)]
    L.swm_status_reads.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong)]
    L.swm_stage_min_h.argtypes = [vp, _dp]
    L.swm_debug_weno5.argtypes = [ctypes.c_int, _dp, _dp, _dp]
    L.swm_profile_enable.argtypes = [vp, ctypes.c_int]
    L.swm_profile_categories.argtypes = [vp, ctypes.c_uint]
    L.swm_profile_read.argtypes = [vp, _dp, ctypes.POINTER(ctypes.c_longlong)]
    L.swm_profile_name.restype = ctypes.c_char_p
    L.swm_profile_name.argtypes = [ctypes.c_int]
    L.swm_fill_family.argtypes = [vp, ctypes.POINTER(Family)]
    L.swm_debug_exp.argtypes = [ctypes.c_int, _dp, _dp]
    L.swm_timestep_noalloc.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 4 + \
        [vp, ctypes.c_int, vp, vp]
    L.swm_debug_sync_cost.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
    _lib = L
    return L


# the largest grid swm_timestep_noalloc accepts (include/swm.h)
TIMESTEP_NOALLOC_MAX_CELLS = 1 << 24


def sync_cost(nblk: int, rounds: int = 1000):
    """Microseconds per round of a step boundary with nblk blocks: (one launch
    per round, launches replayed from a HIP graph, a grid barrier inside one
    cooperative launch)."""
    out = np.zeros(3)
    rc = lib().swm_debug_sync_cost(int(nblk), int(rounds), _ptr(out))
    if rc != SWM_OK:
        raise SwmError(f"swm_debug_sync_cost: {lib().swm_last_error(None).decode()} (rc={rc})")
    return tuple(float(v) for v in out)


def footprint_bytes(nx, ny, wellbalanced, bathymetry) -> int:
    return int(lib().swm_footprint_bytes(nx, ny, int(bool(wellbalanced)), int(bool(bathymetry))))


# time schemes by the reference's timescheme number: entry point of one step
# and Jacobi solves per step (1-4: the explicit schemes of the build without
# PATANKAR; 5-7: the modified-Patankar ones)
SCHEMES = {1: ("swm_step_forward_euler", 0), 2: ("swm_step_ssprk4", 0), 3: ("swm_step_rk65", 0),
           4: ("swm_step_dec5", 0), 5: ("swm_step_mpdec5", 13), 6: ("swm_step_mpeuler", 1),
           7: ("swm_step_mpdec2", 2)}
HP_AS_REFERENCE, HP_BY_NODE = 0, 1


class Solver:
    """One grid on the device; mirrors the reference module state of one run."""

    def __init__(self, nx, ny, dx, dy, gravity, cfl, wellbalanced, wbnd):
        L = lib()
        self._h = ctypes.c_void_p()
        w = np.ascontiguousarray(wbnd, dtype=np.float64)
        rc = L.swm_create(nx, ny, dx, dy, gravity, cfl, int(bool(wellbalanced)), _ptr(w),
                          ctypes.byref(self._h))
        if rc != SWM_OK:
            raise SwmError(f"swm_create: {L.swm_last_error(None).decode()} (rc={rc})")
        self.nx, self.ny = nx, ny
        self.timescheme = 5

    @classmethod
    def from_problem(cls, p, recfix=4):
        """A handle holding problem p; recfix: ReconstructionFix (set before
        the well-balanced state is built, which depends on it)."""
        m = p.mesh
        s = cls(p.n, p.rows, m.dx, p.mesh_y.dx, p.gravity, p.cfl, p.wellbalanced, m.wbnd)
        if recfix != 4:
            s.set_reconstruction_fix(recfix)
        if p.dbdx is not None:
            s.set_bathy_grad(p.dbdx, p.dbdy)
        if p.wellbalanced:
            s.init_wb(p.Uwb)
        s.upload(p.U0)
        return s

    @classmethod
    def from_family(cls, spec, recfix=4):
        """A handle holding the family problem spec (a problems.FamilySpec),
        its data made on the device (swm_fill_family)."""
        m = spec.mesh
        s = cls(spec.n, spec.rows, m.dx, spec.mesh_y.dx, spec.gravity, spec.cfl,
                spec.wellbalanced, m.wbnd)
        if recfix != 4:
            s.set_reconstruction_fix(recfix)
        s.fill_family(spec)
        return s

    def fill_family(self, spec):
        f = Family.of(spec)
        self._check(lib().swm_fill_family(self._h, ctypes.byref(f)), "swm_fill_family")

    @classmethod
    def from_setup(cls, c, recfix=4):
        """A handle holding the reference's initial condition c (a
        problems.ReferenceSetup), its data made on the device."""
        s = cls(c.nx, c.ny, c.mesh.dx, c.mesh_y.dx, c.gravity, c.cfl, c.wellbalanced,
                c.mesh.wbnd)
        if recfix != 4:
            s.set_reconstruction_fix(recfix)
        s.fill_initial_data(c.ic, c.bathy_flag, c.x0, c.x1)
        return s

    def _check(self, rc, what):
        if rc < 0:
            raise SwmError(f"{what}: {lib().swm_last_error(self._h).decode()} (rc={rc})")
        return rc

    def close(self):
        if self._h:
            lib().swm_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_bathy_grad(self, dbdx, dbdy):
        self._check(lib().swm_set_bathy_grad(self._h, _ptr(np.ascontiguousarray(dbdx)),
                                             _ptr(np.ascontiguousarray(dbdy))), "swm_set_bathy_grad")

    def init_wb(self, Uwb, ghosts=0):
        self._check(lib().swm_init_wb(self._h, _ptr(np.ascontiguousarray(Uwb)), ghosts),
                    "swm_init_wb")

    def upload(self, U, ghosts=0):
        self._check(lib().swm_upload(self._h, _ptr(np.ascontiguousarray(U)), ghosts), "swm_upload")

    def download(self):
        U = np.empty((self.ny, self.nx, 3))
        self._check(lib().swm_download(self._h, _ptr(U)), "swm_download")
        return U

    def timestep(self):
        dt, lx, ly = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(lib().swm_timestep(self._h, ctypes.byref(dt), ctypes.byref(lx),
                                       ctypes.byref(ly)), "swm_timestep")
        return dt.value, lx.value, ly.value

    def set_timescheme(self, timescheme):
        """The scheme of step/run_steps/advance: 1 ForwardEuler, 2 SSPRK4,
        3 RK65, 4 DeC5, 5 mPDeC5, 6 mPEuler, 7 mPDeC2."""
        self._check(lib().swm_set_timescheme(self._h, int(timescheme)), "swm_set_timescheme")
        self.timescheme = int(timescheme)

    def set_hp_association(self, mode):
        self._check(lib().swm_set_hp_association(self._h, int(mode)), "swm_set_hp_association")

    def fill_initial_data(self, ic, bathy_flag, x0, x1):
        """The reference's initial data of InitialCondition ic made on the
        device (state, bathymetry tables, well-balanced state) on the domain
        [x0[0], x1[0]] x [x0[1], x1[1]]."""
        a0 = np.ascontiguousarray(x0, dtype=np.float64)
        a1 = np.ascontiguousarray(x1, dtype=np.float64)
        self._check(lib().swm_fill_initial_data(self._h, int(ic), int(bathy_flag), _ptr(a0),
                                                _ptr(a1)), "swm_fill_initial_data")

    def l1_error(self, ic, bathy_flag, x0, x1, t):
        """PostProcessing's L1 error of the state against ExactFunction(ic, t)."""
        a0 = np.ascontiguousarray(x0, dtype=np.float64)
        a1 = np.ascontiguousarray(x1, dtype=np.float64)
        err = np.zeros(3)
        self._check(lib().swm_l1_error(self._h, int(ic), int(bathy_flag), float(t), _ptr(a0),
                                       _ptr(a1), _ptr(err)), "swm_l1_error")
        return err

    def download_primitive(self):
        """(ny, nx, 3) primitives (h, u, v) of the state."""
        V = np.empty((self.ny, self.nx, 3))
        self._check(lib().swm_download_primitive(self._h, _ptr(V)), "swm_download_primitive")
        return V

    def keep_output_state(self, on=True):
        """Keep, from now on, the input of each step's last FVTD (the state
        the reference's writer prints)."""
        self._check(lib().swm_keep_output_state(self._h, int(bool(on))), "swm_keep_output_state")

    def download_output_primitive(self):
        """(ny, nx, 3): the reference writer's V after the last step."""
        V = np.empty((self.ny, self.nx, 3))
        self._check(lib().swm_download_output_primitive(self._h, _ptr(V)),
                    "swm_download_output_primitive")
        return V

    def download_primitive_async(self, V):
        """Starts copying a snapshot of the primitives into V ((ny, nx, 3)
        float64, C order; kept alive until output_wait)."""
        if V.shape != (self.ny, self.nx, 3) or V.dtype != np.float64 or not V.flags.c_contiguous:
            raise ValueError("V must be a C-contiguous (ny, nx, 3) float64 array")
        self._out = getattr(self, "_out", []) + [V]  # alive until the copies are done
        self._check(lib().swm_download_primitive_async(self._h, _ptr(V)),
                    "swm_download_primitive_async")

    def output_wait(self):
        self._check(lib().swm_output_wait(self._h), "swm_output_wait")
        self._out = []

    def set_reconstruction_fix(self, recfix):
        """ReconstructionFix for the cells the shock indicator flags: 4 (WENO5,
        the reference's setting), 1 first order, 2 MUSCL.  Discards the
        well-balanced state (init_wb again)."""
        self._check(lib().swm_set_reconstruction_fix(self._h, int(recfix)),
                    "swm_set_reconstruction_fix")

    def step(self, dt, t=0.0, timescheme=None):
        ts = self.timescheme if timescheme is None else int(timescheme)
        name, ns = SCHEMES[ts]
        if ns == 0:
            self._check(getattr(lib(), name)(self._h, t, dt), name)
            return [], 0
        it = (ctypes.c_int * ns)()
        nonconv = self._check(getattr(lib(), name)(self._h, t, dt, it), name)
        return list(it), nonconv

    def run(self, steps):
        """steps CFL steps; returns per-step (dt, lmx, lmy, iters) and final U."""
        hist = []
        for _ in range(steps):
            dt, lx, ly = self.timestep()
            it, _ = self.step(dt)
            hist.append((dt, lx, ly, it))
        return hist, self.download()

    def run_steps(self, steps):
        ns = SCHEMES[self.timescheme][1]
        dts = np.zeros(steps)
        its = np.zeros(ns * steps, dtype=np.int32)
        nonconv = self._check(lib().swm_run_steps(self._h, steps, _ptr(dts),
                                                  its.ctypes.data_as(_ip)), "swm_run_steps")
        return dts, its.reshape(steps, ns), nonconv

    def advance(self, t, t_analyze, t_end, max_steps=1 << 30):
        """The reference's time loop to the next output time (swm_advance):
        returns (t, dts of the steps taken, non-converged solves)."""
        tt = ctypes.c_double(t)
        done = ctypes.c_int(0)
        cap = min(max_steps, 1 << 20)
        dts = np.zeros(cap)
        nonconv = self._check(lib().swm_advance(self._h, ctypes.byref(tt), t_analyze, t_end,
                                                cap, ctypes.byref(done), _ptr(dts)),
                              "swm_advance")
        return tt.value, dts[:done.value].copy(), nonconv

    def set_execution(self, use_graph=-1, margin=-1):
        """Step execution mode (results are identical in every mode): use_graph
        0/1, margin >= 0 extra speculative sweeps; -1 = autotuned; margin -2 =
        no speculation."""
        self._check(lib().swm_set_execution(self._h, use_graph, margin), "swm_set_execution")

    def execution_stats(self):
        out = (ctypes.c_longlong * 6)()
        self._check(lib().swm_execution_stats(self._h, out), "swm_execution_stats")
        return dict(zip(EXEC_STATS, list(out)))

    def status_reads(self):
        """Host waits for the device status so far."""
        n = ctypes.c_longlong(0)
        self._check(lib().swm_status_reads(self._h, ctypes.byref(n)), "swm_status_reads")
        return n.value

    def set_tile_shape(self, shape=-1):
        """Face-kernel tile shape: 0 (16x16), 1 (32x16 / 16x32), -1 autotuned."""
        self._check(lib().swm_set_tile_shape(self._h, int(shape)), "swm_set_tile_shape")

    def tile_shape(self):
        """(shape in use, (x-kernel w, h, y-kernel w, h))."""
        d = (ctypes.c_int * 4)()
        sh = self._check(lib().swm_tile_shape(self._h, d), "swm_tile_shape")
        return sh, tuple(d)

    def track_stage_min(self, on=True):
        """Record the minimum depth over every DeC stage solution from now on."""
        self._check(lib().swm_track_stage_min(self._h, int(bool(on))), "swm_track_stage_min")

    def stage_min_h(self):
        v = ctypes.c_double()
        self._check(lib().swm_stage_min_h(self._h, ctypes.byref(v)), "swm_stage_min_h")
        return v.value

    def fvtd(self):
        n, m = self.nx, self.ny
        out = dict(Ut=np.empty((m, n, 3)), FX=np.empty((m, n + 1, 3)), FY=np.empty((m + 1, n, 3)),
                   S=np.empty((m, n, 3)), Prod=np.empty(5 * n * m), Dest=np.empty(5 * n * m))
        self._check(lib().swm_fvtd(self._h, *(_ptr(out[k]) for k in
                                               ("Ut", "FX", "FY", "S", "Prod", "Dest"))), "swm_fvtd")
        return out

    def set_jacobi_limits(self, max_iter, tol):
        self._check(lib().swm_set_jacobi_limits(self._h, max_iter, tol), "swm_set_jacobi_limits")

    def ambiguous_decisions(self):
        return int(lib().swm_ambiguous_decisions(self._h))

    def sync(self):
        self._check(lib().swm_sync(self._h), "swm_sync")

    def profile(self, on=True, categories=None):
        """HIP events around every launch (on), or only around the launches
        of the named categories (PROF_NAMES entries)."""
        if categories is None:
            self._check(lib().swm_profile_enable(self._h, int(bool(on))), "swm_profile_enable")
        else:
            self._check(lib().swm_profile_categories(self._h, _prof_mask(categories) if on else 0),
                        "swm_profile_categories")

    def profile_read(self):
        """{kernel category: (milliseconds, launches)} since the last read."""
        ms = (ctypes.c_double * PROF_NCAT)()
        cnt = (ctypes.c_longlong * PROF_NCAT)()
        self._check(lib().swm_profile_read(self._h, ms, cnt), "swm_profile_read")
        return {lib().swm_profile_name(c).decode(): (ms[c], int(cnt[c])) for c in range(PROF_NCAT)}


def debug_divsqrt(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    q = np.empty_like(a)
    r = np.empty_like(a)
    rc = lib().swm_debug_divsqrt(a.size, _ptr(a), _ptr(b), _ptr(q), _ptr(r))
    if rc != SWM_OK:
        raise SwmError(f"swm_debug_divsqrt rc={rc}")
    return q, r


def debug_weno5(q):
    """Device WENO5 first/second sweeps of stencils q (n, 5) -> (n, 2), (n, 4)."""
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 5)
    first = np.empty((q.shape[0], 2))
    second = np.empty((q.shape[0], 4))
    rc = lib().swm_debug_weno5(q.shape[0], _ptr(q), _ptr(first), _ptr(second))
    if rc != SWM_OK:
        raise SwmError(f"swm_debug_weno5 rc={rc}")
    return first, second


def debug_exp(x):
    """The device's exponential of the benchmark families (swm_debug_exp)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    rc = lib().swm_debug_exp(x.size, _ptr(x), _ptr(y))
    if rc != SWM_OK:
        raise SwmError(f"swm_debug_exp rc={rc}")
    return y


def debug_shared_div(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    q = np.empty_like(a)
    rc = lib().swm_debug_shared_div(a.size, _ptr(a), _ptr(b), _ptr(q))
    if rc != SWM_OK:
        raise SwmError(f"swm_debug_shared_div rc={rc}")
    return q


# ---------------------------------------------------------------------------
# batched / served solves
# ---------------------------------------------------------------------------
SWM_E_ARG, SWM_E_HIP, SWM_E_STATE, SWM_E_NOMEM, SWM_E_BUDGET = -1, -2, -3, -4, -5  # include/swm.h
JOB_RUNNING, JOB_DONE, JOB_FAILED = 0, 1, 2


class JobDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double),
                ("gravity", ctypes.c_double), ("cfl", ctypes.c_double),
                ("wellbalanced", ctypes.c_int), ("wgp_bnd", ctypes.c_double * 4),
                ("U0", _dp), ("Uwb", _dp), ("dbdx", _dp), ("dbdy", _dp),
                ("nsteps", ctypes.c_int), ("timescheme", ctypes.c_int),
                ("hp_association", ctypes.c_int), ("reconstruction_fix", ctypes.c_int),
                ("initial_condition", ctypes.c_int), ("bathymetry_flag", ctypes.c_int),
                ("mesh_x0", ctypes.c_double * 2), ("mesh_x1", ctypes.c_double * 2),
                ("family", ctypes.POINTER(Family))]


class JobInfo(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int), ("steps_done", ctypes.c_int),
                ("steps_left", ctypes.c_int), ("t", ctypes.c_double),
                ("jacobi_sweeps", ctypes.c_longlong), ("nonconverged", ctypes.c_int),
                ("bytes", ctypes.c_longlong)]


def _engine_lib():
    L = lib()
    if getattr(L, "_engine_bound", False):
        return L
    vp = ctypes.c_void_p
    L.swm_engine_create.argtypes = [ctypes.c_longlong, ctypes.POINTER(vp)]
    L.swm_engine_destroy.argtypes = [vp]
    L.swm_engine_destroy.restype = None
    L.swm_engine_last_error.restype = ctypes.c_char_p
    L.swm_engine_last_error.argtypes = [vp]
    L.swm_job_footprint_bytes.restype = ctypes.c_longlong
    L.swm_job_footprint_bytes.argtypes = [ctypes.c_int] * 4
    L.swm_engine_bytes_in_use.restype = ctypes.c_longlong
    L.swm_engine_bytes_in_use.argtypes = [vp]
    L.swm_engine_submit.argtypes = [vp, ctypes.POINTER(JobDesc), ctypes.POINTER(ctypes.c_int)]
    L.swm_engine_step.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.swm_engine_run.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                 ctypes.POINTER(ctypes.c_int)]
    L.swm_engine_job_info.argtypes = [vp, ctypes.c_int, ctypes.POINTER(JobInfo)]
    L.swm_engine_fetch.argtypes = [vp, ctypes.c_int, _dp]
    L.swm_engine_l1_error.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      _dp, _dp, _dp]
    L.swm_engine_keep_output_state.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.swm_engine_fetch_output_primitive.argtypes = [vp, ctypes.c_int, _dp]
    L.swm_engine_release.argtypes = [vp, ctypes.c_int]
    L.swm_engine_set_jacobi_limits.argtypes = [vp, ctypes.c_int, ctypes.c_double]
    L.swm_engine_profile_enable.argtypes = [vp, ctypes.c_int]
    L.swm_engine_profile_categories.argtypes = [vp, ctypes.c_uint]
    L.swm_engine_profile_read.argtypes = [vp, _dp, ctypes.POINTER(ctypes.c_longlong)]
    L.swm_engine_track_stage_min.argtypes = [vp, ctypes.c_int]
    L.swm_engine_set_execution.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.swm_engine_set_tile_shape.argtypes = [vp, ctypes.c_int]
    L.swm_engine_tile_shape.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.swm_engine_execution_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong)]
    L.swm_engine_status_reads.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong)]
    L.swm_engine_stage_min_h.argtypes = [vp, ctypes.c_int, _dp]
    L._engine_bound = True
    return L


def job_footprint_bytes(nx, ny, wellbalanced, bathymetry) -> int:
    return int(_engine_lib().swm_job_footprint_bytes(nx, ny, int(bool(wellbalanced)),
                                                     int(bool(bathymetry))))


class Engine:
    """Many independent grids on one GPU, advanced in lockstep (one mPDeC5
    step of every running grid per step()); grids can join and leave between
    steps."""

    def __init__(self, budget_bytes: int = 0):
        L = _engine_lib()
        self._e = ctypes.c_void_p()
        rc = L.swm_engine_create(int(budget_bytes), ctypes.byref(self._e))
        if rc != SWM_OK:
            raise SwmError(f"swm_engine_create: {L.swm_engine_last_error(None).decode()}")
        self._keep = {}

    def _check(self, rc, what):
        if rc < 0:
            raise SwmError(f"{what}: {_engine_lib().swm_engine_last_error(self._e).decode()} "
                           f"(rc={rc})")
        return rc

    def close(self):
        if self._e:
            _engine_lib().swm_engine_destroy(self._e)
            self._e = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, p, nsteps, timescheme=5, hp_association=HP_AS_REFERENCE, recfix=4):
        """Copies problem p to the device; returns its job id, or None when it
        does not fit the memory budget right now."""
        m = p.mesh
        arrs = [np.ascontiguousarray(p.U0, dtype=np.float64)]
        d = JobDesc()
        d.nx, d.ny = p.n, p.rows
        d.dx, d.dy = m.dx, p.mesh_y.dx
        d.gravity = p.gravity
        d.cfl = p.cfl
        d.wellbalanced = int(bool(p.wellbalanced))
        for i in range(4):
            d.wgp_bnd[i] = m.wbnd[i]
        d.U0 = _ptr(arrs[0])
        if p.wellbalanced:
            arrs.append(np.ascontiguousarray(p.Uwb, dtype=np.float64))
            d.Uwb = _ptr(arrs[-1])
        if p.dbdx is not None:
            arrs.append(np.ascontiguousarray(p.dbdx, dtype=np.float64))
            d.dbdx = _ptr(arrs[-1])
            arrs.append(np.ascontiguousarray(p.dbdy, dtype=np.float64))
            d.dbdy = _ptr(arrs[-1])
        d.nsteps = int(nsteps)
        d.timescheme = int(timescheme)
        d.hp_association = int(hp_association)
        d.reconstruction_fix = int(recfix)
        jid = ctypes.c_int(-1)
        rc = _engine_lib().swm_engine_submit(self._e, ctypes.byref(d), ctypes.byref(jid))
        if rc == SWM_E_BUDGET:
            return None
        self._check(rc, "swm_engine_submit")
        self._keep[jid.value] = (p.rows, p.n)
        return jid.value

    def submit_family(self, spec, nsteps, timescheme=5, hp_association=HP_AS_REFERENCE,
                      recfix=4):
        """Submits the family problem spec (a problems.FamilySpec), its data
        made on the device; returns the job id, or None when it does not fit
        the memory budget right now."""
        m = spec.mesh
        f = Family.of(spec)
        d = JobDesc()
        d.nx, d.ny = spec.n, spec.rows
        d.dx, d.dy = m.dx, spec.mesh_y.dx
        d.gravity = spec.gravity
        d.cfl = spec.cfl
        d.wellbalanced = int(bool(spec.wellbalanced))
        for i in range(4):
            d.wgp_bnd[i] = m.wbnd[i]
        d.nsteps = int(nsteps)
        d.timescheme = int(timescheme)
        d.hp_association = int(hp_association)
        d.reconstruction_fix = int(recfix)
        d.family = ctypes.pointer(f)
        jid = ctypes.c_int(-1)
        rc = _engine_lib().swm_engine_submit(self._e, ctypes.byref(d), ctypes.byref(jid))
        if rc == SWM_E_BUDGET:
            return None
        self._check(rc, "swm_engine_submit")
        self._keep[jid.value] = (spec.rows, spec.n)
        return jid.value

    def submit_setup(self, c, nsteps, timescheme=5, hp_association=HP_AS_REFERENCE, recfix=4):
        """Submits the reference's initial condition c (a
        problems.ReferenceSetup), its data made on the device; returns the job
        id, or None when it does not fit the memory budget right now."""
        d = JobDesc()
        d.nx, d.ny = c.nx, c.ny
        d.dx, d.dy = c.mesh.dx, c.mesh_y.dx
        d.gravity = c.gravity
        d.cfl = c.cfl
        d.wellbalanced = int(bool(c.wellbalanced))
        for i in range(4):
            d.wgp_bnd[i] = c.mesh.wbnd[i]
        d.nsteps = int(nsteps)
        d.timescheme = int(timescheme)
        d.hp_association = int(hp_association)
        d.reconstruction_fix = int(recfix)
        d.initial_condition = int(c.ic)
        d.bathymetry_flag = int(c.bathy_flag)
        for i in range(2):
            d.mesh_x0[i] = c.x0[i]
            d.mesh_x1[i] = c.x1[i]
        jid = ctypes.c_int(-1)
        rc = _engine_lib().swm_engine_submit(self._e, ctypes.byref(d), ctypes.byref(jid))
        if rc == SWM_E_BUDGET:
            return None
        self._check(rc, "swm_engine_submit")
        self._keep[jid.value] = (c.ny, c.nx)
        return jid.value

    def step(self):
        n = ctypes.c_int(0)
        self._check(_engine_lib().swm_engine_step(self._e, ctypes.byref(n)), "swm_engine_step")
        return n.value

    def run(self, max_steps):
        """Up to max_steps engine steps with one host read per batch; returns
        (jobs still running, steps run)."""
        n, k = ctypes.c_int(0), ctypes.c_int(0)
        self._check(_engine_lib().swm_engine_run(self._e, int(max_steps), ctypes.byref(n),
                                                 ctypes.byref(k)), "swm_engine_run")
        return n.value, k.value

    def info(self, jid):
        i = JobInfo()
        self._check(_engine_lib().swm_engine_job_info(self._e, jid, ctypes.byref(i)),
                    "swm_engine_job_info")
        return {k: getattr(i, k) for k, _ in JobInfo._fields_}

    def fetch(self, jid):
        rows, n = self._keep[jid]
        U = np.empty((rows, n, 3))
        self._check(_engine_lib().swm_engine_fetch(self._e, jid, _ptr(U)), "swm_engine_fetch")
        return U

    def l1_error(self, jid, ic, bathy_flag, x0, x1, t):
        """PostProcessing's L1 error of job jid's state against
        ExactFunction(ic, t)."""
        a0 = np.ascontiguousarray(x0, dtype=np.float64)
        a1 = np.ascontiguousarray(x1, dtype=np.float64)
        err = np.zeros(3)
        self._check(_engine_lib().swm_engine_l1_error(self._e, int(jid), int(ic), int(bathy_flag),
                                                      float(t), _ptr(a0), _ptr(a1), _ptr(err)),
                    "swm_engine_l1_error")
        return err

    def keep_output_state(self, jid, on=True):
        self._check(_engine_lib().swm_engine_keep_output_state(self._e, int(jid), int(bool(on))),
                    "swm_engine_keep_output_state")

    def fetch_output_primitive(self, jid):
        rows, n = self._keep[jid]
        V = np.empty((rows, n, 3))
        self._check(_engine_lib().swm_engine_fetch_output_primitive(self._e, int(jid), _ptr(V)),
                    "swm_engine_fetch_output_primitive")
        return V

    def release(self, jid):
        self._check(_engine_lib().swm_engine_release(self._e, jid), "swm_engine_release")
        self._keep.pop(jid, None)

    def bytes_in_use(self):
        return int(_engine_lib().swm_engine_bytes_in_use(self._e))

    def set_jacobi_limits(self, max_iter, tol):
        self._check(_engine_lib().swm_engine_set_jacobi_limits(self._e, max_iter, tol),
                    "swm_engine_set_jacobi_limits")

    def advance(self, t, t_analyze, t_end, max_steps=1 << 30):
        """The reference's time loop to the next output time (swm_advance):
        returns (t, dts of the steps taken, non-converged solves)."""
        tt = ctypes.c_double(t)
        done = ctypes.c_int(0)
        cap = min(max_steps, 1 << 20)
        dts = np.zeros(cap)
        nonconv = self._check(lib().swm_advance(self._h, ctypes.byref(tt), t_analyze, t_end,
                                                cap, ctypes.byref(done), _ptr(dts)),
                              "swm_advance")
        return tt.value, dts[:done.value].copy(), nonconv

    def set_execution(self, use_graph=-1, margin=-1):
        self._check(_engine_lib().swm_engine_set_execution(self._e, use_graph, margin),
                    "swm_engine_set_execution")

    def execution_stats(self):
        out = (ctypes.c_longlong * 6)()
        self._check(_engine_lib().swm_engine_execution_stats(self._e, out),
                    "swm_engine_execution_stats")
        return dict(zip(EXEC_STATS, list(out)))

    def status_reads(self):
        n = ctypes.c_longlong(0)
        self._check(_engine_lib().swm_engine_status_reads(self._e, ctypes.byref(n)),
                    "swm_engine_status_reads")
        return n.value

    def set_tile_shape(self, shape=-1):
        self._check(_engine_lib().swm_engine_set_tile_shape(self._e, int(shape)),
                    "swm_engine_set_tile_shape")

    def tile_shape(self):
        d = (ctypes.c_int * 4)()
        sh = self._check(_engine_lib().swm_engine_tile_shape(self._e, d), "swm_engine_tile_shape")
        return sh, tuple(d)

    def track_stage_min(self, on=True):
        self._check(_engine_lib().swm_engine_track_stage_min(self._e, int(bool(on))),
                    "swm_engine_track_stage_min")

    def stage_min_h(self, job_id):
        v = ctypes.c_double()
        self._check(_engine_lib().swm_engine_stage_min_h(self._e, job_id, ctypes.byref(v)),
                    "swm_engine_stage_min_h")
        return v.value

    def profile(self, on=True, categories=None):
        if categories is None:
            self._check(_engine_lib().swm_engine_profile_enable(self._e, int(bool(on))),
                        "swm_engine_profile_enable")
        else:
            self._check(_engine_lib().swm_engine_profile_categories(
                self._e, _prof_mask(categories) if on else 0), "swm_engine_profile_categories")

    def profile_read(self):
        ms = (ctypes.c_double * PROF_NCAT)()
        cnt = (ctypes.c_longlong * PROF_NCAT)()
        self._check(_engine_lib().swm_engine_profile_read(self._e, ms, cnt),
                    "swm_engine_profile_read")
        return {lib().swm_profile_name(c).decode(): (ms[c], int(cnt[c])) for c in range(PROF_NCAT)}
