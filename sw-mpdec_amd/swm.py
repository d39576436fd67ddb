"""ctypes binding of libswm.so, the C ABI of include/swm.h.

The product path: every call goes to the HIP library.  There is no CPU
fallback -- importing this module on a machine without the built library, or
creating a solver without a GPU, raises.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "lib" / "libswm.so"

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

SWM_OK = 0
PROF_NCAT = 8


class SwmError(RuntimeError):
    pass


def _ptr(a):
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_dp)


_lib = None


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
    L.swm_run_steps.argtypes = [vp, ctypes.c_int, _dp, _ip]
    L.swm_fvtd.argtypes = [vp] + [_dp] * 6
    L.swm_set_jacobi_limits.argtypes = [vp, ctypes.c_int, ctypes.c_double]
    L.swm_ambiguous_decisions.restype = ctypes.c_longlong
    L.swm_ambiguous_decisions.argtypes = [vp]
    L.swm_sync.argtypes = [vp]
    L.swm_debug_divsqrt.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
    L.swm_profile_enable.argtypes = [vp, ctypes.c_int]
    L.swm_profile_read.argtypes = [vp, _dp, ctypes.POINTER(ctypes.c_longlong)]
    L.swm_profile_name.restype = ctypes.c_char_p
    L.swm_profile_name.argtypes = [ctypes.c_int]
    _lib = L
    return L


def footprint_bytes(nx, ny, wellbalanced, bathymetry) -> int:
    return int(lib().swm_footprint_bytes(nx, ny, int(bool(wellbalanced)), int(bool(bathymetry))))


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

    @classmethod
    def from_problem(cls, p):
        m = p.mesh
        s = cls(p.n, p.n, m.dx, m.dx, p.gravity, p.cfl, p.wellbalanced, m.wbnd)
        if p.dbdx is not None:
            s.set_bathy_grad(p.dbdx, p.dbdy)
        if p.wellbalanced:
            s.init_wb(p.Uwb)
        s.upload(p.U0)
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

    def step(self, dt, t=0.0):
        it = (ctypes.c_int * 13)()
        nonconv = self._check(lib().swm_step_mpdec5(self._h, t, dt, it), "swm_step_mpdec5")
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
        dts = np.zeros(steps)
        its = np.zeros(13 * steps, dtype=np.int32)
        nonconv = self._check(lib().swm_run_steps(self._h, steps, _ptr(dts),
                                                  its.ctypes.data_as(_ip)), "swm_run_steps")
        return dts, its.reshape(steps, 13), nonconv

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

    def profile(self, on=True):
        self._check(lib().swm_profile_enable(self._h, int(bool(on))), "swm_profile_enable")

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
