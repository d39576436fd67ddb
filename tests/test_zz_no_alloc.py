"""swm_timestep_noalloc (TimeStep with no device allocation) against
swm_timestep on the same states, through the C ABI: two calls in a row, the
second on a smaller grid (and with the module array's ghost layers), then a
state with a NaN wave speed, where both replay the reference's sequential
fold.  Device buffers come from the HIP runtime libswm.so itself uses."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import problems as P
import swm
from parity_util import same_bits

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libswm.so is linked to."""

    def __init__(self):
        swm.lib()
        self.L = ctypes.CDLL("libamdhip64.so.7")
        vp = ctypes.c_void_p
        self.L.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        self.L.hipFree.argtypes = [vp]
        self.L.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, n):
        out = np.empty(n)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self, p):
        assert self.L.hipFree(p) == 0


def with_ghosts(U):
    """(rows, n, 3) -> the module array U(1:3, -2:n+3, -2:rows+3), periodic."""
    return np.pad(U, ((3, 3), (3, 3), (0, 0)), mode="wrap")


def noalloc_timestep(hip, p, U, ghosts):
    m = p.mesh
    dU, dout = hip.put(U), hip.alloc(3 * 8)
    try:
        rc = swm.lib().swm_timestep_noalloc(p.n, p.rows, m.dx, p.mesh_y.dx, p.gravity, p.cfl,
                                            dU, ghosts, dout, None)
        assert rc == swm.SWM_OK, swm.lib().swm_last_error(None)
        assert hip.L.hipDeviceSynchronize() == 0
        return hip.get(dout, 3)
    finally:
        hip.free(dU)
        hip.free(dout)


def handle_timestep(p, U):
    s = swm.Solver.from_problem(p)
    s.upload(U)
    return np.array(s.timestep())


def test_noalloc_timestep_matches_swm_timestep():
    hip = Hip()
    # first call: 4.2M cells, the state made on the device
    p1 = P.vortex_spec(2048, 31)
    s1 = swm.Solver.from_family(p1)
    ref1 = np.array(s1.timestep())
    got1 = noalloc_timestep(hip, p1, s1.download(), 0)
    del s1
    assert same_bits(got1, ref1), (got1, ref1)
    # second call, smaller, from the module array with its ghost layers
    p2 = P.lake_bumps(40, 9)
    ref2 = handle_timestep(p2, p2.U0)
    got2 = noalloc_timestep(hip, p2, with_ghosts(p2.U0), 3)
    assert same_bits(got2, ref2), (got2, ref2)
    assert not same_bits(got2, got1)
    # a negative depth makes sqrt(g*h) NaN: both take the sequential fold
    p3 = P.vortex(37, 8, ny=23)
    U3 = p3.U0.copy()
    U3[11, 20, 0] = -0.25
    ref3 = handle_timestep(p3, U3)
    got3 = noalloc_timestep(hip, p3, U3, 0)
    assert same_bits(got3, ref3), (got3, ref3)
