"""The C ABI library: it loads without a GPU, exports every entry point that
include/swm.h declares, and rejects bad arguments without touching a device."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import swm

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    txt = (ROOT / "include" / "swm.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(swm_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_hot_path():
    names = declared_functions()
    for must in ("swm_create", "swm_destroy", "swm_set_bathy_grad", "swm_init_wb", "swm_upload",
                 "swm_download", "swm_timestep", "swm_step_mpdec5", "swm_fvtd",
                 "swm_set_jacobi_limits", "swm_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = swm.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_no_reference_or_oracle_symbols_in_product():
    out = Path(swm.LIB_PATH).read_bytes()
    assert b"orc_" not in out  # the oracle is never linked into the product


def test_footprint_is_sane():
    f = swm.footprint_bytes
    a = f(2048, 2048, 1, 1)
    assert 0 < f(64, 64, 0, 0) < f(64, 64, 1, 0) < f(64, 64, 1, 1)
    # about 0.6-1 KB per cell with bathymetry tables
    assert 400 * 2048 * 2048 < a < 1200 * 2048 * 2048
    assert f(2, 5, 0, 0) < 0


def test_create_rejects_bad_arguments():
    L = swm.lib()
    h = ctypes.c_void_p()
    w = np.ones(4)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = L.swm_create(2, 8, 0.1, 0.1, 9.81, 0.5, 0, w.ctypes.data_as(dp), ctypes.byref(h))
    assert rc == -1 and not h.value
    assert b"nx, ny" in L.swm_last_error(None)
    rc = L.swm_create(8, 8, -0.1, 0.1, 9.81, 0.5, 0, w.ctypes.data_as(dp), ctypes.byref(h))
    assert rc == -1
    assert L.swm_timestep(None, None, None, None) == -1


def test_profile_names():
    L = swm.lib()
    names = [L.swm_profile_name(c).decode() for c in range(swm.PROF_NCAT)]
    assert names[0] == "fvtd_x" and names[5] == "jacobi_sweep"
    assert tuple(names) == swm.PROF_NAMES


def test_profile_categories_arguments():
    """The category mask names categories by bit; a null handle is refused
    before anything touches the GPU."""
    L = swm.lib()
    assert L.swm_profile_categories(None, 3) == -1
    assert L.swm_engine_profile_categories(None, 3) == -1
    assert swm._prof_mask(["fvtd_x", "fvtd_y"]) == 3
    assert swm._prof_mask(["other"]) == 1 << 7
    with pytest.raises(ValueError):
        swm._prof_mask(["fvtd_z"])


def test_header_declares_the_rows_around_the_path():
    names = declared_functions()
    for must in ("swm_set_timescheme", "swm_set_hp_association", "swm_set_reconstruction_fix",
                 "swm_fill_initial_data", "swm_l1_error", "swm_download_primitive",
                 "swm_download_primitive_async", "swm_output_wait", "swm_engine_l1_error",
                 "swm_run_steps", "swm_advance"):
        assert must in names


def test_null_handles_are_rejected_without_a_device():
    """The new entry points check their handle before any HIP call."""
    L = swm._engine_lib()
    x = (ctypes.c_double * 2)(0.0, 0.0)
    err = (ctypes.c_double * 3)()
    assert L.swm_set_reconstruction_fix(None, 2) == swm.SWM_E_ARG
    assert L.swm_fill_initial_data(None, 1, 0, x, x) == swm.SWM_E_ARG
    assert L.swm_l1_error(None, 1, 0, ctypes.c_double(0.0), x, x, err) == swm.SWM_E_ARG
    assert L.swm_download_primitive(None, err) == swm.SWM_E_ARG
    assert L.swm_download_primitive_async(None, err) == swm.SWM_E_ARG
    assert L.swm_output_wait(None) == swm.SWM_E_ARG
    assert L.swm_engine_l1_error(None, 0, 1, 0, ctypes.c_double(0.0), x, x, err) == swm.SWM_E_ARG


def test_reference_setups_match_the_reference_mesh():
    """problems.ReferenceSetup carries the reference's domains (src/parameters.f90:62-125)
    and Mesh_DX = |X1 - X0| / n (src/mesh.f90:39-41)."""
    import problems as P
    for ic, (x0, x1, g, bf) in P.REFERENCE_SETUPS.items():
        c = P.ReferenceSetup(ic, 40, 24)
        assert c.mesh.dx == abs(abs(x1[0] - x0[0])) / 40.0
        assert c.mesh_y.dx == abs(abs(x1[1] - x0[1])) / 24.0
        assert (c.n, c.rows) == (40, 24) and c.gravity == g and c.bathy_flag == bf
    assert P.REFERENCE_SETUPS[3][0] == (-5.0, -2.0) and P.REFERENCE_SETUPS[5][1] == (50.0, 50.0)


def test_reference_request_mix_is_seeded():
    import serve
    a = serve.reference_mix(12, 3, rate=5.0)
    b = serve.reference_mix(12, 3, rate=5.0)
    assert [(r.problem.ic, r.problem.nx, r.nsteps, r.arrival) for r in a] == \
           [(r.problem.ic, r.problem.nx, r.nsteps, r.arrival) for r in b]
    assert {r.problem.ic for r in a} <= set(range(1, 7))
    assert all(r.problem.wellbalanced == (r.problem.ic in (2, 3)) for r in a)


def test_timestep_noalloc_rejects_bad_arguments_without_a_device():
    """swm_timestep_noalloc checks sizes, ghosts and pointers before any HIP
    call; a grid above its fixed scratch array is refused with a message."""
    L = swm.lib()
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: every case fails its checks first
    big = swm.TIMESTEP_NOALLOC_MAX_CELLS
    assert L.swm_timestep_noalloc(4097, 4096, .1, .1, 9.81, .5, fake, 0, fake, None) == swm.SWM_E_ARG
    msg = L.swm_last_error(None).decode()
    assert "4097 x 4096" in msg and str(big) in msg and "scratch" in msg
    assert L.swm_timestep_noalloc(0, 8, .1, .1, 9.81, .5, fake, 0, fake, None) == swm.SWM_E_ARG
    assert L.swm_timestep_noalloc(8, 8, .1, .1, 9.81, .5, fake, 2, fake, None) == swm.SWM_E_ARG
    assert b"ghosts" in L.swm_last_error(None)
    assert L.swm_timestep_noalloc(8, 8, .1, .1, 9.81, .5, None, 0, fake, None) == swm.SWM_E_ARG
    assert L.swm_timestep_noalloc(8, 8, .1, .1, 9.81, .5, fake, 0, None, None) == swm.SWM_E_ARG
    assert 4096 * 4096 == big  # the largest grid of the GPU tests fits


def test_sync_cost_checks_arguments_before_the_device():
    L = swm.lib()
    out = (ctypes.c_double * 3)()
    assert L.swm_debug_sync_cost(0, 10, out) == swm.SWM_E_ARG
    assert L.swm_debug_sync_cost(8, 0, out) == swm.SWM_E_ARG
    assert L.swm_debug_sync_cost(8, 100001, out) == swm.SWM_E_ARG
    assert b"rounds" in L.swm_last_error(None)
