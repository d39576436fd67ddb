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
