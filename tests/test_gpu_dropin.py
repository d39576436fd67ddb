"""The drop-in on the reference's own program.

oracle/Makefile's dropin target builds sw-mpdec's program (its main.f90,
parameters.f90 and modules) twice: unchanged (swmain_*), and from a copy
edited by oracle/dropin_patch.py so that TimeDiscretization,
InitializeWBVariables and FinalizeFiniteVolume call
sw-mpdec_amd/fortran/swm_dropin.f90, which runs TimeStep and the mPDeC5 step
through libswm (swdropin_*).  Each case runs the reference's own initial
condition at a small size (command line: InitialCondition nElemsX) and checks

* with the switch off (SWM_DROPIN=0) the edited program writes every file of
  the unchanged one byte for byte (the edits change nothing else);
* on the GPU it writes the same files byte for byte (Tecplot solutions and
  errors at the 4 output times, ErrorL1), and its final state U is bitwise
  the CPU run's (raw dump of SWM_DROPIN_DUMP).
"""
from __future__ import annotations

import filecmp
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from parity_util import describe_diff, same_bits

REF = Path(__file__).resolve().parents[1] / "oracle" / "_ref"

# (build, InitialCondition, nElemsX): the vortex, the lake at rest over
# BathymetryFlag 1 (well-balanced build), a circular dam break and the
# wet-dry lake over BathymetryFlag 2 (well-balanced build)
CASES = [("p", 1, 48), ("pwb", 2, 32), ("p", 4, 40), ("pwb", 3, 40)]


def programs(build):
    main, dropin = REF / f"swmain_{build}", REF / f"swdropin_{build}"
    if not (main.exists() and dropin.exists()):
        pytest.skip("the reference programs are not built (oracle/Makefile target dropin, "
                    "which needs the reference sources and amdflang)")
    return main, dropin


def run(exe, ic, n, workdir, env=None):
    workdir.mkdir(parents=True, exist_ok=True)
    e = dict(os.environ)
    e.update(env or {})
    r = subprocess.run([str(exe), str(ic), str(n)], cwd=workdir, env=e, capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, f"{exe.name} {ic} {n}: rc={r.returncode}\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}"
    return {p.name: p for p in workdir.iterdir() if p.name != "U.bin"}


def same_files(a, b):
    assert sorted(a) == sorted(b), (sorted(a), sorted(b))
    differ = [k for k in a if not filecmp.cmp(a[k], b[k], shallow=False)]
    assert not differ, differ
    assert any(k.startswith("SOLUTION_TEC") for k in a) and any(k.startswith("ErrorL1") for k in a)


@pytest.mark.parametrize("build,ic,n", CASES, ids=[f"{b}_ic{i}_{n}" for b, i, n in CASES])
def test_switch_off_is_the_reference_program(build, ic, n, tmp_path):
    main, dropin = programs(build)
    same_files(run(main, ic, n, tmp_path / "main"),
               run(dropin, ic, n, tmp_path / "cpu", {"SWM_DROPIN": "0"}))


@pytest.mark.gpu
@pytest.mark.parametrize("build,ic,n", CASES, ids=[f"{b}_ic{i}_{n}" for b, i, n in CASES])
def test_dropin_on_the_gpu_writes_the_reference_files(build, ic, n, tmp_path):
    main, dropin = programs(build)
    print(f"dropin case {build} ic={ic} n={n}", flush=True)
    ref = run(main, ic, n, tmp_path / "main")
    cpu = run(dropin, ic, n, tmp_path / "cpu", {"SWM_DROPIN": "0", "SWM_DROPIN_DUMP": "U.bin"})
    gpu = run(dropin, ic, n, tmp_path / "gpu", {"SWM_DROPIN": "1", "SWM_DROPIN_DUMP": "U.bin"})
    same_files(ref, cpu)
    same_files(ref, gpu)
    U_cpu = np.fromfile(tmp_path / "cpu" / "U.bin", dtype=np.float64)
    U_gpu = np.fromfile(tmp_path / "gpu" / "U.bin", dtype=np.float64)
    assert U_cpu.size == 3 * n * n  # the command line's size sets nElemsX = nElemsY
    assert same_bits(U_gpu, U_cpu), describe_diff(U_gpu, U_cpu)
