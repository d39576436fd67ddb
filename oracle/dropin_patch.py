#!/usr/bin/env python3
"""Applies the drop-in edits to a copy of the reference's Fortran sources.

    python oracle/dropin_patch.py REF_SRC_DIR OUT_DIR

Copies every .f90 of REF_SRC_DIR (the reference's fv-solver-sw/src) into
OUT_DIR, a scratch directory outside this repository, and edits four of them
so that the reference's own program calls MOD_SwmDropin
(sw-mpdec_amd/fortran/swm_dropin.f90) for the mPDeC5 time step:

  timediscretization.f90  TimeDiscretization: TimeStep() and CASE(5) go to the
                          device behind swm_dropin_on(); the state and the
                          writer's V come back before each WriteSolutionToDisk
                          inside the loop (src/timediscretization.f90:57-95)
  finitevolume.f90        InitializeWBVariables hands the well-balanced state
                          to the device at its end (:339-396);
                          FinalizeFiniteVolume releases the handle (:768)
  equation.f90            Bathymetry_X/_Y become PUBLIC (:49-58), so the
                          module can tabulate them at the Gauss points

Each edit is anchored on text that must occur exactly once; a reference whose
text differs fails here, loudly, instead of building a different program.
Only these anchors and the inserted lines are this script's text; the
reference's sources are read from REF_SRC_DIR at build time and are never
kept in the repository.  This is test infrastructure (it builds the program
tests/test_gpu_dropin.py runs); the patch it applies is the one a maintainer
would make by hand.
"""
from __future__ import annotations

import sys
from pathlib import Path

USE = "USE MOD_SwmDropin\n"

EDITS = {
    "timediscretization.f90": [
        # the module in scope of TimeDiscretization
        ("SUBROUTINE TimeDiscretization()\n",
         "SUBROUTINE TimeDiscretization()\n" + USE),
        # TimeStep on the device
        ("  dt_min = TimeStep()\n",
         "  IF (swm_dropin_on()) THEN\n"
         "    dt_min = swm_dropin_timestep()\n"
         "  ELSE\n"
         "    dt_min = TimeStep()\n"
         "  END IF\n"),
        # the mPDeC5 step on the device
        ("    CASE(5)\n      CALL TimeDiscretizationByMPDeC5(t)\n",
         "    CASE(5)\n"
         "      IF (swm_dropin_on()) THEN\n"
         "        CALL swm_dropin_step(t)\n"
         "      ELSE\n"
         "        CALL TimeDiscretizationByMPDeC5(t)\n"
         "      END IF\n"),
        # U and the writer's V on the host before an output inside the loop
        ("  IF (ABS(tAnalyze-t) .LT. 1.0E-10) THEN\n    CALL WriteSolutionToDisk(t)\n",
         "  IF (ABS(tAnalyze-t) .LT. 1.0E-10) THEN\n"
         "    IF (swm_dropin_on()) CALL swm_dropin_output()\n"
         "    CALL WriteSolutionToDisk(t)\n"),
    ],
    "finitevolume.f90": [
        ("SUBROUTINE InitializeWBVariables()\n",
         "SUBROUTINE InitializeWBVariables()\n" + USE),
        ("END SUBROUTINE InitializeWBVariables\n",
         "IF (swm_dropin_on()) CALL swm_dropin_init_wb()\n"
         "END SUBROUTINE InitializeWBVariables\n"),
        ("SUBROUTINE FinalizeFiniteVolume()\n",
         "SUBROUTINE FinalizeFiniteVolume()\n" + USE),
    ],
    "equation.f90": [
        ("PUBLIC :: Bathymetry\n",
         "PUBLIC :: Bathymetry\nPUBLIC :: Bathymetry_X\nPUBLIC :: Bathymetry_Y\n"),
    ],
}


def patch_finalize(text: str) -> str:
    """Insert CALL swm_dropin_finalize() as FinalizeFiniteVolume's first
    executable statement (after its IMPLICIT NONE and declarations), so the
    handle is released before the module arrays are deallocated."""
    start = text.index("SUBROUTINE FinalizeFiniteVolume()\n")
    end = text.index("END SUBROUTINE FinalizeFiniteVolume\n", start)
    body = text[start:end]
    key = "IMPLICIT NONE\n"
    if body.count(key) != 1:
        raise SystemExit("dropin_patch: FinalizeFiniteVolume has no single IMPLICIT NONE")
    i = body.index(key) + len(key)
    # skip comment / blank / declaration lines up to the first executable one
    lines = body[i:].splitlines(keepends=True)
    k = 0
    while k < len(lines):
        s = lines[k].strip()
        if s == "" or s.startswith("!") or s.upper().startswith(("INTEGER", "REAL", "LOGICAL",
                                                                   "CHARACTER")):
            k += 1
            continue
        break
    new_body = body[:i] + "".join(lines[:k]) + "CALL swm_dropin_finalize()\n" + "".join(lines[k:])
    return text[:start] + new_body + text[end:]


def apply(text: str, name: str) -> str:
    for old, new in EDITS.get(name, []):
        n = text.count(old)
        if n != 1:
            raise SystemExit(f"dropin_patch: {name}: anchor found {n} times: {old!r}")
        text = text.replace(old, new)
    if name == "finitevolume.f90":
        text = patch_finalize(text)
    return text


def main(argv: list[str]) -> int:
    if len(argv) != 3:
        print(__doc__)
        return 2
    src, out = Path(argv[1]), Path(argv[2])
    repo = Path(__file__).resolve().parents[1]
    if out.resolve() == repo or repo in out.resolve().parents:
        raise SystemExit("dropin_patch: OUT_DIR must lie outside the repository")
    out.mkdir(parents=True, exist_ok=True)
    for f in sorted(src.glob("*.f90")):
        text = f.read_text()
        patched = apply(text, f.name)
        (out / f.name).write_text(patched)
    for name in EDITS:
        if not (src / name).exists():
            raise SystemExit(f"dropin_patch: {name} missing in {src}")
    print(f"dropin_patch: {len(list(src.glob('*.f90')))} sources -> {out}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
