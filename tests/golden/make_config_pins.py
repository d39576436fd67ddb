"""Writes config_pins.json: hash pins of two benchmark configurations at their
full sizes, where a fixture of the arrays themselves would be too large.

* 2048 x 2048 lake at rest over 8 seeded Gaussian bumps (seed 2024) with a
  1e-3 surface perturbation, well-balanced, CFL 0.5 (the headline workload):
  SHA-256 of the inputs U0, Uwb, dbdx, dbdy as the host problem builder makes
  them, then LAKE_STEPS mPDeC5 steps of the C oracle: each step's dt, its 13
  Jacobi sweep counts and the SHA-256 of the state after it.
* 1024 x 1024 circular dam break over a dry bed (h = 1 inside radius 0.25,
  h = 0 outside): SHA-256 of U0, then one mPDeC5 step of the C oracle AND of
  the reference Fortran itself (oracle/_ref/swref_p, built unchanged from its
  sources); the script fails unless the two agree bitwise.  The velocities
  turn NaN in that step (the reference's own behaviour), so the state hash
  takes every NaN as one canonical pattern, and the NaN count is recorded.

The inputs depend only on IEEE arithmetic (the families' exponential is the
fixed method of problems.det_exp), so any host rebuilds them bitwise; the GPU
test checks the input hashes first.  Run from the repository root after
__graft_entry__.build() (about 25 minutes on one core):

    python tests/golden/make_config_pins.py
"""
from __future__ import annotations

import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT / "sw-mpdec_amd"))
sys.path.insert(0, str(ROOT / "tests"))

import problems as P  # noqa: E402
from oracle_api import Oracle, run_swref  # noqa: E402

OUT = HERE / "config_pins.json"
LAKE_STEPS = 8
NAN = np.frombuffer(np.uint64(0x7FF8000000000000).tobytes(), dtype=np.float64)[0]


def sha(a) -> str:
    """SHA-256 of a float64 array's bytes (C order), NaNs made canonical."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    if np.isnan(a).any():
        a = np.where(np.isnan(a), NAN, a)
    return hashlib.sha256(a.tobytes()).hexdigest()


def lake_2048() -> dict:
    t0 = time.time()
    spec = P.lake_bumps_spec(2048, 2024, steps=100)
    p = spec.build()
    rec = {"workload": "lake_bumps 2048x2048 seed 2024, well-balanced, CFL 0.5",
           "params": spec.params,
           "inputs": {k: sha(getattr(p, k)) for k in ("U0", "Uwb", "dbdx", "dbdy")}}
    print(f"lake 2048: inputs built in {time.time() - t0:.1f} s", flush=True)
    o = Oracle(p)
    steps = []
    for k in range(LAKE_STEPS):
        t0 = time.time()
        dt, lmx, lmy = o.timestep()
        iters, nonconv = o.step(dt)
        U = o.get_U()
        steps.append({"dt": float(dt).hex(), "iters": [int(i) for i in iters],
                      "nonconv": int(nonconv), "U": sha(U)})
        print(f"lake 2048: oracle step {k + 1} dt={dt!r} iters={list(iters)} "
              f"({time.time() - t0:.1f} s)", flush=True)
    rec["steps"] = steps
    return rec


def dam_1024() -> dict:
    p = P.dam_break(1024, steps=1)
    rec = {"workload": "dam_break 1024x1024 dry bed (h = 0 outside r = 0.25), CFL 0.5",
           "inputs": {"U0": sha(p.U0)}}
    t0 = time.time()
    o = Oracle(p)
    dt, _, _ = o.timestep()
    iters, nonconv = o.step(dt)
    U = o.get_U()
    print(f"dam 1024: oracle step dt={dt!r} iters={list(iters)} ({time.time() - t0:.1f} s)",
          flush=True)
    t0 = time.time()
    hist, U_ref = run_swref(p, 2, steps=1, timeout=3600)
    print(f"dam 1024: reference step ({time.time() - t0:.1f} s)", flush=True)
    dt_ref = hist[0][0]
    assert float(dt_ref).hex() == float(dt).hex(), (dt_ref, dt)
    assert sha(U_ref) == sha(U), "oracle and reference differ on the 1024 dam break"
    rec["steps"] = [{"dt": float(dt).hex(), "iters": [int(i) for i in iters],
                     "nonconv": int(nonconv), "U": sha(U), "h": sha(U[..., 0]),
                     "nan_count": int(np.isnan(U).sum()),
                     "h_min": float(np.nanmin(U[..., 0])).hex(),
                     "reference_agrees": True}]
    return rec


def main():
    pins = {"lake_bumps_2048": lake_2048(), "dam_break_1024": dam_1024()}
    OUT.write_text(json.dumps(pins, indent=1) + "\n")
    print(f"wrote {OUT}", flush=True)


if __name__ == "__main__":
    main()
