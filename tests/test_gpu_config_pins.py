"""GPU parity at the benchmark configurations' full sizes, against hash pins
the C oracle wrote (tests/golden/config_pins.json, by make_config_pins.py),
and the benchmark families made on the device against the host builder.

* The families' exponential: the device's (swm_debug_exp) is bitwise the
  host builder's det_exp on 10^7 arguments over its whole range.
* Family data: the state, the well-balanced state and the bathymetry tables
  that swm_fill_family makes are bitwise the host builder's (directly for the
  state; through the FVTD outputs and a step for the others).
* 2048^2 lake over bumps (the headline workload): the host-built inputs hash
  to the pinned values (so this host rebuilt them bitwise); eight mPDeC5
  steps from them, and from the same data made on the device, give the
  oracle's dt, its 104 Jacobi sweep counts and its state hashes after every
  step.  This covers the certified Jacobi stop at n = 4.19M cells, the
  speculative steps after the first, and the 64-bit indexing at that size.
* 1024^2 dam break over a dry bed: the oracle's (and the reference
  Fortran's, equal to it) first step, NaN velocities included.
"""
from __future__ import annotations

import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

import problems as P
import swm
from parity_util import describe_diff, same_bits

pytestmark = pytest.mark.gpu

PINS = json.loads((Path(__file__).resolve().parent / "golden" / "config_pins.json").read_text())
NAN = np.frombuffer(np.uint64(0x7FF8000000000000).tobytes(), dtype=np.float64)[0]


def sha(a) -> str:
    a = np.ascontiguousarray(a, dtype=np.float64)
    if np.isnan(a).any():
        a = np.where(np.isnan(a), NAN, a)
    return hashlib.sha256(a.tobytes()).hexdigest()


def test_family_exp_bitwise_host_and_device():
    rng = np.random.default_rng(77)
    x = np.concatenate([rng.uniform(-746.5, 710.0, 4_000_000), rng.uniform(-60.0, 2.0, 4_000_000),
                        rng.uniform(-0.4, 0.4, 2_000_000),
                        np.array([0.0, -0.0, 709.78, 709.8, 709.80001, -745.13, -745.14, -746.0,
                                  -746.0001, -708.4, 1e-300, -1e-300, np.inf, -np.inf, np.nan])])
    y_host = P.det_exp(x)
    y_dev = swm.debug_exp(x)
    assert same_bits(y_dev, y_host), describe_diff(y_dev, y_host)


FAMILY_CASES = {
    "vortex_37x23": lambda: P.vortex_spec(37, 5, ny=23),
    "lake_40": lambda: P.lake_bumps_spec(40, 9),
    "lake_33_amp0": lambda: P.lake_bumps_spec(33, 10, amp=0.0),
    "dam_dry_31": lambda: P.dam_break_spec(31),
    "dam_film_28": lambda: P.dam_break_spec(28, dry=0.1),
}


@pytest.mark.parametrize("case", sorted(FAMILY_CASES))
def test_family_data_bitwise_host_builder(case):
    spec = FAMILY_CASES[case]()
    p = spec.build()
    dev = swm.Solver.from_family(spec)
    host = swm.Solver.from_problem(p)
    U = dev.download()
    assert same_bits(U, p.U0), describe_diff(U, p.U0)
    # the tables and the well-balanced state enter S, FX, FY and the PDS
    a, b = dev.fvtd(), host.fvtd()
    for k in a:
        assert same_bits(a[k], b[k]), (k, describe_diff(a[k], b[k]))
    dt_a, _, _ = dev.timestep()
    it_a, _ = dev.step(dt_a)
    dt_b, _, _ = host.timestep()
    it_b, _ = host.step(dt_b)
    assert same_bits([dt_a], [dt_b]) and it_a == it_b
    Ua, Ub = dev.download(), host.download()
    assert same_bits(Ua, Ub), describe_diff(Ua, Ub)


def test_family_rejects_a_non_unit_domain():
    spec = P.vortex_spec(16, 1)
    p = P.lake_flag1(16, False)  # unit square, but set up with another dx below
    s = swm.Solver(16, 16, 0.5 / 16, 1.0 / 16, 9.81, 0.5, False, p.mesh.wbnd)
    with pytest.raises(swm.SwmError, match="unit square"):
        s.fill_family(spec)
    wb = swm.Solver(16, 16, 1.0 / 16, 1.0 / 16, 9.81, 0.5, True, p.mesh.wbnd)
    with pytest.raises(swm.SwmError, match="well-balanced"):
        wb.fill_family(spec)


def _run_pinned(s, pin):
    for k, ref in enumerate(pin["steps"]):
        dt, _, _ = s.timestep()
        iters, nonconv = s.step(dt)
        assert float(dt).hex() == ref["dt"], (k, float(dt).hex(), ref["dt"])
        assert list(iters) == ref["iters"], (k, list(iters), ref["iters"])
        assert nonconv == ref["nonconv"]
        U = s.download()
        assert sha(U) == ref["U"], f"state after step {k + 1} differs from the oracle's"
    return U


def test_lake_bumps_2048_matches_oracle_pins():
    pin = PINS["lake_bumps_2048"]
    spec = P.lake_bumps_spec(2048, 2024, steps=100)
    assert spec.params == pin["params"]
    p = spec.build()
    for k, h in pin["inputs"].items():
        assert sha(getattr(p, k)) == h, f"{k}: this host did not rebuild the inputs bitwise"
    s = swm.Solver.from_problem(p)
    del p
    _run_pinned(s, pin)
    s.close()
    # the same data made on the device
    d = swm.Solver.from_family(spec)
    assert sha(d.download()) == pin["inputs"]["U0"]
    _run_pinned(d, pin)
    assert d.ambiguous_decisions() >= 0


def test_dam_break_1024_dry_bed_matches_oracle_and_reference_pins():
    pin = PINS["dam_break_1024"]
    spec = P.dam_break_spec(1024, steps=1)
    p = spec.build()
    assert sha(p.U0) == pin["inputs"]["U0"]
    for s in (swm.Solver.from_problem(p), swm.Solver.from_family(spec)):
        U = _run_pinned(s, pin)
        ref = pin["steps"][0]
        assert sha(U[..., 0]) == ref["h"]
        assert int(np.isnan(U).sum()) == ref["nan_count"]
        assert float(np.nanmin(U[..., 0])).hex() == ref["h_min"]
        s.close()
