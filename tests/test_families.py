"""The benchmark families' host builder (problems.py) on the CPU: its
exponential against a 60-digit exp, the family formulas against an
independent point-by-point evaluation with libm, the bathymetry draw's
margin, and the pinned hashes of the headline inputs (host independence)."""
from __future__ import annotations

import hashlib
import json
import math
from decimal import Decimal, getcontext
from pathlib import Path

import numpy as np
import pytest

import problems as P

PINS = json.loads((Path(__file__).resolve().parent / "golden" / "config_pins.json").read_text())


def ulps(x: float, y: float) -> float:
    """|y - exp(x)| in ulps of y, exp(x) to 60 digits."""
    getcontext().prec = 60
    e = Decimal(x).exp()
    return float(abs(Decimal(y) - e) / Decimal(math.ulp(y)))


def test_det_exp_is_faithful():
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-700.0, 709.0, 1500), rng.uniform(-40.0, 1.0, 1500),
                         rng.uniform(-0.35, 0.35, 1500), [0.0, 1.0, -1.0, 0.5, 1e-10, -1e-10]])
    ys = P.det_exp(xs)
    worst = max(ulps(float(a), float(b)) for a, b in zip(xs, ys))
    assert worst < 0.85, worst
    assert P.det_exp(0.0) == 1.0 and P.det_exp(-0.0) == 1.0


def test_det_exp_edges():
    x = np.array([710.0, 709.8, 709.79, np.inf, -np.inf, -746.5, -800.0, np.nan, -745.0, -740.0])
    y = P.det_exp(x)
    assert y[0] == np.inf and y[1] == np.inf and y[2] == np.inf and y[3] == np.inf
    assert y[4] == 0.0 and y[5] == 0.0 and y[6] == 0.0 and np.isnan(y[7])
    # subnormal results: the single rounding of p * 2^k, within 1 ulp of libm
    assert 0.0 < y[8] < 2.3e-308 and abs(y[8] - math.exp(-745.0)) <= 5e-324
    assert abs(y[9] - math.exp(-740.0)) <= 2 * 5e-324
    # libm agrees to 1 ulp everywhere on the normal range
    xs = np.random.default_rng(4).uniform(-708.0, 709.0, 200_000)
    lib = np.array([math.exp(v) for v in xs])
    assert np.max(np.abs(P.det_exp(xs) - lib) / np.spacing(lib)) <= 1.0


def test_vortex_matches_a_pointwise_libm_evaluation():
    spec = P.vortex_spec(9, 17, ny=7)
    p = spec.build()
    eps, g = spec.params["eps"], P.GRAVITY
    nodes, w = P.gauss4()
    for (j, i) in [(0, 0), (3, 4), (6, 8)]:
        acc = np.zeros(3)
        for ig in range(4):
            for jg in range(4):
                x = (i + 0.5) / 9 + nodes[ig] / 9
                y = (j + 0.5) / 7 + nodes[jg] / 7
                xp, yp = (x - 0.5) / 0.1, (y - 0.5) / 0.1
                r2 = xp * xp + yp * yp
                h = 1.0 - eps * eps / (2 * g) * math.exp(1.0 - r2)
                s = eps * math.exp(0.5 * (1.0 - r2))
                acc += w[ig] * w[jg] * np.array([h, h * (-s * yp), h * (s * xp)])
        assert np.allclose(p.U0[j, i], acc, rtol=1e-14, atol=1e-15), (i, j, p.U0[j, i], acc)


def test_lake_bumps_matches_a_pointwise_libm_evaluation():
    spec = P.lake_bumps_spec(6, 23, amp=0.01)
    p = spec.build()
    q = spec.params
    nodes, w = P.gauss4()

    def G(x, c, s):
        return sum(math.exp(-((x - c - a) ** 2) / s ** 2) for a in (-1, 0, 1))

    def dG(x, c, s):
        return sum(math.exp(-((x - c - a) ** 2) / s ** 2) * (-2 * (x - c - a) / s ** 2)
                   for a in (-1, 0, 1))

    def b(x, y):
        return sum(A * G(x, cx, s) * G(y, cy, s) for A, cx, cy, s in
                   zip(q["A"], q["cx"], q["cy"], q["s"]))

    for (j, i) in [(0, 0), (2, 5), (5, 3)]:
        h0 = hw = 0.0
        for ig in range(4):
            for jg in range(4):
                x, y = (i + 0.5) / 6 + nodes[ig] / 6, (j + 0.5) / 6 + nodes[jg] / 6
                bb = b(x, y)
                hw += w[ig] * w[jg] * (1 - bb)
                h0 += w[ig] * w[jg] * (1 - bb + q["amp"] * G(x, q["px"], q["ps"]) *
                                       G(y, q["py"], q["ps"]))
                bx = sum(A * dG(x, cx, s) * G(y, cy, s) for A, cx, cy, s in
                         zip(q["A"], q["cx"], q["cy"], q["s"]))
                by = sum(A * G(x, cx, s) * dG(y, cy, s) for A, cx, cy, s in
                         zip(q["A"], q["cx"], q["cy"], q["s"]))
                assert math.isclose(p.dbdx[jg * 4 + ig, j, i], bx, rel_tol=1e-13, abs_tol=1e-14)
                assert math.isclose(p.dbdy[jg * 4 + ig, j, i], by, rel_tol=1e-13, abs_tol=1e-14)
        assert math.isclose(p.U0[j, i, 0], h0, rel_tol=1e-14)
        assert math.isclose(p.Uwb[j, i, 0], hw, rel_tol=1e-14)
        assert p.U0[j, i, 1] == 0.0 and p.U0[j, i, 2] == 0.0


def test_bump_draws_keep_the_margin():
    for seed in range(12):
        spec = P.lake_bumps_spec(8, seed)
        q = spec.params
        gb = P.GaussBumps(np.array(q["A"]), np.array(q["cx"]), np.array(q["cy"]),
                          np.array(q["s"]))
        xs = np.linspace(0.0, 1.0, 2049)  # finer than the draw's own check
        assert gb.value(xs[None, :], xs[:, None]).max() < 0.95 + 4e-5


def test_dam_break_family():
    p = P.dam_break_spec(20, dry=0.1).build()
    # the Gauss weights sum to 1 within rounding
    assert abs(p.U0[..., 0].max() - 1.0) <= 4e-16 and abs(p.U0[..., 0].min() - 0.1) <= 4e-17
    assert not p.U0[..., 1:].any()


@pytest.mark.slow
def test_headline_inputs_rebuild_to_the_pinned_hashes():
    pin = PINS["lake_bumps_2048"]
    spec = P.lake_bumps_spec(2048, 2024, steps=100)
    assert spec.params == pin["params"]
    p = spec.build()
    for k, h in pin["inputs"].items():
        a = np.ascontiguousarray(getattr(p, k), dtype=np.float64)
        assert hashlib.sha256(a.tobytes()).hexdigest() == h, k
