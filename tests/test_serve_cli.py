"""Host logic of the served-stream driver (sw-mpdec_amd/serve.py run as a
program): the sequential queue's latencies from measured service times, and
the comparison of runs result by result."""
from __future__ import annotations

import json

import pytest

import serve


def test_queue_latencies_follow_the_fifo_recursion():
    recs = [{"arrival": 0.0, "service": 2.0}, {"arrival": 1.0, "service": 1.0},
            {"arrival": 1.5, "service": 1.0}, {"arrival": 10.0, "service": 0.5}]
    # finishes at 2, 3, 4 and 10.5
    assert serve.queue_latencies(recs) == [2.0, 2.0, 2.5, 0.5]


def _run(mode, recs, **kw):
    base = {"mode": mode, "jobs": len(recs), "jobs_per_s": 2.0, "makespan_s": 1.0,
            "p50_latency_s": 0.5, "p99_latency_s": 0.9, "peak_running": 2, "peak_bytes": 10,
            "failed": 0, "requests": recs}
    base.update(kw)
    return base


def test_compare_reports_bitwise_agreement_and_speedup(tmp_path, capsys):
    seq = [{"rid": k, "arrival": float(k), "service": 1.0, "sha256": f"h{k}"} for k in range(4)]
    fifo = [{"rid": k, "sha256": f"h{k}"} for k in range(4)]
    bad = [{"rid": k, "sha256": "x" if k == 2 else f"h{k}"} for k in range(4)]
    files = []
    for name, run in (("s0", _run("sequential", seq[:2])), ("s1", _run("sequential", seq[2:])),
                      ("f", _run("fifo", fifo)), ("j", _run("sjf", bad))):
        f = tmp_path / f"{name}.json"
        f.write_text(json.dumps(run))
        files.append(str(f))
    assert serve.compare(files) == 0
    out = json.loads(capsys.readouterr().out)
    assert out["sequential"]["requests"] == 4 and out["sequential"]["jobs_per_s"] == 1.0
    assert out["fifo"]["bitwise_identical_to_sequential"] is True
    assert out["fifo"]["speedup_vs_sequential"] == pytest.approx(2.0)
    assert out["sjf"]["bitwise_identical_to_sequential"] is False
    assert out["fifo"]["results_compared"] == 4


def test_stream_is_the_seeded_2000_request_mix():
    reqs = serve.stream(0, 5)
    again = serve.stream(0, 5)
    assert [(r.problem.n, r.nsteps, r.arrival) for r in reqs] == \
           [(r.problem.n, r.nsteps, r.arrival) for r in again]
    assert all(64 <= r.problem.n <= 1024 and 50 <= r.nsteps <= 500 for r in reqs)
    part = serve.stream(3, 2)
    assert [r.rid for r in part] == [3, 4]
    assert [(r.problem.n, r.nsteps) for r in part] == [(r.problem.n, r.nsteps) for r in reqs[3:]]
