"""Host-side parts of the serving loop (no GPU): the arrival stream."""
import numpy as np

from psblas3_amd.serve import latency_summary, poisson_arrivals


def test_poisson_arrivals_statistics():
    req = poisson_arrivals(1000.0, 20000, 64, seed=4)
    t = np.array([r[0] for r in req])
    assert np.all(np.diff(t) > 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    # exponential gaps: mean 1/rate, coefficient of variation 1
    assert abs(gaps.mean() * 1000.0 - 1.0) < 0.03
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.05
    m = np.array([r[1] for r in req])
    assert m.min() == 0 and m.max() == 63
    counts = np.bincount(m, minlength=64)
    assert counts.min() > 0.7 * 20000 / 64 and counts.max() < 1.3 * 20000 / 64
    assert len({r[2] for r in req}) == len(req)


def test_poisson_arrivals_deterministic():
    assert poisson_arrivals(200.0, 50, 8) == poisson_arrivals(200.0, 50, 8)
    assert poisson_arrivals(200.0, 50, 8) != poisson_arrivals(200.0, 50, 8, seed=5)


def test_latency_summary():
    rec = [{"arrival": 0.0, "admitted": 0.0, "done": 0.010, "latency": 0.010, "status": 1,
            "relres": 1e-11, "iterations": 10},
           {"arrival": 0.005, "admitted": 0.011, "done": 0.020, "latency": 0.015, "status": 1,
            "relres": 2e-11, "iterations": 20},
           {"arrival": 0.006, "status": 0}]
    s = latency_summary(rec)
    assert s["finished"] == 2 and s["requests"] == 3 and s["converged"] == 2
    assert abs(s["p50_ms"] - 12.5) < 1e-9
    assert abs(s["solves_per_s"] - 100.0) < 1e-9
    # latency = queue wait (arrival -> admitted) + service (admitted -> done)
    assert abs(s["queue_wait_p50_ms"] - 3.0) < 1e-9
    assert abs(s["service_p50_ms"] - 9.5) < 1e-9
