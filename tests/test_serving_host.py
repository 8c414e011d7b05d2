"""Serving host logic that needs no GPU: latency accounting."""
from fluidloom_amd.serving import Request, request_specs, summary


def test_unfinished_requests_enter_latency_as_lower_bounds():
    done = [Request(0, 0.0, "cavity", {}, admitted=0.0, finished=1.0)]
    s = summary(done, 10.0, 3, pending_wait=[5.0, 6.0, 7.0])
    assert s["served"] == 1 and s["censored"] == 3
    assert s["p50_s"] == 5.5 and s["p99_s"] > 6.9


def test_request_trace_is_seeded_and_sized():
    a = request_specs(4, 60, seed=42, scale=2)
    b = request_specs(4, 60, seed=42, scale=2)
    assert [(r.arrival, r.kind, r.spec) for r in a] == [(r.arrival, r.kind, r.spec) for r in b]
    assert all(r.arrival < 60 for r in a) and 150 < len(a) < 330
    for r in a:
        dims = [r.spec["n0"]] if r.kind == "cavity" else [r.spec["nx"], r.spec["ny"], r.spec["nz"]]
        assert all(d % 4 == 0 and d >= 8 for d in dims)
