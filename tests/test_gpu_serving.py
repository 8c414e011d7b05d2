"""Serving loop on the GPU: continuous batching must give every request the
result it gets when served alone (same populations bit for bit, same step
count), admission must respect the memory budget, and finished requests
must be retired."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def small_requests():
    from fluidloom_amd.serving import Request
    reqs = [
        Request(0, 0.0, "cavity", dict(n0=16, lid=0.08, tau=0.6, max_level=1)),
        Request(1, 0.0, "sphere", dict(nx=32, ny=16, nz=16, r=2.5, re=60.0, u_in=0.05, max_level=2)),
        Request(2, 0.0, "cavity", dict(n0=12, lid=0.05, tau=0.8, max_level=0)),
        Request(3, 0.05, "cavity", dict(n0=20, lid=0.03, tau=0.7, max_level=2)),
    ]
    return reqs


def test_served_alone_equals_in_mix(fl):
    from fluidloom_amd.serving import Server
    mix = Server(storage="f16", check_every=20, tol=1e-7, max_steps=120, keep_results=True)
    done = {r.rid: r for r in mix.run(small_requests())}
    assert sorted(done) == [0, 1, 2, 3] and mix.unserved == 0
    for rid in range(4):
        solo = Server(storage="f16", check_every=20, tol=1e-7, max_steps=120, keep_results=True)
        r = [q for q in small_requests() if q.rid == rid][0]
        r.arrival = 0.0
        (d,) = solo.run([r])
        assert d.steps == done[rid].steps, rid
        assert np.array_equal(d.result["f"], done[rid].result["f"]), rid


def test_admission_respects_budget(fl):
    from fluidloom_amd.serving import Server, prepare
    reqs = small_requests()
    for r in reqs:
        prepare(r, "f16")
    srv = Server(storage="f16", check_every=20, tol=0.0, max_steps=40)
    # room for the largest request but not for all of them at once
    srv.budget = max(r.footprint for r in reqs) + 1
    peak = []
    orig = srv._admit

    def admit(waiting, now):
        n = orig(waiting, now)
        peak.append(srv.used)
        return n

    srv._admit = admit
    done = srv.run(reqs)
    assert len(done) == 4 and srv.unserved == 0
    assert max(peak) <= srv.budget * 1.05
    assert srv.used == 0 and not srv.cohorts
    assert all(d.steps == 40 and not d.converged for d in done)


def test_steady_state_retires(fl):
    from fluidloom_amd.serving import Request, Server
    # a cavity at rest with a resting lid is already steady
    r = Request(0, 0.0, "cavity", dict(n0=8, lid=0.0, tau=0.7, max_level=0))
    srv = Server(storage="f32", check_every=10, tol=1e-6, max_steps=1000)
    (d,) = srv.run([r])
    assert d.converged and d.steps == 10


def test_plain_server_one_at_a_time_same_results(fl):
    """The plain baseline serves one request at a time in arrival order, and
    (FP32 both ways) gives each request the populations and step count that
    continuous batching gives it."""
    from fluidloom_amd.serving import Server, plain_server
    plain = plain_server(check_every=20, tol=1e-7, max_steps=120)
    plain.keep_results = True
    peak = []
    orig = plain._admit

    def admit(waiting, now):
        n = orig(waiting, now)
        peak.append(len(plain.cohorts))
        return n

    plain._admit = admit
    done = plain.run(small_requests())
    assert [d.rid for d in done] == [0, 1, 2, 3] and plain.unserved == 0 and not plain.pending_wait
    assert max(peak) == 1
    for a, b in zip(done, done[1:]):
        assert b.admitted >= a.finished
    mix = Server(storage="f32", check_every=20, tol=1e-7, max_steps=120, keep_results=True)
    batched = {r.rid: r for r in mix.run(small_requests())}
    for d in done:
        assert d.steps == batched[d.rid].steps
        assert np.array_equal(d.result["f"], batched[d.rid].result["f"]), d.rid



def test_served_request_matches_oracle(fl, oracle):
    """Requests served inside a mix (a refined cavity and a refined sphere,
    batched with two other cavities) against the FP64 oracle run of each on
    its own mesh for the same step count: velocity relative L2 within 1e-4
    in FP32 storage and 1e-3 in FP16 storage."""
    from fluidloom_amd.serving import Server
    from tests.lbm_cases import OUT, WALL
    cv, _, _ = oracle.lattice()
    for storage, tol in (("f32", 1e-4), ("f16", 1e-3)):
        mix = Server(storage=storage, check_every=20, tol=0.0, max_steps=200, keep_results=True)
        done = {r.rid: r for r in mix.run(small_requests())}
        for rid in (0, 1):
            d = done[rid]
            s, m = d.spec, d.mesh
            if d.kind == "cavity":
                n = (s["n0"],) * 3
                bcu = [(0.0, 0.0, 0.0)] * 5 + [(s["lid"], 0.0, 0.0)]
                o = oracle.Lbm(*n, s["max_level"], s["tau"], [WALL] * 6, bcu, m["x"], m["y"], m["z"], m["level"],
                               m["state"])
            else:
                n = (s["nx"], s["ny"], s["nz"])
                tau0 = 3 * s["u_in"] * 2 * s["r"] / s["re"] + 0.5
                bcu = [(s["u_in"], 0.0, 0.0)] + [(0.0, 0.0, 0.0)] * 5
                o = oracle.Lbm(*n, s["max_level"], tau0, [WALL, OUT, WALL, WALL, WALL, WALL], bcu, m["x"], m["y"],
                               m["z"], m["level"], m["state"], u_init=(s["u_in"], 0.0, 0.0))
            o.step(d.steps)
            o.fx, o.fy, o.fz, o.flv = (np.ascontiguousarray(a) for a in d.result["fluid"])
            _, uo = o.moments()
            f = d.result["f"]
            u = (f @ cv) / f.sum(1)[:, None]
            err = np.linalg.norm(u - uo) / np.linalg.norm(uo)
            print(f"served {d.kind} ({storage}, {d.steps} steps) vs FP64 oracle: velocity rel L2 {err:.2e}", flush=True)
            assert d.steps == 200 and err < tol, (storage, rid, err)
