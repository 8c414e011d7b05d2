"""The sphere workload at its full size (base 256x128x128, r = 16, three
levels within 2r: 76.4M leaves), checked through properties that do not
depend on size:
  * the layout holds exactly the mesh's fluid leaves per level;
  * graph replay and eager stepping give bit-identical populations (a sample
    of 200k leaves), FP64 totals and force;
  * two independent solvers on the same mesh agree bit for bit (no
    dependence on allocation addresses or launch timing);
  * the flow stays finite and below the lattice speed limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_sphere_full_size_properties(fl):
    from bench import sphere_params
    from fluidloom_amd import meshes
    from fluidloom_amd.solver import LbmSolver

    nx0, ny0, nz0, r, levels, u_in = 256, 128, 128, 16.0, 3, 0.05
    print("full-size sphere: building mesh", flush=True)
    mesh = meshes.sphere(nx0, ny0, nz0, (64.0, 64.0, 64.0), r, 2 * r, levels)
    tau0, bc, bcu = sphere_params(nx0, ny0, nz0, r, u_in, 100.0)
    lv, st = np.asarray(mesh["level"]), np.asarray(mesh["state"])
    expect = [int(((lv == l) & (st == 0)).sum()) for l in range(levels + 1)]

    def make():
        return LbmSolver(nx0, ny0, nz0, levels, tau0, bc, bcu, mesh["x"], mesh["y"], mesh["z"], mesh["level"],
                         mesh["state"], storage="f32", u_init=(u_in, 0.0, 0.0))

    a = make()
    s = a.stats()
    assert [int(s.leaves[l]) for l in range(levels + 1)] == expect
    assert int(s.updates_per_step) == sum(n << l for l, n in enumerate(expect))
    b = make()
    b.set_graph(True)
    print("full-size sphere: stepping eager and graph", flush=True)
    a.step(6)
    b.step(6)
    x, y, z, lvl = a.fluid[0]
    rng = np.random.default_rng(5)
    pick = rng.choice(len(x), 200_000, replace=False)
    q = (np.ascontiguousarray(np.asarray(x)[pick]), np.ascontiguousarray(np.asarray(y)[pick]),
         np.ascontiguousarray(np.asarray(z)[pick]), np.ascontiguousarray(np.asarray(lvl)[pick]))
    fa, fb = a.get_f(*q), b.get_f(*q)
    assert np.array_equal(fa, fb)
    ma, ja = a.totals()
    mb, jb = b.totals()
    assert ma == mb and np.array_equal(ja, jb)
    assert np.array_equal(a.force(), b.force())
    _, u = a.moments(*q)
    assert np.isfinite(u).all() and np.abs(u).max() < 0.3
    a.close()
    b.close()


def test_sphere_quarter_scale_f32_matches_oracle(fl, oracle):
    """The sphere configuration at a quarter of its size (the bench's
    --scale 4 case: base 64x32x32, r = 4 at (16,16,16), 3 levels within
    2r, Re 100; 1.3M leaves), FP32 storage, 100 coarse steps (800 level-3
    sub-steps) against the FP64 oracle on the same leaves: density and
    velocity within 1e-4 relative L2, the momentum-exchange force within
    1e-4 relative.  The leaf set comes from the same device mesh builder as
    the bench (its regrid passes are checked against the oracle's in the
    integer tests)."""
    import time

    from bench import sphere_params
    from fluidloom_amd import meshes
    from fluidloom_amd.solver import LbmSolver

    nx0, ny0, nz0, r, levels, u_in = 64, 32, 32, 4.0, 3, 0.05
    mesh = meshes.sphere(nx0, ny0, nz0, (16.0, 16.0, 16.0), r, 2 * r, levels)
    tau0, bc, bcu = sphere_params(nx0, ny0, nz0, r, u_in, 100.0)
    g = LbmSolver(nx0, ny0, nz0, levels, tau0, bc, bcu, mesh["x"], mesh["y"], mesh["z"], mesh["level"],
                  mesh["state"], storage="f32", u_init=(u_in, 0.0, 0.0))
    bci = [{"wall": 0, "outflow": 1}[b] for b in bc]
    o = oracle.Lbm(nx0, ny0, nz0, levels, tau0, bci, bcu, mesh["x"], mesh["y"], mesh["z"], mesh["level"],
                   mesh["state"], u_init=(u_in, 0.0, 0.0))
    print(f"quarter-scale sphere: {len(o.fx)} fluid leaves; oracle stepping", flush=True)
    t0 = time.perf_counter()
    for k in range(4):
        o.step(25)
        print(f"  oracle {25 * (k + 1)} steps, {time.perf_counter() - t0:.1f} s", flush=True)
    g.step(100)
    ro, uo = o.moments()
    rg, ug = g.moments(o.fx, o.fy, o.fz, o.flv)
    e_rho = np.linalg.norm(rg - ro) / np.linalg.norm(ro)
    e_u = np.linalg.norm(ug - uo) / np.linalg.norm(uo)
    Fo, Fg = o.force(), np.asarray(g.force())
    e_f = np.abs(Fg - Fo).max() / np.abs(Fo[0])
    print(f"quarter-scale sphere FP32 vs FP64 oracle, 100 steps: rho {e_rho:.2e} u {e_u:.2e} "
          f"force {e_f:.2e} (Fx {Fg[0]:.6e} vs {Fo[0]:.6e})", flush=True)
    g.close()
    assert e_rho < 1e-4 and e_u < 1e-4 and e_f < 1e-4


def test_sphere_full_size_f32_matches_fp64_run(fl):
    """The sphere configuration at its full size (base 256x128x128, r = 16,
    3 levels within 2r, Re 100; 76.4M leaves), FP32 storage against the
    FP64-storage run of the same kernels (the GPU oracle run, held to the
    CPU oracle within 1e-11 on the smaller cases), 200 coarse steps (1,600
    level-3 sub-steps) from the same start: density and velocity within 1e-4
    relative L2 over every leaf, and the drag coefficient within 1e-4
    relative.  The same comparison over the configuration's 5,000 steps is
    `python -m fluidloom_amd.sphere --compare f32`."""
    from fluidloom_amd import sphere

    print("full-size sphere: FP32 and FP64 runs, 200 steps", flush=True)
    out = sphere.precision(steps=200, report_every=100, forms=("f32",), tune=False)
    r = out["results"]["f32"]
    print(f"full-size sphere FP32 vs FP64, 200 steps: rho {r['rho_l2']:.2e} u {r['vel_l2']:.2e} "
          f"C_d {r['cd']:.6f} vs {out['results']['f64']['cd']:.6f} (rel {r['cd_rel_err']:.2e})", flush=True)
    assert out["leaves"] > 76_000_000
    assert r["rho_l2"] < 1e-4 and r["vel_l2"] < 1e-4 and r["cd_rel_err"] < 1e-4


def test_sphere_precision_driver_small(fl):
    """The storage-form comparison on a small sphere (base 32x16x16, r = 2,
    3 levels, 40 steps): every form reports; FP32 stays within the FP32
    tolerance of the FP64 run, and the 16-bit forms within 1e-2 (about 25
    times the 3.8e-4 / 1.8e-4 they show at twice this size after 200
    steps); each form's mass drift matches the FP64 run's to 1e-4."""
    from fluidloom_amd import sphere

    out = sphere.precision(steps=40, report_every=20, scale=8, forms=("f32", "f16", "f16c"), tune=False)
    res = out["results"]
    assert set(res) == {"f64", "f32", "f16", "f16c"}
    assert len(out["history"]) == 2 and all(np.isfinite(v) for row in out["history"] for v in row.values())
    assert res["f32"]["vel_l2"] < 1e-4 and res["f32"]["rho_l2"] < 1e-4 and res["f32"]["cd_rel_err"] < 1e-4
    for st in ("f16", "f16c"):
        assert res[st]["vel_l2"] < 1e-2 and res[st]["rho_l2"] < 1e-3, (st, res[st])
    for st in ("f32", "f16", "f16c"):
        assert abs(res[st]["mass_drift"] - res["f64"]["mass_drift"]) < 1e-4, (st, res[st])
