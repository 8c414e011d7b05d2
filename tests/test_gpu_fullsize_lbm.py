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
