"""Batched solver on the GPU: every instance of a batch must give exactly
the result it gives alone (same kernels, same per-cell arithmetic, per-
instance reductions in a fixed order), graph replay must equal eager
stepping, all collide variants must agree bit for bit, and non-live
instances must stay frozen."""
import numpy as np
import pytest

from tests.lbm_cases import case, gpu_solver, oracle_solver, perturb

pytestmark = pytest.mark.gpu
CASES = ["periodic", "cavity", "cavity2", "patch2", "sphere3"]


def params_of(cs, storage):
    from fluidloom_amd.solver import make_params
    nx, ny, nz = cs["n0"]
    return make_params(nx, ny, nz, cs["max_level"], cs["tau0"], cs["bc"], cs["bc_u"], storage,
                       u_init=cs["u_init"])


def mesh_of(cs):
    c = cs["cells"]
    return dict(x=c.x, y=c.y, z=c.z, level=c.level, state=c.state)


def batch_of(oracle, names, storage, graph=False):
    from fluidloom_amd.solver import LbmBatch
    css = [case(n, oracle) for n in names]
    return css, LbmBatch([params_of(cs, storage) for cs in css], [mesh_of(cs) for cs in css], graph=graph)


@pytest.mark.parametrize("storage", ["f32", "f16", "f16c", "f64"])
def test_batch_equals_single(fl, oracle, storage):
    names = CASES + ["cavity2", "sphere3"]
    css, b = batch_of(oracle, names, storage)
    singles = [gpu_solver(cs, storage) for cs in css]
    for k, cs in enumerate(css):
        f0 = perturb(singles[k].get_f(), 11 + k, 1e-3)
        singles[k].set_f(f0)
        b.set_f(f0, inst=k)
    b.step(5)
    for s in singles:
        s.step(5)
    m, j = b.totals_batch()
    F = b.force_batch()
    for k, s in enumerate(singles):
        assert np.array_equal(b.get_f(k), s.get_f()), names[k]
        ms, js = s.totals()
        assert m[k] == ms and np.array_equal(j[k], js), names[k]
        assert np.array_equal(F[k], s.force()), names[k]
    mt, jt = b.totals_batch()
    assert np.array_equal(mt, m)


def test_batch_instance_matches_oracle(fl, oracle):
    names = ["sphere3", "cavity2", "patch2"]
    css, b = batch_of(oracle, names, "f64")
    b.step(4)
    for k, cs in enumerate(css):
        o = oracle_solver(oracle, cs)
        o.step(4)
        fo = o.get_f()
        assert np.abs(b.get_f(k) - fo).max() / np.abs(fo).max() < 1e-11


def test_graph_replay_equals_eager(fl, oracle):
    names = ["sphere3", "cavity2", "periodic"]
    _, g = batch_of(oracle, names, "f32", graph=True)
    _, e = batch_of(oracle, names, "f32", graph=False)
    for chunk in (3, 4, 1, 2, 6):
        g.step(chunk)
        e.step(chunk)
        for k in range(len(names)):
            assert np.array_equal(g.get_f(k), e.get_f(k)), (chunk, names[k])
        assert np.array_equal(g.totals_batch()[0], e.totals_batch()[0])
    assert g.stats().graph == 1 and g.stats().graph_launches >= 6


@pytest.mark.parametrize("storage", ["f32", "f16", "f16c", "f64"])
def test_staged_f2c_equals_gathered(fl, oracle, storage):
    # From the second step of a stepping pass on, the f2c fills read the child
    # sums that the finer level's collide staged; a pass's first step gathers
    # them from the fine populations.  Six steps in one pass against six
    # single-step passes (which never stage at the coarsest level): the same
    # bits.
    names = ["sphere3", "cavity2", "patch2"]
    _, a = batch_of(oracle, names, storage)
    _, b = batch_of(oracle, names, storage)
    for k in range(len(names)):
        f0 = perturb(a.get_f(k), 41 + k, 1e-3)
        a.set_f(f0, inst=k)
        b.set_f(f0, inst=k)
    a.step(6)
    for _ in range(6):
        b.step(1)
    for k in range(len(names)):
        assert np.array_equal(a.get_f(k), b.get_f(k)), names[k]
    assert np.array_equal(a.totals_batch()[0], b.totals_batch()[0])


def test_collide_variants_agree(fl, oracle):
    from fluidloom_amd.solver import variant_names
    names = ["sphere3", "patch2"]
    ref = None
    nv = len(variant_names())
    assert nv >= 6
    for v in range(nv):
        css, b = batch_of(oracle, names, "f32")
        for l in range(b.stats().nlevels):
            b.set_variant(l, v)
        b.step(3)
        got = [b.get_f(k) for k in range(len(names))]
        if ref is None:
            ref = got
        for k in range(len(names)):
            assert np.array_equal(got[k], ref[k]), (variant_names()[v], names[k])
        b.close()


def test_autotune_keeps_results(fl, oracle):
    names = ["sphere3", "cavity"]
    _, a = batch_of(oracle, names, "f32")
    _, b = batch_of(oracle, names, "f32")
    a.step(1)
    b.step(1)
    f0 = [a.get_f(k) for k in range(2)]
    chosen = a.autotune(4)
    assert len(chosen) == a.stats().nlevels
    # timing leaves both buffers of every level as they were
    for k in range(2):
        assert np.array_equal(a.get_f(k), f0[k])
    a.step(4)
    b.step(4)
    for k in range(2):
        assert np.array_equal(a.get_f(k), b.get_f(k))


def test_live_flag_freezes_instance(fl, oracle):
    names = ["cavity", "cavity2", "sphere3"]
    css, b = batch_of(oracle, names, "f32")
    b.step(2)
    frozen = b.get_f(1)
    b.set_live(1, False)
    b.step(3)
    assert np.array_equal(b.get_f(1), frozen)
    s = gpu_solver(css[2], "f32")
    s.step(5)
    assert np.array_equal(b.get_f(2), s.get_f())


def test_velocity_change(fl, oracle):
    names = ["cavity", "sphere3"]
    _, b = batch_of(oracle, names, "f32")
    b.step(3)
    du0 = b.velocity_change()
    assert np.all(np.isinf(du0))
    u_prev = [b.moments(k)[1] for k in range(2)]
    b.step(2)
    du = b.velocity_change()
    for k in range(2):
        u_now = b.moments(k)[1]
        host = np.abs(u_now.astype(np.float32).astype(np.float64) - u_prev[k].astype(np.float32).astype(np.float64))
        assert abs(du[k] - host.max()) <= 1e-6 * max(host.max(), 1e-12) + 1e-9, k
    assert np.all(du > 0)


def test_batched_regrid_equals_single_regrids(fl, oracle):
    """Two cavities of different size, lid and tau stepped as one batch with a
    vorticity-driven regrid of both (fl_lbm_regrid_batch): every instance
    ends bit-identical to the same cavity run alone with fl_lbm_regrid."""
    from fluidloom_amd import batched
    from fluidloom_amd.solver import LbmBatch
    specs = [dict(n0=16, lid=0.08, tau=0.6, levels=1), dict(n0=12, lid=0.05, tau=0.7, levels=1)]
    meshes = batched.build_meshes([dict(s, levels=0) for s in specs])
    for storage in ("f32", "f16"):
        b = LbmBatch([batched.cavity_params(s, storage) for s in specs], meshes)
        singles = [LbmBatch([batched.cavity_params(s, storage)], [m]) for s, m in zip(specs, meshes)]
        for rnd in range(2):
            b.step(30)
            for s in singles:
                s.step(30)
            # flags from the batch, applied to both forms
            new = batched.regrid_meshes(b, specs)
            b.regrid_batch(new)
            for k, s in enumerate(singles):
                s.regrid_batch([new[k]])
        b.step(10)
        for s in singles:
            s.step(10)
        for k, s in enumerate(singles):
            assert np.array_equal(b.get_f(k), s.get_f(0)), (storage, k)
            assert (np.asarray(b.meshes[k]["level"]) == 1).any()
