"""GPU: the lattice layout built by kernels from device leaf arrays equals
the host builder's layout table for table, bit for bit (brick order,
neighbour rows, link words, ghost lists, c2f groups, capture lists, f2c
children, force links, per-instance ranges), on the test cases, a batch of
several instances and the full-size sphere; and a regrid through the device
path gives the same populations as one through the host path."""
import ctypes as C
import time

import numpy as np
import pytest

from tests.lbm_cases import case

pytestmark = pytest.mark.gpu

_TABLES = (("nbr", np.int32, lambda L, ni: L.nb * 27), ("bcoord", np.int32, lambda L, ni: L.nb * 3),
           ("binst", np.uint16, lambda L, ni: L.nb), ("bflags", np.uint32, lambda L, ni: L.nb),
           ("cflags", np.uint64, lambda L, ni: L.nb * 64), ("c2f_gchild", np.uint32, lambda L, ni: L.n_groups * 8),
           ("c2f_gcap", np.uint32, lambda L, ni: L.n_groups * 7), ("c2f_gmask", np.uint32, lambda L, ni: L.n_groups * 8),
           ("cap_slots", np.uint32, lambda L, ni: L.n_cap), ("f2c_ghost", np.uint32, lambda L, ni: L.n_f2c),
           ("f2c_child", np.uint32, lambda L, ni: L.n_f2c * 8), ("force_links", np.uint32, lambda L, ni: L.n_links),
           ("abeg", np.uint32, lambda L, ni: ni + 1), ("lbeg", np.uint32, lambda L, ni: ni + 1))


def layout(fl, batch, level):
    L = fl.LbmLayout()
    fl.check(fl.lib().fl_lbm_layout_get(batch.h, level, C.byref(L)))
    out = {k: int(getattr(L, k)) for k in ("nb", "nactive", "n_groups", "n_cap", "n_f2c", "n_links")}
    for name, dt, size in _TABLES:
        n = int(size(L, batch.ninst))
        a = np.empty(n, dt)
        if n:
            fl.check(fl.lib().fl_memcpy_d2h(a.ctypes.data, getattr(L, name), a.nbytes, None))
        out[name] = a
    return out


def build(fl, params, meshes, host):
    from fluidloom_amd.solver import LbmBatch
    fl.check(fl.lib().fl_lbm_layout_on_host(1 if host else 0))
    try:
        t0 = time.perf_counter()
        b = LbmBatch(params, meshes)
        b.sync()
        return b, time.perf_counter() - t0
    finally:
        fl.check(fl.lib().fl_lbm_layout_on_host(0))


def same_layouts(fl, a, b):
    assert a.stats().nlevels == b.stats().nlevels
    for l in range(a.stats().nlevels):
        la, lb = layout(fl, a, l), layout(fl, b, l)
        for k in la:
            assert np.array_equal(la[k], lb[k]), (l, k)


def _params(cs, storage="f32"):
    from fluidloom_amd.solver import make_params
    return make_params(*cs["n0"], cs["max_level"], cs["tau0"], cs["bc"], cs["bc_u"], storage, u_init=cs["u_init"])


def _mesh(cs):
    c = cs["cells"]
    return dict(x=c.x, y=c.y, z=c.z, level=c.level, state=c.state)


@pytest.mark.parametrize("name", ["periodic", "cavity", "cavity2", "patch2", "sphere3", "outper2"])
def test_device_layout_equals_host_layout(fl, oracle, name):
    cs = case(name, oracle)
    h, _ = build(fl, [_params(cs)], [_mesh(cs)], host=True)
    d, _ = build(fl, [_params(cs)], [_mesh(cs)], host=False)
    same_layouts(fl, h, d)
    # the same populations after a few steps (the tables drive everything)
    h.step(4)
    d.step(4)
    assert np.array_equal(h.get_f(0), d.get_f(0))
    h.close()
    d.close()


def test_device_layout_equals_host_layout_batch(fl, oracle):
    names = ["sphere3", "cavity2", "patch2", "outper2", "cavity"]
    cs = [case(n, oracle) for n in names]
    h, _ = build(fl, [_params(c) for c in cs], [_mesh(c) for c in cs], host=True)
    d, _ = build(fl, [_params(c) for c in cs], [_mesh(c) for c in cs], host=False)
    same_layouts(fl, h, d)
    h.close()
    d.close()


def test_device_layout_full_sphere(fl):
    """The sphere workload at full size (base 256x128x128, r = 16, three
    levels within 2r: 67.6M leaves, 87M slots): the device builder's tables
    equal the host builder's, and it takes a small fraction of its time."""
    from bench import sphere_params
    from fluidloom_amd import meshes
    from fluidloom_amd.solver import make_params
    nx0, ny0, nz0, r, levels = 256, 128, 128, 16.0, 3
    mesh = meshes.sphere(nx0, ny0, nz0, (64.0, 64.0, 64.0), r, 2 * r, levels)
    tau0, bc, bcu = sphere_params(nx0, ny0, nz0, r, 0.05, 100.0)
    p = make_params(nx0, ny0, nz0, levels, tau0, bc, bcu, "f32", u_init=(0.05, 0.0, 0.0))
    d, td = build(fl, [p], [mesh], host=False)
    d2, td2 = build(fl, [p], [mesh], host=False)
    h, th = build(fl, [p], [mesh], host=True)
    print(f"full sphere layout: device builder {td2:.3f} s (first {td:.3f} s), host builder {th:.3f} s "
          f"(both including buffers and initial state)", flush=True)
    same_layouts(fl, h, d)
    for b in (h, d, d2):
        b.close()


def test_regrid_device_path_equals_host_path(fl, oracle):
    """A refined cavity regridded (split + merge) through the device path and
    through the host path: same layouts, same populations."""
    from fluidloom_amd import amr
    from fluidloom_amd.cavity import regrid_flags
    cs = case("cavity2", oracle)
    out = []
    for host in (True, False):
        b, _ = build(fl, [_params(cs)], [_mesh(cs)], host=host)
        b.step(30)
        m = b.meshes[0]
        w = b.vorticity(0)
        flags = regrid_flags(np.asarray(m["level"]), np.asarray(m["state"]), w, 0.5, 1)
        cells = amr.GpuCells.from_host(x=m["x"], y=m["y"], z=m["z"], level=m["level"], state=m["state"], flag=flags)
        new, st = amr.regrid(cells, balance_mode=1, max_split_level=1, root=(8, 8, 8), merge_balance_check=1)
        assert st.num_split > 0 or st.num_groups > 0
        if host:
            fl.check(fl.lib().fl_lbm_layout_on_host(1))
            try:
                hm = new.to_host()
                b.regrid_batch([hm])
            finally:
                fl.check(fl.lib().fl_lbm_layout_on_host(0))
        else:
            b.regrid_batch_device(new, [0, new.n])
            # host copies of the leaves, made on demand, before `new` is freed
            assert np.array_equal(b.meshes[0]["x"], new.to_host()["x"])
        b.step(10)
        out.append(b)
        cells.free()
        new.free()
    h, d = out
    same_layouts(fl, h, d)
    assert np.array_equal(h.get_f(0), d.get_f(0))
    h.close()
    d.close()


def test_device_layout_mesh_errors(fl, oracle):
    """The device builder reports the host builder's mesh errors: a duplicate
    leaf, a leaf outside the domain, a leaf above the instance's max_level
    (the unbalanced-mesh error is covered by test_mesh_errors_are_reported)."""
    from fluidloom_amd._lib import FL_E_MESH, FluidLoomError
    from fluidloom_amd.solver import LbmBatch
    cs = case("cavity2", oracle)
    m = _mesh(cs)
    bad = []
    dup = {k: np.concatenate([np.asarray(v), np.asarray(v)[:1]]) for k, v in m.items()}
    bad.append(("duplicate leaf", dup))
    out = {k: np.asarray(v).copy() for k, v in m.items()}
    out["x"][0] = 64
    bad.append(("outside the domain", out))
    hi = {k: np.asarray(v).copy() for k, v in m.items()}
    hi["level"][0] = 3
    bad.append(("max_level", hi))
    for what, mesh in bad:
        for host in (True, False):
            fl.check(fl.lib().fl_lbm_layout_on_host(1 if host else 0))
            try:
                with pytest.raises(FluidLoomError) as e:
                    LbmBatch([_params(cs)], [mesh])
            finally:
                fl.check(fl.lib().fl_lbm_layout_on_host(0))
            assert e.value.status == FL_E_MESH and what in str(e.value), (what, host, str(e.value))


def test_block_cache_reuse_and_release(fl, oracle):
    """Freed solver buffers stay in the library's block cache and are reused
    by the next solver of the same size (no new device memory); releasing
    the cache returns them to HIP."""
    cs = case("cavity2", oracle)
    lib = fl.lib()
    held = C.c_uint64(0)
    b, _ = build(fl, [_params(cs)], [_mesh(cs)], host=False)
    b.close()
    fl.check(lib.fl_cached_memory_bytes(C.byref(held)))
    assert held.value > 0
    before = held.value
    b, _ = build(fl, [_params(cs)], [_mesh(cs)], host=False)
    fl.check(lib.fl_cached_memory_bytes(C.byref(held)))
    assert held.value < before  # the new solver took its blocks from the cache
    b.step(5)
    b.close()
    fl.check(lib.fl_release_cached_memory())
    fl.check(lib.fl_cached_memory_bytes(C.byref(held)))
    assert held.value == 0
    # a solver built after the release allocates afresh and steps
    b, _ = build(fl, [_params(cs)], [_mesh(cs)], host=False)
    b.step(5)
    assert np.isfinite(b.get_f(0)).all()
    b.close()
