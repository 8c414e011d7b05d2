"""CPU: the FP64 lattice oracle pinned by closed forms (the reference has no
LBM arithmetic to compare with; see the oracle module docstring)."""
import numpy as np
import pytest

from tests.amr_util import box_region, refine_mesh, uniform_cells
from tests.lbm_cases import OUT, PER, WALL, ZERO, case, couette_wall, oracle_solver, perturb


def test_lattice_constants(oracle):
    c, w, opp = oracle.lattice()
    # the reference fixture's first 7 vectors and weights (test_parsing.cpp:23-30)
    assert c[:7].tolist() == [[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]]
    assert np.allclose(w[:7], [1 / 3] + [1 / 18] * 6)
    assert abs(w.sum() - 1.0) < 1e-12  # LatticeDescriptor::validate (LatticeDescriptor.cpp:9-28)
    assert np.allclose((w[:, None] * c).sum(0), 0)
    assert np.allclose(np.einsum("q,qa,qb->ab", w, c, c), np.eye(3) / 3)
    assert np.array_equal(c[opp], -c)
    assert len({tuple(v) for v in c}) == 19


def test_equilibrium_moments_and_fixed_point(oracle):
    cs = case("periodic", oracle)
    s = oracle_solver(oracle, cs)
    c, w, _ = oracle.lattice()
    f = s.get_f()
    u = np.array(cs["u_init"])
    assert np.allclose(f.sum(1), 1.0, atol=1e-14)
    assert np.allclose(f @ c, u, atol=1e-14)
    Pi = np.einsum("nq,qa,qb->nab", f, c, c)
    assert np.allclose(Pi, np.eye(3) / 3 + np.outer(u, u), atol=1e-14)
    s.step(5)
    assert np.allclose(s.get_f(), f, atol=1e-14)


def test_rest_state_stays_at_rest(oracle):
    cs = case("cavity", oracle)
    cs["bc_u"] = ZERO
    s = oracle_solver(oracle, cs)
    s.step(20)
    _, w, _ = oracle.lattice()
    assert np.abs(s.get_f() - w).max() < 1e-15


def test_periodic_conservation(oracle):
    cs = case("periodic", oracle)
    s = oracle_solver(oracle, cs)
    s.set_f(perturb(s.get_f(), 1, 1e-2))
    m0, j0 = s.totals()
    s.step(10)
    m1, j1 = s.totals()
    assert abs(m1 - m0) < 1e-12 and np.abs(j1 - j0).max() < 1e-13


def test_taylor_green_decay(oracle):
    N, tau, U = 32, 0.8, 0.01
    c = uniform_cells(oracle, (N, N, 4))
    s = oracle.Lbm(N, N, 4, 0, tau, [PER] * 6, ZERO, c.x, c.y, c.z, c.level, c.state)
    k = 2 * np.pi / N
    x, y = c.x + 0.5, c.y + 0.5
    ux, uy = U * np.sin(k * x) * np.cos(k * y), -U * np.cos(k * x) * np.sin(k * y)
    cv, w, _ = oracle.lattice()
    uu = np.stack([ux, uy, np.zeros_like(ux)], 1)
    cu = uu @ cv.T
    feq = w * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (uu ** 2).sum(1)[:, None])
    s.set_f(feq)
    T = 400
    s.step(T)
    _, u = s.moments()
    nu = (tau - 0.5) / 3
    amp = np.sqrt((u[:, 0] ** 2 + u[:, 1] ** 2).mean() / (ux ** 2 + uy ** 2).mean())
    assert abs(amp / np.exp(-2 * nu * k * k * T) - 1) < 0.01


def _couette(oracle, N, refined, steps):
    U = 0.02
    bcu = list(ZERO)
    bcu[5] = (U, 0.0, 0.0)
    base = uniform_cells(oracle, (4, 4, N))
    if refined:
        band = (0, 0, N // 2 - 1), (4, 4, N // 2 + 1)  # two coarse cells thick
        c = refine_mesh(oracle, base, (4, 4, N), [(1, box_region(*band))], 1)
        ml = 1
    else:
        c, ml = base, 0
    s = oracle.Lbm(4, 4, N, ml, 1.0, [PER, PER, PER, PER, WALL, WALL], bcu, c.x, c.y, c.z, c.level, c.state)
    s.step(steps)
    _, u = s.moments()
    zc = (s.fz + 0.5) / (2.0 ** s.flv)
    exact = U * zc / N
    assert np.abs(u[:, 1:]).max() < 1e-12
    return np.abs(u[:, 0] - exact).max() / U


def test_couette_profile_exact_on_uniform_grid(oracle):
    # half-way bounce-back with the moving-wall term reproduces the linear
    # profile to round-off
    assert _couette(oracle, 8, False, 1500) < 1e-9


def test_couette_profile_exact_across_interfaces(oracle):
    # the coarse->fine fill interpolates density and velocity linearly from
    # the coarse neighbours and rescales the (uniform) shear stress, so the
    # linear Couette profile crosses a refined band unchanged
    assert _couette(oracle, 8, True, 1500) < 1e-11
    assert _couette(oracle, 16, True, 12000) < 1e-11


def test_couette_detects_a_zeroth_order_fill(oracle):
    # the same band with the fill's slopes switched off (planted in the
    # oracle) gives the O(dx) interface error of a parent copy
    U, N = 0.02, 8
    bcu = list(ZERO)
    bcu[5] = (U, 0.0, 0.0)
    base = uniform_cells(oracle, (4, 4, N))
    c = refine_mesh(oracle, base, (4, 4, N), [(1, box_region((0, 0, N // 2 - 1), (4, 4, N // 2 + 1)))], 1)
    s = oracle.Lbm(4, 4, N, 1, 1.0, [PER, PER, PER, PER, WALL, WALL], bcu, c.x, c.y, c.z, c.level, c.state)
    s.plant_coupling_error(1.0, 1.0, slopes=0.0)
    s.step(1500)
    _, u = s.moments()
    err = np.abs(u[:, 0] - U * ((s.fz + 0.5) / (2.0 ** s.flv)) / N).max() / U
    assert err > 1e-3


def test_uniform_flow_through_coarse_fine_interfaces(oracle):
    cs = case("patch2", oracle)
    s = oracle_solver(oracle, cs)
    s.step(6)
    rho, u = s.moments()
    assert np.abs(rho - 1).max() < 1e-13
    assert np.abs(u - np.array(cs["u_init"])).max() < 1e-13


def _feq(oracle, rho, u):
    cv, w, _ = oracle.lattice()
    cu = u @ cv.T
    return rho[:, None] * w * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (u ** 2).sum(1)[:, None])


def test_outflow_with_periodic_axes_is_shift_invariant(oracle):
    """Outflow (pressure outlet) on both x faces, periodic y and z, a refined
    box against the y seam and a flow that varies in y and z: shifting the
    mesh and the flow by 3 coarse cells along y, across the seam, must shift
    the result and change nothing else.  The edge links that leave through an
    outflow face and a periodic face at once must take the periodic image
    along y/z; clamping y to the seam instead makes the seam special and
    breaks the invariance there."""
    n0, shift = (8, 8, 8), 3
    base = uniform_cells(oracle, n0)
    c = refine_mesh(oracle, base, n0, [(1, box_region((0, 0, 2), (8, 4, 6)))], 1)
    ny = lambda lv: n0[1] << lv.astype(np.int64)  # noqa: E731
    runs = []
    for sh in (0, shift):
        y = ((c.y + (sh << c.level.astype(np.int64))) % ny(c.level)).astype(np.int32)
        s = oracle.Lbm(*n0, 1, 0.6, [OUT, OUT, PER, PER, PER, PER], ZERO, c.x, y, c.z, c.level, c.state)
        h = 2.0 ** s.flv
        yy, zz = (s.fy + 0.5) / h - sh, (s.fz + 0.5) / h
        k = 2 * np.pi / 8
        u = np.stack([0.03 * np.sin(k * yy + 0.4) * np.cos(k * zz), 0.01 * np.cos(k * zz + 0.2),
                      0.02 * np.sin(k * yy - 0.3)], 1)
        s.set_f(_feq(oracle, 1 + 0.01 * np.cos(k * yy) * np.sin(k * zz), u))
        s.step(6)
        # unshifted coordinates of each leaf, then a common order
        y0 = (s.fy - (sh << s.flv.astype(np.int64))) % ny(s.flv)
        order = np.lexsort((s.fz, y0, s.fx, s.flv))
        runs.append(s.get_f()[order])
    assert np.abs(runs[0] - runs[1]).max() < 1e-14
    # and the state is not trivially uniform: the y/z structure survives
    assert np.ptp(runs[0][:, 1]) > 1e-4


def test_refined_mass_drift_small(oracle):
    cs = case("patch2", oracle)
    s = oracle_solver(oracle, cs)
    s.set_f(perturb(s.get_f(), 3, 1e-3))
    m0, _ = s.totals()
    s.step(10)
    m1, _ = s.totals()
    assert abs(m1 - m0) / m0 < 1e-4


@pytest.mark.parametrize("refined", [False, True])
def test_pressure_outlet_uniform_channel(oracle, refined):
    """Velocity inlet on -x, pressure outlet on +x, periodic y and z: the
    uniform flow at the outlet density rho0 = 1 and the inlet velocity is the
    exact steady state.  It is a fixed point to round-off, and a state started
    at 1.03 times it returns to it (a plain copy at the outlet leaves the
    density level free and would keep the excess; a wrong sign of the density
    correction runs away).  The refined variant puts a level-1 box against the
    outlet and the y seam."""
    u = 0.04
    base = uniform_cells(oracle, (16, 4, 4))
    c = refine_mesh(oracle, base, (16, 4, 4), [(1, box_region((10, 0, 1), (16, 2, 3)))], 1) if refined else base
    bcu = list(ZERO)
    bcu[0] = (u, 0.0, 0.0)
    args = (16, 4, 4, 1 if refined else 0, 0.6, [WALL, OUT, PER, PER, PER, PER], bcu, c.x, c.y, c.z, c.level,
            c.state)
    s = oracle.Lbm(*args, u_init=(u, 0.0, 0.0))
    f0 = s.get_f()
    s.step(200)
    assert np.abs(s.get_f() - f0).max() < 1e-13
    s = oracle.Lbm(*args, u_init=(u, 0.0, 0.0))
    s.set_f(1.03 * f0)
    s.step(4000)
    rho, v = s.moments()
    assert np.abs(rho - 1).max() < 1e-6 and np.abs(v[:, 0] - u).max() < 1e-6
    assert np.abs(s.get_f() - f0).max() < 1e-6


@pytest.mark.parametrize("N,tau,steps,refined", [(8, 1.0, 2000, False), (8, 0.8, 4000, False),
                                                  (12, 0.9, 6000, False), (8, 1.0, 3000, True)])
def test_momentum_exchange_force_couette_closed_form(oracle, N, tau, steps, refined):
    """The momentum-exchange force on a stationary wall of solid cells in
    steady plane Couette flow: shear F_x = rho nu U / H A (which pins the
    viscosity, the factor 2 of each bounced link and the sign) and wall
    pressure F_z = -rho cs^2 A (which pins the per-level area scaling when
    the wall layer is refined).  Both hold exactly for half-way
    bounce-back, so the tolerance is round-off."""
    cs, exact = couette_wall(oracle, N, tau, refined)
    s = oracle_solver(oracle, cs)
    s.step(steps)
    F = s.force()
    assert abs(F[0] / exact[0] - 1) < 1e-9
    assert abs(F[1]) < 1e-12 and abs(F[2] / exact[2] - 1) < 1e-12


def test_sphere_case_runs_and_drags(oracle):
    cs = case("sphere3", oracle)
    s = oracle_solver(oracle, cs)
    s.step(3)
    F = s.force()
    rho, u = s.moments()
    assert np.isfinite(F).all() and F[0] > 0
    assert np.isfinite(u).all() and np.abs(rho - 1).max() < 0.1


def test_cavity_regrid_refines_lid_corners_and_conserves(oracle):
    from tests.cavity_util import OracleCavity
    cav = OracleCavity(oracle, n0=16, max_level=1)
    cav.run(121, 40)
    s = cav.s
    assert (s.level == 1).any() and (s.level == 0).any()
    # refinement sits where the shear is: next to the moving lid
    zc = (s.z + 0.5) / 2.0 ** s.level
    assert zc[s.level == 1].mean() > zc[s.level == 0].mean()
    rho, u = s.moments()
    assert np.isfinite(u).all() and np.abs(rho - 1).max() < 0.05
    # a regrid with unchanged leaves is the identity on the populations
    f0 = s.get_f()
    s.regrid(s.x, s.y, s.z, s.level, s.state)
    assert np.array_equal(s.get_f(), f0)


def test_regrid_split_and_merge_rules(oracle):
    """Splitting a uniform-flow region and merging it back keeps the flow
    uniform (equilibrium states carry no f_neq)."""
    cs_u = (0.03, 0.0, 0.01)
    c = uniform_cells(oracle, (8, 8, 8))
    s = oracle.Lbm(8, 8, 8, 1, 0.7, [PER] * 6, ZERO, c.x, c.y, c.z, c.level, c.state, u_init=cs_u)
    ref = refine_mesh(oracle, c, (8, 8, 8), [(1, box_region((2, 2, 2), (6, 6, 6)))], 1)
    s.regrid(ref.x, ref.y, ref.z, ref.level, ref.state)
    s.step(2)
    rho, u = s.moments()
    assert np.abs(u - np.array(cs_u)).max() < 1e-13
    s.regrid(c.x, c.y, c.z, c.level, c.state)
    rho, u = s.moments()
    assert np.abs(u - np.array(cs_u)).max() < 1e-13 and np.abs(rho - 1).max() < 1e-13
