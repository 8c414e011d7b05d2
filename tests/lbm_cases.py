"""Lattice test cases shared by the oracle (CPU) and parity (GPU) tests."""
import numpy as np

from tests.amr_util import ball_region, box_region, mark_solid_ball, refine_mesh, uniform_cells

WALL, OUT, PER = 0, 1, 2
ZERO = [(0.0, 0.0, 0.0)] * 6


def case(name, O):
    """Returns dict(params..., cells=CellSet)."""
    if name == "periodic":
        c = uniform_cells(O, (8, 8, 8))
        return dict(n0=(8, 8, 8), max_level=0, tau0=0.7, bc=[PER] * 6, bc_u=ZERO, cells=c,
                    u_init=(0.02, -0.01, 0.005))
    if name == "cavity":
        c = uniform_cells(O, (8, 8, 8))
        bcu = list(ZERO)
        bcu[5] = (0.05, 0.0, 0.0)
        return dict(n0=(8, 8, 8), max_level=0, tau0=0.6, bc=[WALL] * 6, bc_u=bcu, cells=c, u_init=(0, 0, 0))
    if name == "cavity2":
        base = uniform_cells(O, (8, 8, 8))
        c = refine_mesh(O, base, (8, 8, 8), [(1, box_region((0, 0, 5), (8, 8, 8)))], 1)
        bcu = list(ZERO)
        bcu[5] = (0.05, 0.0, 0.0)
        return dict(n0=(8, 8, 8), max_level=1, tau0=0.6, bc=[WALL] * 6, bc_u=bcu, cells=c, u_init=(0, 0, 0))
    if name == "patch2":
        base = uniform_cells(O, (12, 8, 8))
        c = refine_mesh(O, base, (12, 8, 8), [(1, box_region((4, 2, 2), (8, 6, 6)))], 1)
        return dict(n0=(12, 8, 8), max_level=1, tau0=0.62, bc=[PER] * 6, bc_u=ZERO, cells=c,
                    u_init=(0.03, 0.01, -0.02))
    if name == "outper2":
        # outflow on +x with periodic y and z: edge links of the +x face that
        # leave through a periodic face take the periodic image, clamped only
        # along x; the refined box touches the +x face and the y seam
        base = uniform_cells(O, (16, 8, 8))
        c = refine_mesh(O, base, (16, 8, 8), [(1, box_region((10, 0, 2), (16, 4, 6)))], 1)
        bcu = list(ZERO)
        bcu[0] = (0.04, 0.0, 0.0)
        return dict(n0=(16, 8, 8), max_level=1, tau0=0.6, bc=[WALL, OUT, PER, PER, PER, PER], bc_u=bcu, cells=c,
                    u_init=(0.04, 0.0, 0.0))
    if name == "sphere3":
        base = uniform_cells(O, (16, 8, 8))
        ctr = (5.0, 4.0, 4.0)
        c = refine_mesh(O, base, (16, 8, 8), [(3, ball_region(ctr, 1.5))], 3)
        c = mark_solid_ball(c, ctr, 1.0)
        bcu = list(ZERO)
        bcu[0] = (0.05, 0.0, 0.0)
        return dict(n0=(16, 8, 8), max_level=3, tau0=0.56, bc=[WALL, OUT, WALL, WALL, WALL, WALL], bc_u=bcu,
                    cells=c, u_init=(0.05, 0.0, 0.0))
    raise KeyError(name)


def couette_wall(O, N, tau, refined=False, U=0.02):
    """Plane Couette flow on 4 x 4 x N base cells, periodic in x and y: the
    stationary wall is a layer of solid cells below z = 1 (level-0 units),
    the lid (+z face) moves at U along x.  With half-way bounce-back the walls
    sit at z = 1 and z = N, so H = N - 1, and the steady force on the solid
    layer is exactly F = (rho nu U / H * A, 0, -rho cs^2 A), A = 16,
    nu = (tau - 1/2) / 3, cs^2 = 1/3.  `refined` puts the bottom three base
    layers (the solid included) on level 1."""
    base = uniform_cells(O, (4, 4, N))
    if refined:
        c = refine_mesh(O, base, (4, 4, N), [(1, box_region((0, 0, 0), (4, 4, 3)))], 1)
    else:
        c = base
    zc = (c.z + 0.5) * 2.0 ** (-c.level.astype(np.float64))
    c.state[:] = np.where(zc < 1.0, 1, 0)
    bcu = list(ZERO)
    bcu[5] = (U, 0.0, 0.0)
    nu = (tau - 0.5) / 3
    exact = np.array([nu * U / (N - 1) * 16, 0.0, -16.0 / 3])
    return dict(n0=(4, 4, N), max_level=1 if refined else 0, tau0=tau, bc=[PER, PER, PER, PER, WALL, WALL],
                bc_u=bcu, cells=c, u_init=(0.0, 0.0, 0.0)), exact


def oracle_solver(O, cs):
    c = cs["cells"]
    nx, ny, nz = cs["n0"]
    return O.Lbm(nx, ny, nz, cs["max_level"], cs["tau0"], cs["bc"], cs["bc_u"], c.x, c.y, c.z, c.level, c.state,
                 u_init=cs["u_init"])


def gpu_solver(cs, storage):
    from fluidloom_amd.solver import LbmSolver
    c = cs["cells"]
    nx, ny, nz = cs["n0"]
    return LbmSolver(nx, ny, nz, cs["max_level"], cs["tau0"], cs["bc"], cs["bc_u"], c.x, c.y, c.z, c.level, c.state,
                     storage=storage, u_init=cs["u_init"])


def perturb(f, seed, amp=1e-3):
    rng = np.random.default_rng(seed)
    return f * (1.0 + amp * rng.standard_normal(f.shape))
