"""Leaf sets of the benchmark configurations, built with the device regrid.

sphere(): flow past a sphere -- base grid refined to `levels` inside a ball
of `refine_radius` (level-0 units) around the sphere centre by repeated
mark -> fl_regrid passes (D3Q19 neighbour-cube balance), then a
balance-only pass until nothing splits; leaves whose centre lies inside the
sphere become SOLID.  cavity_base(): the uniform level-0 grid of the
lid-driven cavity.
"""
from __future__ import annotations

import numpy as np

from . import amr
from ._lib import FluidLoomError, FL_E_CAPACITY


def uniform(nx0, ny0, nz0):
    z, y, x = np.meshgrid(np.arange(nz0), np.arange(ny0), np.arange(nx0), indexing="ij")
    n = nx0 * ny0 * nz0
    return dict(x=x.ravel().astype(np.int32), y=y.ravel().astype(np.int32), z=z.ravel().astype(np.int32),
                level=np.zeros(n, np.uint8), state=np.zeros(n, np.uint8))


def _regrid_retry(cells, root, max_level, marked):
    n = cells.n
    cap = n + 8 * marked + n // 2 + 64
    while True:
        try:
            return amr.regrid(cells, balance_mode=1, max_split_level=max_level, root=root,
                              out_capacity=cap)
        except FluidLoomError as e:
            if e.status != FL_E_CAPACITY or cap >= 8 * n:
                raise
            cap = min(8 * n, 2 * cap)


def refine_ball(base: dict, root, center, refine_radius, levels, max_passes=32, log=None):
    """Device regrid of `base` (host leaf dict) until every cell within
    refine_radius of center is at `levels` and the mesh is balanced."""
    cells = amr.GpuCells.from_host(**base)
    for _ in range(max_passes):
        marked = amr.mark_ball(cells, center, refine_radius, levels)
        new, st = _regrid_retry(cells, root, levels, marked)
        if log:
            log(f"regrid: {cells.n} -> {new.n} cells (marked {marked}, split {st.num_split}, "
                f"balance iters {st.balance_iterations})")
        cells.free()
        cells = new
        if marked == 0 and st.num_split == 0:
            break
    return cells


def sphere(nx0=256, ny0=128, nz0=128, center=(64.0, 64.0, 64.0), radius=16.0, refine_radius=32.0,
           levels=3, log=None):
    base = uniform(nx0, ny0, nz0)
    c = np.array(center, np.float64)
    cells = refine_ball(base, (nx0, ny0, nz0), c, refine_radius, levels, log=log)
    nsolid = amr.mark_solid_ball(cells, c, radius)
    h = cells.to_host()
    cells.free()
    if log:
        log(f"sphere mesh: {len(h['x'])} leaves, {nsolid} solid")
    return h


def cavity_base(n0=32):
    return uniform(n0, n0, n0)


def cavity(n0, levels, log=None):
    """Lid-driven cavity of n0^3 base cells with `levels` refinement levels
    in a ball of radius 0.3 n0 centred on the middle of the lid (z = n0),
    where the shear layer sits."""
    base = uniform(n0, n0, n0)
    if levels == 0:
        return base
    cells = refine_ball(base, (n0, n0, n0), np.array([n0 / 2, n0 / 2, float(n0)]), 0.3 * n0, levels, log=log)
    h = cells.to_host()
    cells.free()
    return h
