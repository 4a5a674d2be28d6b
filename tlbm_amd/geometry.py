"""Node classification (ndType) in integer lattice units.

The reference's pre-processing (python/FluidChannel.py) builds geometry in
physical float32 coordinates and classifies nodes by these rules, which are
kept here on the integer lattice (x, y, z are node indices):

* walls (FluidChannel.py:1517-1538): 'right' x == 0, 'left' x == nx-1,
  'bottom' y == 0, 'top' y == ny-1 -> solid;
* obstacles: sphere and crossflow cylinder strictly inside (d^2 < r^2,
  FluidChannel.py:266-276, 411-416), prolate spheroid on or inside
  (g <= 0, FluidChannel.py:378-393);
* inlet = plane z == 0 minus walls and obstacles -> 2, outlet = plane
  z == nz-1 minus walls and obstacles -> 3 (FluidChannel.py:1358-1397);
* solid = obstacles union walls -> 1, fluid -> 0;
* lid-driven cavity (FluidChannel.py:1179-1243): walls x == 0, x == nx-1,
  y == 0, z == 0, z == nz-1 solid, then the lid y == ny-1 -> 5.

Arrays are returned flat in the reference's global order
g = x + nx*y + nx*ny*z (TLBM_Partition.cpp:730-737) as uint8.
"""
from __future__ import annotations

import numpy as np

FLUID, SOLID, INLET, OUTLET, LID = 0, 1, 2, 3, 5
VALID_TYPES = (FLUID, SOLID, INLET, OUTLET, LID)


def grid(nx: int, ny: int, nz: int):
    """Index grids shaped (nz, ny, nx), so ravel() is global order."""
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return x, y, z


def sphere(xc, yc, zc, r):
    def mask(x, y, z):
        return (x - xc) ** 2 + (y - yc) ** 2 + (z - zc) ** 2 < r * r
    return mask


def cylinder_x(yc, zc, r):
    """Crossflow cylinder with its axis along x."""
    def mask(x, y, z):
        return (y - yc) ** 2 + (z - zc) ** 2 < r * r
    return mask


def prolate_spheroid(xc, yc, zc, ab, c):
    """Semi-axis ab across the flow (x, y) and c along z, zero incidence."""
    def mask(x, y, z):
        dx = (x - xc).astype(np.float64)
        dy = (y - yc).astype(np.float64)
        dz = (z - zc).astype(np.float64)
        g = (dx * dx + dy * dy) / float(ab * ab) + (dz * dz) / float(c * c) - 1.0
        return g <= 0.0
    return mask


def channel(nx: int, ny: int, nz: int, walls=("left", "right", "top", "bottom"),
            obstacles=(), inlet: bool = True, outlet: bool = True) -> np.ndarray:
    """Channel along z with optional side walls, obstacles, inlet and outlet."""
    x, y, z = grid(nx, ny, nz)
    wall = np.zeros(x.shape, dtype=bool)
    for w in walls:
        if w == "right":
            wall |= x == 0
        elif w == "left":
            wall |= x == nx - 1
        elif w == "bottom":
            wall |= y == 0
        elif w == "top":
            wall |= y == ny - 1
        else:
            raise ValueError(f"unknown wall {w!r}")
    obst = np.zeros(x.shape, dtype=bool)
    for m in obstacles:
        obst |= m(x, y, z)
    nd = np.zeros(x.shape, dtype=np.uint8)
    nd[wall | obst] = SOLID
    free = ~(wall | obst)
    if inlet:
        nd[(z == 0) & free] = INLET
    if outlet:
        nd[(z == nz - 1) & free] = OUTLET
    return nd.ravel()


def lid_cavity(nx: int, ny: int, nz: int) -> np.ndarray:
    x, y, z = grid(nx, ny, nz)
    nd = np.zeros(x.shape, dtype=np.uint8)
    nd[(x == 0) | (x == nx - 1) | (y == 0) | (z == 0) | (z == nz - 1)] = SOLID
    nd[y == ny - 1] = LID
    return nd.ravel()


def periodic_box(nx: int, ny: int, nz: int, obstacles=()) -> np.ndarray:
    """All-fluid periodic box, optionally with solid obstacles."""
    return channel(nx, ny, nz, walls=(), obstacles=obstacles, inlet=False, outlet=False)


def rsa_spheres(nx: int, ny: int, nz: int, r_min: float, r_max: float, porosity: float,
                seed: int, max_attempts: int = 2_000_000, walls=True, solid_mask=None,
                z_margin: float = 0.0):
    """Seeded random sequential addition of non-overlapping spheres.

    Radii are uniform in [r_min, r_max], centres uniform in the box; a
    candidate is rejected if it overlaps an accepted sphere.  Spheres are
    added until the solid fraction reaches 1 - porosity or max_attempts
    candidates have been drawn.  The solid fraction is the spheres' volume
    4/3 pi r^3, or, with solid_mask (a (nz, ny, nx) bool array of solid
    nodes to start from, e.g. the walls; updated in place), the nodes of
    that mask together with the nodes inside the spheres (d^2 < r^2), so
    that the lattice's own porosity is the one reached.  z_margin > 0 keeps
    every sphere that many nodes clear of the planes z = 0 and z = nz - 1
    (centre z uniform in [z_margin + r, nz - 1 - z_margin - r]).  A cell
    list keyed on the largest diameter keeps the overlap test local.
    Returns (centres, radii) as float64 arrays.
    """
    rng = np.random.default_rng(seed)
    cell = 2.0 * r_max
    ncx, ncy, ncz = (max(1, int(np.ceil(n / cell))) for n in (nx, ny, nz))
    # cell list as a dense array, padded by one cell on every side so that
    # the 27 cells around any cell are one slice; K slots per cell (no two
    # accepted centres are closer than 2 r_min, so a cell of edge 2 r_max
    # holds at most ~(r_max / r_min + 1)^3 of them)
    K = int(np.ceil((r_max / r_min + 1.0) ** 3)) + 1
    grid = np.full((ncx + 2, ncy + 2, ncz + 2, K), -1, np.int64)
    fill = np.zeros((ncx + 2, ncy + 2, ncz + 2), np.int64)
    cap = 1024
    C = np.empty((cap, 3), np.float64)
    R = np.empty(cap, np.float64)
    n_acc = 0
    target = (1.0 - porosity) * nx * ny * nz
    solid = 0.0 if solid_mask is None else float(solid_mask.sum())
    attempts = 0
    # the generator's doubles are drawn in batches and mapped exactly as
    # Generator.uniform maps them (low + (high - low) * u, one double for the
    # radius, then three for the centre), so the sequence is the one
    # per-candidate uniform() calls give
    batch = np.empty(0)
    pos = 0
    lo_c = np.zeros(3)
    while solid < target and attempts < max_attempts:
        if pos + 4 > batch.size:
            batch = rng.random(4 * 4096)
            pos = 0
        attempts += 1
        u = batch[pos:pos + 4]
        pos += 4
        r = r_min + (r_max - r_min) * u[0]
        if z_margin > 0:
            lo = (0.0, 0.0, z_margin + r)
            hi = (float(nx), float(ny), nz - 1 - z_margin - r)
        else:
            lo = (0.0, 0.0, 0.0)
            hi = (float(nx), float(ny), float(nz))
        lo_c[:] = lo
        c = lo_c + (np.array(hi) - lo_c) * u[1:4]
        ci = (min(int(c[0] / cell), ncx - 1), min(int(c[1] / cell), ncy - 1),
              min(int(c[2] / cell), ncz - 1))
        nb = grid[ci[0]:ci[0] + 3, ci[1]:ci[1] + 3, ci[2]:ci[2] + 3].ravel()
        nb = nb[nb >= 0]
        if nb.size:
            d2 = ((C[nb] - c) ** 2).sum(axis=1)
            if np.any(d2 < (R[nb] + r) ** 2):
                continue
        if n_acc == cap:
            cap *= 2
            C = np.resize(C, (cap, 3))
            R = np.resize(R, cap)
        C[n_acc] = c
        R[n_acc] = r
        g = (ci[0] + 1, ci[1] + 1, ci[2] + 1)
        if fill[g] >= K:
            raise RuntimeError("rsa_spheres: cell list overflow")
        grid[g + (fill[g],)] = n_acc
        fill[g] += 1
        n_acc += 1
        if solid_mask is None:
            solid += 4.0 / 3.0 * np.pi * r ** 3
        else:
            box, inside = _sphere_box(nx, ny, nz, c, r)
            sub = solid_mask[box]
            solid += float(np.count_nonzero(inside & ~sub))
            sub |= inside
    centres, radii = C[:n_acc], R[:n_acc]
    return np.array(centres, dtype=np.float64).reshape(-1, 3), np.array(radii, dtype=np.float64)


def _sphere_box(nx, ny, nz, c, r):
    """(bounding-box slices, nodes strictly inside the sphere there)."""
    cx, cy, cz = c
    x0, x1 = max(0, int(np.floor(cx - r))), min(nx, int(np.ceil(cx + r)) + 1)
    y0, y1 = max(0, int(np.floor(cy - r))), min(ny, int(np.ceil(cy + r)) + 1)
    z0, z1 = max(0, int(np.floor(cz - r))), min(nz, int(np.ceil(cz + r)) + 1)
    xx = (np.arange(x0, x1) - cx)[None, None, :] ** 2
    yy = (np.arange(y0, y1) - cy)[None, :, None] ** 2
    zz = (np.arange(z0, z1) - cz)[:, None, None] ** 2
    return (slice(z0, z1), slice(y0, y1), slice(x0, x1)), xx + yy + zz < r * r


def sphere_pack_mask(nx, ny, nz, centres, radii):
    """Obstacle mask for a sphere pack, evaluated per sphere bounding box."""
    mask = np.zeros((nz, ny, nx), dtype=bool)
    for c, r in zip(centres, radii):
        box, inside = _sphere_box(nx, ny, nz, c, r)
        mask[box] |= inside
    return mask
