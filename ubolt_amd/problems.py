"""Seeded synthetic problems: meshes and per-cell cross sections.

Random numbers come from numpy's PCG64 (np.random.default_rng(seed)), so a
(config, seed) pair is the same problem on every machine. Meshes:

* box_quads / box_tris / box_hexes: uniform boxes, vertices x fastest;
* delaunay_square: the Delaunay triangulation (scipy/qhull) of a jittered
  point lattice on the unit square with evenly spaced boundary points.

Labels follow PETSc's box "Face Sets" (core.Mesh labels the bounding box).
"""
import numpy as np

from . import core

VACUUM_2D = {1: "vacuum", 2: "vacuum", 3: "vacuum", 4: "vacuum"}
VACUUM_3D = {k: "vacuum" for k in range(1, 7)}


def box_quads(nx, ny, lx=1.0, ly=1.0):
    x = np.linspace(0.0, lx, nx + 1)
    y = np.linspace(0.0, ly, ny + 1)
    X, Y = np.meshgrid(x, y)  # [j, i]
    coords = np.stack([X.ravel(), Y.ravel()], axis=1)
    v = lambda i, j: j * (nx + 1) + i
    I, J = np.meshgrid(np.arange(nx), np.arange(ny))
    I, J = I.ravel(), J.ravel()  # cell c = j * nx + i, x fastest
    cells = np.stack([v(I, J), v(I + 1, J), v(I + 1, J + 1), v(I, J + 1)], axis=1)
    return coords, cells


def box_tris(nx, ny, lx=1.0, ly=1.0):
    coords, q = box_quads(nx, ny, lx, ly)
    tris = np.concatenate([q[:, [0, 1, 2]], q[:, [0, 2, 3]]], axis=1).reshape(-1, 3)
    return coords, tris


def box_hexes(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0):
    x = np.linspace(0.0, lx, nx + 1)
    y = np.linspace(0.0, ly, ny + 1)
    z = np.linspace(0.0, lz, nz + 1)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    coords = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    v = lambda i, j, k: (k * (ny + 1) + j) * (nx + 1) + i
    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    cells = np.stack([v(I, J, K), v(I + 1, J, K), v(I + 1, J + 1, K), v(I, J + 1, K),
                      v(I, J, K + 1), v(I + 1, J, K + 1), v(I + 1, J + 1, K + 1), v(I, J + 1, K + 1)], axis=1)
    return coords, cells


def delaunay_square(n_side, seed, jitter=0.3):
    """Delaunay triangles of the unit square: n_side - 1 evenly spaced points
    per side plus the corners, and an (n_side - 1)^2 lattice of interior points
    jittered by up to `jitter` of the spacing. About 2 n_side^2 triangles."""
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    h = 1.0 / n_side
    t = np.linspace(0.0, 1.0, n_side + 1)
    bnd = np.concatenate([np.stack([t, np.zeros_like(t)], 1), np.stack([t, np.ones_like(t)], 1),
                          np.stack([np.zeros_like(t[1:-1]), t[1:-1]], 1), np.stack([np.ones_like(t[1:-1]), t[1:-1]], 1)])
    g = (np.arange(1, n_side) * h)
    GX, GY = np.meshgrid(g, g)
    inner = np.stack([GX.ravel(), GY.ravel()], 1)
    inner = inner + rng.uniform(-jitter * h, jitter * h, size=inner.shape)
    pts = np.concatenate([bnd, inner])
    tri = Delaunay(pts)
    cells = tri.simplices.astype(np.int32)
    # drop any zero-area sliver qhull may emit on the collinear boundary
    p = pts[cells]
    area = 0.5 * np.abs((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) -
                        (p[:, 1, 1] - p[:, 0, 1]) * (p[:, 2, 0] - p[:, 0, 0]))
    cells = cells[area > 1e-14]
    # order cells along a space-filling curve of their centroids so upwind
    # neighbours sit close in memory
    cen = pts[cells].mean(axis=1)
    cells = cells[np.lexsort((cen[:, 0], np.floor(cen[:, 1] / (4 * h))))]
    return pts, cells


# ---------------------------------------------------------------------------
# The benchmark configurations
# ---------------------------------------------------------------------------

class Problem:
    """A built problem: quadrature, discretisation and host cross sections."""

    def __init__(self, name, quad, disc, sigma_t, sigma_s, source, n_groups=1):
        self.name, self.quad, self.disc = name, quad, disc
        self.sigma_t, self.sigma_s, self.source, self.n_groups = sigma_t, sigma_s, source, n_groups

    def operator(self):
        op = core.Operator(self.disc, self.n_groups)
        op.set_xs(self.sigma_t, self.sigma_s, self.source)
        op.set_group(0)
        return op


def fd_box_2d_central_source(n=32, sn=4, sigma_t=1.0, sigma_s=0.5, bcs=None):
    """2D structured FD unit square, n x n cells, uniform sigma_t / sigma_s, unit
    isotropic source on cells whose centre lies in [0.375, 0.625]^2, vacuum."""
    q = core.Quadrature(2, sn)
    disc = core.Disc.structured(q, [n, n], [1.0, 1.0], bcs if bcs is not None else VACUUM_2D)
    h = 1.0 / n
    c = (np.arange(n) + 0.5) * h
    CX, CY = np.meshgrid(c, c)
    inside = (CX.ravel() >= 0.375) & (CX.ravel() <= 0.625) & (CY.ravel() >= 0.375) & (CY.ravel() <= 0.625)
    nc = n * n
    return Problem("fd2d_%d_s%d" % (n, sn), q, disc, np.full(nc, sigma_t), np.full(nc, sigma_s),
                   inside.astype(np.float64))


def dg1_delaunay(n_side=100, sn=8, seed=2, order=1):
    """2D DG1 on a seeded Delaunay triangulation of the unit square (about
    2 n_side^2 triangles), sigma_t ~ U[0.1, 10], scattering ratio ~ U[0.5, 0.99],
    unit source, vacuum."""
    rng = np.random.default_rng(seed)
    pts, cells = delaunay_square(n_side, seed)
    q = core.Quadrature(2, sn)
    mesh = core.Mesh(pts, cells)
    disc = core.Disc.dg(q, mesh, order, VACUUM_2D)
    nc = disc.n_cells
    st = rng.uniform(0.1, 10.0, nc)
    ratio = rng.uniform(0.5, 0.99, nc)
    return Problem("dg%d_delaunay_%d_s%d" % (order, nc, sn), q, disc, st, ratio * st, np.ones(nc))


def downscatter_xs(rng, G, nc, st_lo=0.5, st_hi=5.0, cap=0.9):
    """sigma_t ~ U[st_lo, st_hi] per cell per group; a lower-triangular
    (downscatter) sigma_s(g' -> g), g >= g', whose row sums (total scattering
    out of g') are at most cap * sigma_t(g')."""
    st = rng.uniform(st_lo, st_hi, (G, nc))
    ss = np.zeros((G, G, nc))
    for gf in range(G):
        frac = rng.uniform(0.0, 1.0, (G - gf, nc))
        total = rng.uniform(0.2, cap, nc)
        frac = frac / frac.sum(axis=0) * total
        for k, gt in enumerate(range(gf, G)):
            ss[gf, gt] = frac[k] * st[gf]
    return st, ss


def fd_cube_multigroup(n=128, sn=8, G=8, seed=3):
    """3D structured cube n^3 (the DG0 hex stencil), S8, G groups with seeded
    downscatter, source in group 0 only, vacuum."""
    rng = np.random.default_rng(seed)
    q = core.Quadrature(3, sn)
    disc = core.Disc.structured(q, [n, n, n], [1.0, 1.0, 1.0], VACUUM_3D)
    nc = n ** 3
    st, ss = downscatter_xs(rng, G, nc)
    src = np.zeros((G, nc))
    src[0] = 1.0
    return Problem("fd3d_%d_s%d_g%d" % (n, sn, G), q, disc, st, ss, src, n_groups=G)


def dg_box_2d(nx, ny, sn, order=1, tris=False, seed=4, st_range=(0.1, 10.0), c_range=(0.3, 0.99), G=1, bcs=None):
    rng = np.random.default_rng(seed)
    coords, cells = (box_tris if tris else box_quads)(nx, ny)
    q = core.Quadrature(2, sn)
    mesh = core.Mesh(coords, cells)
    disc = core.Disc.dg(q, mesh, order, bcs if bcs is not None else VACUUM_2D)
    nc = disc.n_cells
    if G == 1:
        st = rng.uniform(*st_range, nc)
        ss = rng.uniform(*c_range, nc) * st
        src = np.ones(nc)
    else:
        st, ss = downscatter_xs(rng, G, nc, st_range[0], st_range[1], c_range[1])
        src = np.zeros((G, nc))
        src[0] = 1.0
    return Problem("dg%d_box_%dx%d_s%d" % (order, nx, ny, sn), q, disc, st, ss, src, n_groups=G)


def dg_box_3d(n, sn, order=1, seed=5, G=1, block=None, bcs=None):
    rng = np.random.default_rng(seed)
    coords, cells = box_hexes(n, n, n)
    q = core.Quadrature(3, sn)
    mesh = core.Mesh(coords, cells)
    disc = core.Disc.dg(q, mesh, order, bcs if bcs is not None else VACUUM_3D)
    nc = disc.n_cells
    if block:
        nbk = n // block
        bst, bss = downscatter_xs(rng, G, nbk ** 3)
        i = np.arange(n) // block
        K, J, I = np.meshgrid(i, i, i, indexing="ij")
        bid = ((K * nbk + J) * nbk + I).ravel()
        st, ss = bst[:, bid], bss[:, :, bid]
    else:
        st, ss = downscatter_xs(rng, G, nc)
    src = np.zeros((G, nc))
    src[0] = 1.0
    return Problem("dg%d_hex_%d_s%d_g%d" % (order, n, sn, G), q, disc, st, ss, src, n_groups=G)


def batch_quads(n_problems=64, seed=11, side=(32, 256), sns=(4, 6, 8, 10, 12), groups=(1, 4), c_range=(0.3, 0.99)):
    """The batched ensemble: n_problems 2D DG1 problems on quad boxes of the
    unit square, n_x and n_y drawn independently from U{side}, the S_N order
    from sns, the group count from U{groups}, sigma_t ~ U[0.1, 10], scattering
    ratio ~ U[c_range] (downscatter tables when G > 1), unit source in group 0,
    vacuum. Problem k uses seed * 1000 + k for its cross sections."""
    rng = np.random.default_rng(seed)
    probs = []
    for k in range(n_problems):
        nx, ny = (int(v) for v in rng.integers(side[0], side[1] + 1, 2))
        sn = int(rng.choice(sns))
        G = int(rng.integers(groups[0], groups[1] + 1))
        probs.append(dg_box_2d(nx, ny, sn, order=1, seed=seed * 1000 + k, c_range=c_range, G=G))
    return probs
