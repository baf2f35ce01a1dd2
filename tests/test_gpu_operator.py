"""GPU parity of the operator applies against the oracle (FP64 <= 1e-12 relative
in the max-norm over rows, the reference's -check_matfree measure; the FP32
apply's error against the same FP64 oracle is bounded and reported)."""
import numpy as np
import pytest
import torch

import oracle
from helpers import oracle_mult, oracle_rows, oracle_rhs, rel_max, to_dev, to_host, xs_arrays
from ubolt_amd import problems as P
from ubolt_amd import core

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 2e-5


def _cases():
    return {
        "fd2d_vac": lambda: P.fd_box_2d_central_source(n=12, sn=4),
        "fd2d_inflow_window_reflect": lambda: P.fd_box_2d_central_source(
            n=10, sn=6, bcs={4: {"type": "vacuum", "inflow": 3.0, "window": [0.2, 0.7]}, 2: "reflect",
                             1: {"type": "vacuum", "inflow": 1.5}, 3: "reflect"}),
        "fd3d_vac": lambda: _fd3d(6, 4, None),
        "fd3d_reflect": lambda: _fd3d(5, 6, {6: "reflect", 3: "reflect", 1: {"type": "vacuum", "inflow": 2.0,
                                                                              "window": [0.1, 0.6, 0.3, 0.9]}}),
        "dg0_tri": lambda: P.dg_box_2d(6, 5, 4, order=0, tris=True),
        "dg0_quad": lambda: P.dg_box_2d(7, 6, 6, order=0),
        "dg0_hex": lambda: P.dg_box_3d(4, 4, order=0),
        "dg1_tri": lambda: P.dg_box_2d(6, 5, 4, order=1, tris=True),
        "dg1_quad_reflect": lambda: P.dg_box_2d(6, 6, 6, order=1, bcs={1: "reflect", 4: "reflect",
                                                                       2: {"type": "vacuum", "inflow": 2.0}}),
        "dg1_hex": lambda: P.dg_box_3d(4, 4, order=1),
        "dg1_hex_reflect": lambda: P.dg_box_3d(3, 6, order=1, bcs={5: "reflect", 2: "reflect"}),
        "dg1_tet": lambda: _tets(),
        "dg1_delaunay": lambda: P.dg1_delaunay(n_side=12, sn=8, seed=7),
        "dg1_quad_s12": lambda: P.dg_box_2d(4, 3, 12, order=1, bcs={2: "reflect"}),
        "dg0_hex_s8": lambda: P.dg_box_3d(3, 8, order=0, bcs={1: "reflect"}),
        "fd3d_s8_reflect": lambda: _fd3d(4, 8, {5: "reflect", 4: "reflect"}),
    }


def _fd3d(n, sn, bcs):
    q = core.Quadrature(3, sn)
    d = core.Disc.structured(q, [n, n + 1, n - 1], [1.0, 1.3, 0.8], bcs if bcs is not None else P.VACUUM_3D)
    rng = np.random.default_rng(11)
    nc = d.n_cells
    st = rng.uniform(0.5, 5.0, nc)
    return P.Problem("fd3d", q, d, st, rng.uniform(0.1, 0.9, nc) * st, np.ones(nc))


def _tets():
    coords, hexes = P.box_hexes(3, 3, 2)
    # 6 tetrahedra per hex along the 0-6 diagonal
    t = [(0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)]
    tets = np.concatenate([hexes[:, list(x)] for x in t])
    q = core.Quadrature(3, 4)
    mesh = core.Mesh(coords, tets)
    d = core.Disc.dg(q, mesh, 1, P.VACUUM_3D)
    rng = np.random.default_rng(12)
    st = rng.uniform(0.1, 10.0, d.n_cells)
    return P.Problem("dg1_tet", q, d, st, 0.7 * st, np.ones(d.n_cells))


CASES = _cases()


@pytest.fixture(scope="module", params=sorted(CASES))
def built(request):
    prob = CASES[request.param]()
    op = prob.operator()
    print(f"[case] {request.param}: rows={prob.disc.n_rows}", flush=True)
    return request.param, prob, op


# tile, wavefront per cell, multi-cell tile (2 and 4 cells per thread), pencil on
# a structured box (8, 16, 32 cells per workgroup), four-angle tile, box tile
# (2 and 4 rows per thread), strip march, z-march on a 3D FD box (4 and 8
# cells per lane column); a variant that does not apply to a case runs the
# tile kernel
VARIANTS = [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12]


@pytest.mark.parametrize("variant", VARIANTS)
def test_mult_fp64(gpu, built, variant):
    name, prob, op = built
    op.set_variant(core.FP64, variant)
    x = np.random.default_rng(1).standard_normal(prob.disc.n_rows)
    want = oracle_mult(prob, x)
    y = op.vec()
    op.mult(to_dev(x), y)
    torch.cuda.synchronize()
    err = rel_max(to_host(y), want)
    print(f"{name}: fp64 apply (variant {variant}) rel err {err:.3e}")
    assert err <= FP64_TOL


@pytest.mark.parametrize("variant", VARIANTS)
def test_mult_fp32(gpu, built, variant):
    name, prob, op = built
    op.set_variant(core.FP32, variant)
    x = np.random.default_rng(2).standard_normal(prob.disc.n_rows)
    want = oracle_mult(prob, x)
    y = op.vec(core.FP32)
    op.mult(to_dev(x, torch.float32), y)
    torch.cuda.synchronize()
    err = rel_max(to_host(y), want)
    print(f"{name}: fp32 apply (variant {variant}) rel err vs fp64 oracle {err:.3e}")
    assert err <= FP32_TOL


@pytest.mark.parametrize("variant", VARIANTS)
def test_mult_unscattered(gpu, built, variant):
    name, prob, op = built
    op.set_variant(core.FP64, variant)
    op.set_variant(core.FP32, variant)
    x = np.random.default_rng(3).standard_normal(prob.disc.n_rows)
    want = oracle_mult(prob, x, scatter=False)
    y = op.vec()
    op.mult_unscattered(to_dev(x), y)
    y32 = op.vec(core.FP32)
    op.mult_unscattered(to_dev(x, torch.float32), y32)
    torch.cuda.synchronize()
    assert rel_max(to_host(y), want) <= FP64_TOL
    assert rel_max(to_host(y32), want) <= FP32_TOL


def test_tuner_reports(gpu, built):
    name, prob, op = built
    op.set_variant(core.FP64, -1)
    t = op.tuning()
    print(f"{name}: tuned {t}")
    assert 0 <= t["variant_fp64"] < 13 and 0 <= t["variant_fp32"] < 13 and t["us"][0] > 0


def test_diagonal(gpu, built):
    name, prob, op = built
    R = oracle_rows(prob)
    r = np.arange(R.n_rows)
    cols = R.col.reshape(R.n_rows, R.S)
    vals = R.val.reshape(R.n_rows, R.S)
    want = np.where(cols == r[:, None], vals, 0.0).sum(axis=1)
    d = op.vec()
    op.diagonal(d)
    torch.cuda.synchronize()
    assert rel_max(to_host(d), want) <= FP64_TOL


def test_moments(gpu, built):
    name, prob, op = built
    d, q = prob.disc, prob.quad
    psi = np.random.default_rng(4).standard_normal(d.n_rows)
    errs = {}
    for M in (1, 1 + d.dim):
        w = q.w[:, None] * np.concatenate([np.ones((d.n_angles, 1)), q.omega[:, :d.dim]], axis=1)[:, :M]
        want = psi.reshape(d.n_nodes, d.n_angles) @ w
        for prec, tol in ((core.FP64, 1e-13), (core.FP32, 1e-5)):
            phi = torch.zeros(d.n_nodes * M, dtype=core.torch_dtype(prec), device="cuda")
            op.moments(to_dev(psi, core.torch_dtype(prec)), phi, M)
            torch.cuda.synchronize()
            errs[(M, prec)] = (rel_max(to_host(phi).reshape(-1, M), want), tol)
    print(name, "moments rel err", errs)
    assert all(e <= tol for e, tol in errs.values()), errs
    # M = 1 against the oracle's integral
    phi = torch.zeros(d.n_nodes, dtype=torch.float64, device="cuda")
    op.moments(to_dev(psi), phi, 1)
    torch.cuda.synchronize()
    assert rel_max(to_host(phi), oracle.moments(psi, q.w)) <= 1e-13


def test_block_inverse(gpu, built):
    name, prob, op = built
    d = prob.disc
    R = oracle_rows(prob)
    inv, n_sing = oracle.block_setup(R, d.nb, d.n_cells, d.n_angles)
    assert n_sing == 0
    x = np.random.default_rng(5).standard_normal(d.n_rows)
    want = oracle.block_apply(inv, d.nb, d.n_cells, d.n_angles, x)
    bi = core.BlockInverse(op)
    bi.setup()
    y = op.vec()
    bi.apply(to_dev(x), y)
    y32 = op.vec(core.FP32)
    bi.apply(to_dev(x, torch.float32), y32)
    torch.cuda.synchronize()
    assert rel_max(to_host(y), want) <= FP64_TOL
    assert rel_max(to_host(y32), want) <= FP32_TOL


def test_rhs_multigroup(gpu):
    """ghost inflow + source + transfers from the groups above, in one launch,
    and scatter_source, against the oracle, on a 3-group problem"""
    prob = P.dg_box_2d(5, 4, 4, order=1, G=3, bcs={4: {"type": "vacuum", "inflow": 2.0, "window": [0.2, 0.8]}})
    op = prob.operator()
    d = prob.disc
    rng = np.random.default_rng(6)
    psis = [rng.standard_normal(d.n_rows) for _ in range(3)]
    cache = []
    for g in range(3):
        op.set_scalar_flux(g, to_dev(psis[g]))
        cache.append(oracle.moments(psis[g], prob.quad.w))
    for g in range(3):
        b = op.vec()
        op.rhs(g, b)
        b32 = op.vec(core.FP32)
        op.rhs(g, b32)
        torch.cuda.synchronize()
        want = oracle_rhs(prob, g, cache)
        assert rel_max(to_host(b), want) <= 1e-13
        assert rel_max(to_host(b32), want) <= 1e-6
    _, ss, _ = xs_arrays(prob)
    b = op.vec()
    op.scatter_source(0, 2, b)
    torch.cuda.synchronize()
    want = oracle.broadcast_add(d.nb, d.n_angles, ss[0, 2] / prob.quad.sum_weights, cache[0], np.zeros(d.n_rows))
    assert rel_max(to_host(b), want) <= 1e-13


def test_geometry_classes(gpu):
    """uniform boxes collapse to one geometry record per boundary pattern (a
    cell's faces on each axis: low boundary, interior or high boundary, so at
    most 3^dim); a jittered Delaunay mesh keeps one per cell. The applies above
    agree with the oracle's per-cell geometry either way."""
    box = P.dg_box_3d(6, 2, order=1).operator()
    quad = P.dg_box_2d(9, 7, 2, order=1).operator()
    dela = P.dg1_delaunay(n_side=8, sn=2, seed=1)
    d_op = dela.operator()
    print("classes: hex box", box.geometry_classes(), "quad box", quad.geometry_classes(), "delaunay",
          d_op.geometry_classes(), "of", dela.disc.n_cells)
    assert box.geometry_classes() <= 27 and quad.geometry_classes() <= 9
    assert d_op.geometry_classes() == dela.disc.n_cells


# Boxes wide and tall enough for several box tiles per axis (x halos from a
# neighbouring tile, rows beyond a tile, a partial last tile), with vacuum,
# inflow and reflective sides: the box-tile kernel against the oracle.
BOX_CASES = {
    "dg1_quad_tiles": lambda: P.dg_box_2d(23, 13, 4, order=1, bcs={1: "reflect", 3: "reflect",
                                                                   2: {"type": "vacuum", "inflow": 1.5}}),
    "dg1_hex_tiles": lambda: P.dg_box_3d(5, 8, order=1, bcs={3: "reflect", 6: "reflect"}),
    # every side vacuum: no list phase; three bands of the strip march, a
    # partial last strip and band
    "dg1_quad_bands_vacuum": lambda: P.dg_box_2d(23, 150, 6, order=1),
    "dg1_quad_bands_s12": lambda: P.dg_box_2d(9, 131, 12, order=1, bcs={4: "reflect"}),
    "dg1_hex_bands": lambda: P.dg_box_3d(4, 2, order=1),
}


@pytest.mark.parametrize("case", sorted(BOX_CASES))
def test_box_tile_apply(gpu, case):
    prob = BOX_CASES[case]()
    op = prob.operator()
    us = op.tuning()["us"]
    assert all(us[v] > 0 for v in (8, 9, 21, 22)), f"box tile not applicable: {us}"
    # the strip march holds one halo element (nb x angles of a cell) per thread
    strip = prob.disc.nb * prob.disc.n_angles <= 256
    assert (us[10] > 0 and us[23] > 0) == strip, f"strip march applicability: {us}"
    x = np.random.default_rng(4).standard_normal(prob.disc.n_rows)
    want = oracle_mult(prob, x)
    want_l = oracle_mult(prob, x, scatter=False)
    for v in (8, 9, 10) if strip else (8, 9):
        op.set_variant(core.FP64, v)
        op.set_variant(core.FP32, v)
        y, y32, yl = op.vec(), op.vec(core.FP32), op.vec()
        op.mult(to_dev(x), y)
        op.mult(to_dev(x, torch.float32), y32)
        op.mult_unscattered(to_dev(x), yl)
        torch.cuda.synchronize()
        e64, e32, el = rel_max(to_host(y), want), rel_max(to_host(y32), want), rel_max(to_host(yl), want_l)
        print(f"{case} variant {v}: fp64 {e64:.2e} fp32 {e32:.2e} unscattered {el:.2e}")
        assert e64 <= FP64_TOL and el <= FP64_TOL and e32 <= FP32_TOL


# the z-march (3D FD): runs of lanes that end inside the box's last x tile,
# several z chunks with a partial last one, every face reflective or with
# inflow, S4 (ten lane columns per workgroup), S8 (three) and S12 (one)
ZM_CASES = {
    "fd3d_zm_s4_vac": lambda: _fd3d_shape([13, 7, 70], 4, None),
    "fd3d_zm_s8_reflect": lambda: _fd3d_shape([13, 9, 37], 8, {1: "reflect", 2: "reflect", 3: "reflect",
                                                                4: "reflect", 5: "reflect", 6: "reflect"}),
    "fd3d_zm_s12_mixed": lambda: _fd3d_shape([5, 6, 33], 12, {2: "reflect", 3: {"type": "vacuum", "inflow": 1.5},
                                                              5: "reflect"}),
}


def _fd3d_shape(n, sn, bcs):
    q = core.Quadrature(3, sn)
    d = core.Disc.structured(q, n, [1.0, 0.7, 2.1], bcs if bcs is not None else P.VACUUM_3D)
    rng = np.random.default_rng(17)
    nc = d.n_cells
    st = rng.uniform(0.5, 5.0, nc)
    return P.Problem("fd3d_zm", q, d, st, rng.uniform(0.1, 0.9, nc) * st, np.ones(nc))


@pytest.mark.parametrize("case", sorted(ZM_CASES))
def test_fd_zmarch_apply(gpu, case):
    prob = ZM_CASES[case]()
    op = prob.operator()
    us = op.tuning()["us"]
    assert all(us[v] > 0 for v in (11, 12, 24, 25)), f"z-march not applicable: {us}"
    x = np.random.default_rng(5).standard_normal(prob.disc.n_rows)
    want = oracle_mult(prob, x)
    want_l = oracle_mult(prob, x, scatter=False)
    for v in (11, 12):
        op.set_variant(core.FP64, v)
        op.set_variant(core.FP32, v)
        y, y32, yl = op.vec(), op.vec(core.FP32), op.vec()
        op.mult(to_dev(x), y)
        op.mult(to_dev(x, torch.float32), y32)
        op.mult_unscattered(to_dev(x), yl)
        torch.cuda.synchronize()
        e64, e32, el = rel_max(to_host(y), want), rel_max(to_host(y32), want), rel_max(to_host(yl), want_l)
        print(f"[case] {case} variant {v}: fp64 {e64:.2e} fp32 {e32:.2e} unscattered {el:.2e}", flush=True)
        assert e64 <= FP64_TOL and el <= FP64_TOL and e32 <= FP32_TOL
