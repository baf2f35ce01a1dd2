"""Each benchmark configuration run on the GPU at its stated size.

Where the oracle finishes in seconds the HIP path is compared with it directly
(the applies at full size, FP64 <= 1e-12, FP32 <= 2e-5 in the max-norm over
rows; solves through the true residual with the oracle's matrix, or against
the oracle's direct solve where one is affordable). At the sizes where it does
not, size-independent properties are checked at full size:

* the constant solution: with inflow psi0 * W on every boundary face and the
  source W (sigma_t - sigma_s) psi0, psi = psi0 (basis 0 for DG1) solves the
  transport equation exactly, so the apply of the constant vector reproduces
  the right-hand side (ghost inflow + source) to rounding;
* the dot-product identity <A x, y> = <x, A^T y> of the operator and of the
  sweep with their adjoints;
and the oracle comparison runs on the same configuration at a reduced size.
"""
import math

import numpy as np
import pytest
import scipy.sparse.linalg as spl
import torch

import oracle
import ubolt_amd as u
from helpers import oracle_full, oracle_mult, oracle_rhs, oracle_rows, oracle_solve, rel_max, to_dev, to_host, xs_arrays
from ubolt_amd import core
from ubolt_amd import problems as P

pytestmark = pytest.mark.gpu

FOUR_PI = 4.0 * math.pi


def phi_of(prob, psi):
    return oracle.moments(np.asarray(psi), prob.quad.w)


def constant_case(disc, quad, sigma_t, sigma_s, psi0=0.75):
    """an operator whose exact solution is psi0 on basis 0 (inflow psi0 W on
    every face of disc, source W (sigma_t - sigma_s) psi0), the rhs, psi0 as
    a vector"""
    W = quad.sum_weights
    prob = P.Problem("constant", quad, disc, sigma_t, sigma_s, W * (sigma_t - sigma_s) * psi0)
    op = prob.operator()
    b = op.vec()
    op.rhs(0, b)
    psi = torch.zeros(disc.n_rows, dtype=torch.float64, device="cuda")
    psi.view(disc.n_cells, disc.nb, disc.n_angles)[:, 0, :] = psi0
    return op, b, psi


def check_constant(op, b, psi, tag):
    y = torch.empty_like(psi)
    op.mult(psi, y)
    y32 = torch.empty(psi.numel(), dtype=torch.float32, device="cuda")
    op.mult(psi.float(), y32)
    torch.cuda.synchronize()
    scale = float(b.abs().max())
    e64 = float((y - b).abs().max()) / scale
    e32 = float((y32.double() - b).abs().max()) / scale
    print(f"[case] {tag}: constant solution, apply vs rhs fp64 {e64:.2e} fp32 {e32:.2e}", flush=True)
    assert e64 <= 1e-12 and e32 <= 2e-5


def check_dot_identity(fwd, adj, n, tag, tol=1e-12, seed=0):
    g = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    y = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    ax, aty = torch.empty_like(x), torch.empty_like(y)
    fwd(x, ax)
    adj(y, aty)
    torch.cuda.synchronize()
    lhs, rhs = float(torch.dot(ax, y)), float(torch.dot(x, aty))
    # relative to the sum of magnitudes the dot products are rounded against
    mag = float(torch.dot(ax.abs(), y.abs()))
    err = abs(lhs - rhs) / mag
    print(f"[case] {tag}: <Ax,y> {lhs:.12e} <x,A^T y> {rhs:.12e} rel {err:.2e}", flush=True)
    assert err <= tol


# --------------------------------------------------------------------- config 1
def test_c1_fd2d_32_s4_dsa():
    """2D FD 32 x 32, S4, sigma_t 1, sigma_s 0.5, central source, vacuum; FP64
    GMRES with DSA: driven to 1e-12 the scalar flux is within 1e-8 of the
    oracle's direct solve; driven to 1e-8 the true residual is at most 1e-8"""
    prob = P.fd_box_2d_central_source(n=32, sn=4)
    op = prob.operator()
    dsa = u.DSA(op)
    dsa.set_group(0)
    b = op.vec()
    op.rhs(0, b)
    bh = to_host(b)
    want = oracle_solve(prob, bh)
    A = oracle_full(prob)
    for rtol in (1e-12, 1e-8):
        sol = u.Solver(op, dsa, rtol=rtol, max_it=400)
        sol.refresh()
        x = op.vec()
        res = sol.solve(b, x)
        torch.cuda.synchronize()
        xh = to_host(x)
        tr = np.linalg.norm(bh - A @ xh) / np.linalg.norm(bh)
        ferr = np.linalg.norm(phi_of(prob, xh) - phi_of(prob, want)) / np.linalg.norm(phi_of(prob, want))
        print(f"[case] C1 rtol {rtol:g}: its {res['iterations']} true residual {tr:.2e} flux vs oracle {ferr:.2e}",
              flush=True)
        assert res["reason"] == 1 and tr <= rtol * 1.05
        if rtol == 1e-12:
            assert ferr <= 1e-8


# --------------------------------------------------------------------- config 2
@pytest.fixture(scope="module")
def c2():
    prob = P.dg1_delaunay(n_side=100, sn=8, seed=2)
    op = prob.operator()
    return prob, op, oracle_rows(prob)


def test_c2_delaunay_apply_full(c2):
    """the 20,000-triangle S8 mesh: the tuned applies and the tile kernel in
    both precisions against the oracle's rows"""
    prob, op, R = c2
    assert prob.disc.n_cells >= 19000
    x = np.random.default_rng(1).standard_normal(prob.disc.n_rows)
    want = oracle_mult(prob, x, rows=R)
    t = op.tuning()
    for variant in sorted({0, t["variant_fp64"], t["variant_fp32"]}):
        op.set_variant(core.FP64, variant)
        op.set_variant(core.FP32, variant)
        y, y32 = op.vec(), op.vec(core.FP32)
        op.mult(to_dev(x), y)
        op.mult(to_dev(x, torch.float32), y32)
        torch.cuda.synchronize()
        e64, e32 = rel_max(to_host(y), want), rel_max(to_host(y32), want)
        print(f"[case] C2 variant {variant}: fp64 {e64:.2e} fp32 {e32:.2e}", flush=True)
        assert e64 <= 1e-12 and e32 <= 2e-5
    op.set_variant(core.FP64, t["variant_fp64"])
    op.set_variant(core.FP32, t["variant_fp32"])


@pytest.mark.parametrize("mixed", [False, True])
def test_c2_delaunay_solve_full(c2, mixed):
    """FP64 FGMRES and FP32-apply refinement on the full mesh, true residual
    with the oracle's matrix"""
    prob, op, R = c2
    sol = u.Solver(op, None, rtol=1e-8, max_it=400, restart=30, mixed=mixed, pc_dsa=False)
    sol.refresh()
    b = op.vec()
    op.rhs(0, b)
    x = op.vec()
    res = sol.solve(b, x)
    torch.cuda.synchronize()
    bh, xh = to_host(b), to_host(x)
    A = oracle_full(prob)
    tr = np.linalg.norm(bh - A @ xh) / np.linalg.norm(bh)
    print(f"[case] C2 solve mixed={mixed}: its {res['iterations']} reported {res['rel_residual']:.2e} "
          f"true {tr:.2e}", flush=True)
    assert res["reason"] == 1 and tr <= 1.05e-8
    assert abs(tr - res["rel_residual"]) <= 1e-3 * tr + 1e-14


# --------------------------------------------------------------------- config 3
def test_c3_cube128_s8_properties():
    """128^3, S8 (80 directions): the constant solution through the FP64 and
    FP32 applies, and the dot-product identity of the apply and the sweep,
    with group 0's cross sections of the 8-group downscatter problem"""
    rng = np.random.default_rng(3)
    n = 128
    q = core.Quadrature(3, 8)
    bcs = {k: {"type": "vacuum", "inflow": 0.75 * FOUR_PI} for k in range(1, 7)}
    disc = core.Disc.structured(q, [n, n, n], [1.0, 1.0, 1.0], bcs)
    st, ss = P.downscatter_xs(rng, 8, n ** 3)
    op, b, psi = constant_case(disc, q, st[0], ss[0, 0])
    assert disc.n_rows == 167_772_160
    check_constant(op, b, psi, "C3 128^3 S8")
    del b, psi
    check_dot_identity(op.mult, op.mult_transpose, disc.n_rows, "C3 apply")
    sw = u.Sweep(op)
    check_dot_identity(sw.apply, sw.apply_transpose, disc.n_rows, "C3 sweep", tol=1e-11)


def test_c3_cube32_s8_g8_against_oracle():
    """32^3, S8, the 8 downscatter groups: every group's apply and its rhs
    (source in group 0, transfers from the groups above through the scalar
    flux cache) against the oracle, then group Gauss-Seidel solves (FP32 apply
    in FP64 refinement) to a 1e-8 true residual with the oracle's matrix"""
    prob = P.fd_cube_multigroup(n=32, sn=8, G=8, seed=3)
    op = prob.operator()
    d = prob.disc
    rng = np.random.default_rng(5)
    x = rng.standard_normal(d.n_rows)
    cache = []
    for g in range(8):
        op.set_group(g)
        y = op.vec()
        op.mult(to_dev(x), y)
        psi = rng.standard_normal(d.n_rows)
        op.set_scalar_flux(g, to_dev(psi))
        b = op.vec()
        op.rhs(g, b)
        torch.cuda.synchronize()
        assert rel_max(to_host(y), oracle_mult(prob, x, g=g)) <= 1e-12, g
        assert rel_max(to_host(b), oracle_rhs(prob, g, cache)) <= 1e-13, g
        cache.append(phi_of(prob, psi))
    op2 = prob.operator()
    sol = u.Solver(op2, None, rtol=1e-8, max_it=400, mixed=True, pc_dsa=False)
    cache = []
    for g in range(8):
        op2.set_group(g)
        sol.refresh()
        b = op2.vec()
        op2.rhs(g, b)
        xg = op2.vec()
        res = sol.solve(b, xg)
        op2.set_scalar_flux(g, xg)
        torch.cuda.synchronize()
        bh, xh = to_host(b), to_host(xg)
        assert rel_max(bh, oracle_rhs(prob, g, cache)) <= 1e-12, g
        tr = np.linalg.norm(bh - oracle_full(prob, g) @ xh) / np.linalg.norm(bh)
        print(f"[case] C3 32^3 group {g}: its {res['iterations']} true residual {tr:.2e}", flush=True)
        assert res["reason"] == 1 and tr <= 1.05e-8, g
        cache.append(phi_of(prob, xh))


# --------------------------------------------------------------------- config 4
@pytest.fixture(scope="module")
def c4():
    probs = P.batch_quads(n_problems=64, seed=11)
    ops = [p.operator() for p in probs]
    b = u.Batch(ops)
    b.set_group(0)
    return probs, ops, b


def test_c4_ensemble_apply_every_problem(c4):
    """the 64-problem ensemble (S4-S12, up to 84 directions): the tile, box and
    strip bodies in both precisions against the oracle, problem by problem"""
    probs, ops, b = c4
    t = b.tuning()
    x = np.random.default_rng(2).standard_normal(b.n_rows)
    runs = {}
    for body in (0, 2, 3):
        for prec, dt in ((u.FP64, torch.float64), (u.FP32, torch.float32)):
            b.set_body(prec, body)
            y = b.vec(prec)
            b.mult(to_dev(x, dt), y)
            torch.cuda.synchronize()
            runs[(body, prec)] = to_host(y)
            del y
    b.set_body(u.FP64, t["body_fp64"])
    b.set_body(u.FP32, t["body_fp32"])
    worst = {}
    for p, prob in enumerate(probs):
        sl = slice(b.offsets[p], b.offsets[p + 1])
        want = oracle_mult(prob, x[sl])
        for (body, prec), yh in runs.items():
            err = rel_max(yh[sl], want)
            worst[(body, prec)] = max(worst.get((body, prec), 0.0), err)
            assert err <= (1e-12 if prec == u.FP64 else 2e-5), (p, prob.name, body, prec, err)
    print(f"[case] C4 worst per (body, prec): {worst}", flush=True)


def test_c4_ensemble_mixed_solve(c4):
    """group 0 of every problem by the batched FP32-apply refinement: every
    problem's true residual with its oracle matrix"""
    probs, ops, b = c4
    b.set_group(0)
    sol = u.BatchSolver(b, rtol=1e-8, max_it=600, restart=20, mixed=True, gs_passes=1)
    sol.refresh()
    rhs = b.vec()
    b.rhs(rhs)
    x = b.vec()
    res = sol.solve(rhs, x)
    torch.cuda.synchronize()
    bh, xh = to_host(rhs), to_host(x)
    worst = 0.0
    for p, prob in enumerate(probs):
        sl = slice(b.offsets[p], b.offsets[p + 1])
        ax = oracle_mult(prob, xh[sl])
        tr = np.linalg.norm(bh[sl] - ax) / np.linalg.norm(bh[sl])
        worst = max(worst, tr)
        assert res["reason"][p] == 1 and tr <= 1.05e-8, (p, prob.name, tr)
    print(f"[case] C4 mixed solve: max its {max(res['iterations'])}, worst true residual {worst:.2e}", flush=True)


# --------------------------------------------------------------------- config 5
def test_c5_hex64_s6_properties():
    """DG1 hexes 64^3, S6 (48 directions): the constant solution through the
    FP64 and FP32 applies, and the dot-product identity of the operator and
    of the sweep with their adjoints"""
    n = 64
    bcs = {k: {"type": "vacuum", "inflow": 0.75 * FOUR_PI} for k in range(1, 7)}
    base = P.dg_box_3d(n, 6, order=1, G=4, block=8, bcs=bcs)
    st, ss, _ = xs_arrays(base)
    op, b, psi = constant_case(base.disc, base.quad, st[1], ss[1, 1])
    assert base.disc.n_rows == 50_331_648
    check_constant(op, b, psi, "C5 64^3 S6")
    del b, psi
    check_dot_identity(op.mult, op.mult_transpose, base.disc.n_rows, "C5 apply")
    sw = u.Sweep(op)
    check_dot_identity(sw.apply, sw.apply_transpose, base.disc.n_rows, "C5 sweep", tol=1e-11)


def test_c5_hex64_adjoint_vs_fd():
    """the sensitivity workload at full size: adjoint gradients of the
    response against central differences of the forward solve, one cell per
    parameter kind"""
    from ubolt_amd import sensitivity as S

    n = 64
    prob = P.dg_box_3d(n, 6, order=1, G=4, block=n // 8)
    G, nb = 4, prob.disc.nb
    model = S.TransportModel(prob.disc, G, rtol=1e-13, dsa_inner_rtol=1e-3, dsa_inner_max_it=100)
    r = torch.as_tensor(S.box_response_weights(prob.disc, [0.7] * 3, [0.9] * 3), dtype=torch.float64, device="cuda")
    base = [np.ascontiguousarray(a, dtype=np.float64) for a in (prob.sigma_t, prob.sigma_s, prob.source)]
    xs = [torch.tensor(a, requires_grad=True) for a in base]
    J = (S.scalar_flux(model, *xs)[:, ::nb] * r).sum()
    J.backward()
    grads = [v.grad.numpy() for v in xs]
    # a cell inside the response box: strong sensitivities
    i = int(0.8 * n)
    c = (i * n + i) * n + i
    for arr, idx in ((0, (1, c)), (1, (2, 2, c)), (2, (0, c))):
        h = 2e-3 * (abs(base[arr][idx]) if arr != 2 else float(np.abs(base[2][0]).max()))
        Js = []
        for sgn in (1, -1):
            pert = [a.copy() for a in base]
            pert[arr][idx] += sgn * h
            with torch.no_grad():
                _, ph = model.forward_solve(*(torch.as_tensor(a) for a in pert))
                Js.append(float((ph[:, ::nb] * r).sum()))
        fd = (Js[0] - Js[1]) / (2 * h)
        ad = float(grads[arr][idx])
        err = abs(ad - fd) / abs(fd)
        print(f"[case] C5 adjoint vs fd, param {arr} {idx}: {ad:+.8e} {fd:+.8e} rel {err:.2e}", flush=True)
        assert err <= 1e-5


# ------------------------------------------------------------ a5, error paths
@pytest.mark.parametrize("case", ["fd2d", "fd3d"])
def test_fd_ghost_inflow_rhs(case):
    """the structured ghost inflow (src/structured_fd_commonk.hpp:330-343) with
    non-zero, windowed inflow on several faces and reflective sides, plus the
    source, against the oracle"""
    if case == "fd2d":
        prob = P.fd_box_2d_central_source(
            n=10, sn=6, bcs={4: {"type": "vacuum", "inflow": 3.0, "window": [0.2, 0.7]}, 2: "reflect",
                             1: {"type": "vacuum", "inflow": 1.5}, 3: {"type": "vacuum", "inflow": 0.5}})
    else:
        q = core.Quadrature(3, 6)
        bcs = {6: "reflect", 3: {"type": "vacuum", "inflow": 2.5, "window": [0.1, 0.6, 0.3, 0.9]},
               1: {"type": "vacuum", "inflow": 1.25}, 5: {"type": "vacuum", "inflow": 0.5, "window": [0.0, 0.5, 0.5, 1.0]}}
        d = core.Disc.structured(q, [6, 5, 4], [1.0, 1.3, 0.8], bcs)
        rng = np.random.default_rng(2)
        st = rng.uniform(0.5, 5.0, d.n_cells)
        prob = P.Problem("fd3d_inflow", q, d, st, 0.4 * st, rng.uniform(0.0, 2.0, d.n_cells))
    op = prob.operator()
    b, b32 = op.vec(), op.vec(core.FP32)
    op.rhs(0, b)
    op.rhs(0, b32)
    torch.cuda.synchronize()
    want = oracle_rhs(prob, 0, [])
    assert np.abs(want).max() > 0
    e64, e32 = rel_max(to_host(b), want), rel_max(to_host(b32), want)
    print(f"[case] {case} inflow rhs: fp64 {e64:.2e} fp32 {e32:.2e}", flush=True)
    assert e64 <= 1e-13 and e32 <= 1e-6


def test_singular_block_error():
    """an element block with a zero pivot is an error through the C ABI
    (src/block_inversek.kokkos.cxx:273-274): h = 1 and sigma_t = -(|mu| + |eta|)
    zero cell 0's S2 blocks exactly (every S2 ordinate shares |mu|, |eta|)"""
    q = core.Quadrature(2, 2)
    d = core.Disc.structured(q, [3, 3], [3.0, 3.0])
    st = np.full(9, 1.0)
    st[0] = -(abs(q.omega[0, 0]) + abs(q.omega[0, 1]))
    prob = P.Problem("singular", q, d, st, np.zeros(9), None)
    op = prob.operator()
    bi = u.BlockInverse(op)
    with pytest.raises(u.UbtError) as e:
        bi.setup()
    assert e.value.status == "ERR_MAT_LU_ZRPVT", e.value
    assert "4 of 36" in str(e.value)
    # a regular operator afterwards still sets up
    st[0] = 1.0
    op.set_xs(st, np.zeros(9))
    bi.setup()
