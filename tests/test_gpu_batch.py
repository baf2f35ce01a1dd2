"""Batched problems on the GPU: one persistent apply, the batched block
inverse and merged sweep, and the batched FGMRES (device-side per-problem
Hessenberg, graph-replayed cycles) against the oracle, problem by problem."""
import numpy as np
import pytest
import scipy.sparse.linalg as spl
import torch

import oracle
import ubolt_amd as u
from helpers import oracle_full, oracle_mult, oracle_rhs, oracle_rows, rel_max, to_dev, to_host
from ubolt_amd import problems as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ensemble():
    probs = P.batch_quads(n_problems=6, seed=5, side=(3, 9), sns=(2, 4, 6), groups=(1, 3))
    ops = [p.operator() for p in probs]
    b = u.Batch(ops)
    print(f"[case] batch of {len(probs)}: rows {b.n_rows}, tuning {b.tuning()}", flush=True)
    return probs, ops, b


@pytest.mark.parametrize("body", [0, 1, 2, 3])
def test_batch_mult(gpu, ensemble, body):
    """every apply body (tile, four-angle tile, box tile, strip march) against the oracle"""
    probs, ops, b = ensemble
    b.set_group(0)
    t = b.tuning()
    assert t["us"][body] > 0 and t["us"][4 + body] > 0, f"body {body} does not apply: {t}"
    x = np.random.default_rng(1).standard_normal(b.n_rows)
    for dt, tol in ((torch.float64, 1e-12), (torch.float32, 2e-5)):
        prec = u.FP64 if dt == torch.float64 else u.FP32
        b.set_body(prec, body)
        y = b.vec(prec)
        b.mult(to_dev(x, dt), y)
        torch.cuda.synchronize()
        yh = to_host(y)
        for p, prob in enumerate(probs):
            sl = slice(b.offsets[p], b.offsets[p + 1])
            err = rel_max(yh[sl], oracle_mult(prob, x[sl]))
            assert err <= tol, (p, dt, err)
        b.set_body(prec, t["body_fp64" if prec == u.FP64 else "body_fp32"])


def test_batch_inactive_untouched(gpu, ensemble):
    """group 2 exists only in the problems with three groups; the others' spans
    keep what was there"""
    probs, ops, b = ensemble
    live = [p.n_groups > 2 for p in probs]
    if not any(live):
        pytest.skip("no problem with three groups in this draw")
    b.set_group(2)
    x = b.vec()
    x.normal_()
    y = b.vec()
    y.fill_(7.0)
    b.mult(x, y)
    torch.cuda.synchronize()
    yh, xh = to_host(y), to_host(x)
    for p, prob in enumerate(probs):
        sl = slice(b.offsets[p], b.offsets[p + 1])
        if live[p]:
            assert rel_max(yh[sl], oracle_mult(prob, xh[sl], g=2)) <= 1e-12
        else:
            assert np.all(yh[sl] == 7.0)
    b.set_group(0)


def test_batch_blocks_and_sweep(gpu, ensemble):
    probs, ops, b = ensemble
    b.set_group(0)
    x = np.random.default_rng(2).standard_normal(b.n_rows)
    yb, ys = b.vec(), b.vec()
    b.blocks_apply(to_dev(x), yb)
    b.sweep_apply(to_dev(x), ys)
    torch.cuda.synchronize()
    for p, (prob, op) in enumerate(zip(probs, ops)):
        sl = slice(b.offsets[p], b.offsets[p + 1])
        bi = u.BlockInverse(op)
        bi.setup()
        want = op.vec()
        bi.apply(to_dev(x[sl]), want)
        torch.cuda.synchronize()
        assert rel_max(to_host(yb)[sl], to_host(want)) <= 1e-14
        ref = spl.spsolve(oracle_rows(prob).to_scipy().tocsc(), x[sl])
        assert rel_max(to_host(ys)[sl], ref) <= 1e-11


@pytest.mark.parametrize("mixed,gs_passes", [(False, 2), (True, 2), (False, 1), (True, 1)])
def test_batch_solve(gpu, ensemble, mixed, gs_passes):
    """FP64 and mixed, classical Gram-Schmidt twice (default) and once"""
    probs, ops, b = ensemble
    b.set_group(0)
    sol = u.BatchSolver(b, rtol=1e-8 if mixed else 1e-11, max_it=600, restart=20, mixed=mixed, gs_passes=gs_passes)
    sol.refresh()
    rhs = b.vec()
    b.rhs(rhs)
    x = b.vec()
    res = sol.solve(rhs, x)
    torch.cuda.synchronize()
    print(f"batch solve {'mixed' if mixed else 'fp64'}: {res}")
    bh, xh = to_host(rhs), to_host(x)
    for p, prob in enumerate(probs):
        sl = slice(b.offsets[p], b.offsets[p + 1])
        A = oracle_full(prob)
        true_res = np.linalg.norm(bh[sl] - A @ xh[sl]) / np.linalg.norm(bh[sl])
        assert res["reason"][p] == 1, (p, res)
        assert true_res <= (1.05e-8 if mixed else 1e-10), (p, true_res)
        assert abs(true_res - res["rel_residual"][p]) <= 1e-6 * max(true_res, 1e-16) + 1e-14
        if not mixed:
            want = spl.spsolve(A, bh[sl])
            phi, phw = oracle.moments(xh[sl], prob.quad.w), oracle.moments(want, prob.quad.w)
            assert np.linalg.norm(phi - phw) / np.linalg.norm(phw) <= 1e-8


def test_batch_multigroup(gpu, ensemble):
    """group Gauss-Seidel over the batch: rhs with transfers from the solved
    groups of each problem, one batched solve per group"""
    probs, ops, b = ensemble
    G = max(p.n_groups for p in probs)
    caches = [[] for _ in probs]
    for g in range(G):
        b.set_group(g)
        sol = u.BatchSolver(b, rtol=1e-11, max_it=600, restart=20)
        sol.refresh()
        rhs = b.vec()
        b.rhs(rhs)
        x = b.vec()
        res = sol.solve(rhs, x)
        b.set_scalar_flux(x)
        torch.cuda.synchronize()
        xh = to_host(x)
        for p, prob in enumerate(probs):
            if prob.n_groups <= g:
                assert res["reason"][p] == 0
                continue
            sl = slice(b.offsets[p], b.offsets[p + 1])
            want = spl.spsolve(oracle_full(prob, g), oracle_rhs(prob, g, caches[p]))
            caches[p].append(oracle.moments(want, prob.quad.w))
            phi = oracle.moments(xh[sl], prob.quad.w)
            assert np.linalg.norm(phi - caches[p][-1]) / np.linalg.norm(caches[p][-1]) <= 1e-8, (p, g)
    b.set_group(0)


def test_batch_follows_set_xs(gpu):
    """new cross sections on a member after the batch and its solver were
    built (and the solver's cycles captured): the batched apply, sweep and the
    graph-replayed solve use them, and agree with the solve without graphs"""
    probs = P.batch_quads(n_problems=3, seed=7, side=(3, 6), sns=(2, 4), groups=(1, 1))
    ops = [p.operator() for p in probs]
    b = u.Batch(ops)
    b.set_group(0)
    sol = u.BatchSolver(b, rtol=1e-11, max_it=600, restart=20)
    sol.refresh()
    rhs = b.vec()
    b.rhs(rhs)
    x = b.vec()
    sol.solve(rhs, x)  # captures the cycles on the old cross sections
    rng = np.random.default_rng(21)
    for k in (0, 2):
        pr = probs[k]
        pr.sigma_t = np.asarray(pr.sigma_t) * rng.uniform(1.5, 3.0, np.asarray(pr.sigma_t).shape)
        pr.sigma_s = np.asarray(pr.sigma_s) * rng.uniform(0.2, 0.9, np.asarray(pr.sigma_s).shape)
        ops[k].set_xs(pr.sigma_t, pr.sigma_s, pr.source)
    xr = rng.standard_normal(b.n_rows)
    y, ys = b.vec(), b.vec()
    b.mult(to_dev(xr), y)
    b.sweep_apply(to_dev(xr), ys)
    rhs2 = b.vec()
    b.rhs(rhs2)
    xg, xn = b.vec(), b.vec()
    res_g = sol.solve(rhs2, xg)
    sol.set_graph(False)
    res_n = sol.solve(rhs2, xn)
    torch.cuda.synchronize()
    yh, ysh, bh = to_host(y), to_host(ys), to_host(rhs2)
    print(f"[case] set_xs on batch members 0, 2: graph {res_g['iterations']} plain {res_n['iterations']}", flush=True)
    for p, prob in enumerate(probs):
        sl = slice(b.offsets[p], b.offsets[p + 1])
        assert rel_max(yh[sl], oracle_mult(prob, xr[sl])) <= 1e-12, p
        assert rel_max(ysh[sl], spl.spsolve(oracle_rows(prob).to_scipy().tocsc(), xr[sl])) <= 1e-11, p
        want = oracle_full(prob) @ to_host(xg)[sl]
        assert np.linalg.norm(bh[sl] - want) / np.linalg.norm(bh[sl]) <= 1e-10, p
        assert res_g["iterations"][p] == res_n["iterations"][p], (p, res_g, res_n)
        assert rel_max(to_host(xg)[sl], to_host(xn)[sl]) <= 1e-12, p


def test_batch_strip_bands(gpu):
    """the strip march and the box tile over problems of several strips and
    bands (sides 20..140, S4-S12, all vacuum: no list phase), both precisions,
    problem by problem against the oracle"""
    probs = P.batch_quads(n_problems=4, seed=9, side=(20, 140), sns=(4, 8, 12), groups=(1, 1))
    ops = [p.operator() for p in probs]
    b = u.Batch(ops)
    b.set_group(0)
    x = np.random.default_rng(3).standard_normal(b.n_rows)
    for body in (2, 3):
        for dt, tol in ((torch.float64, 1e-12), (torch.float32, 2e-5)):
            prec = u.FP64 if dt == torch.float64 else u.FP32
            b.set_body(prec, body)
            y = b.vec(prec)
            b.mult(to_dev(x, dt), y)
            torch.cuda.synchronize()
            yh = to_host(y)
            for p, prob in enumerate(probs):
                sl = slice(b.offsets[p], b.offsets[p + 1])
                err = rel_max(yh[sl], oracle_mult(prob, x[sl]))
                print(f"[case] body {body} {dt} problem {p} ({prob.name}): {err:.2e}", flush=True)
                assert err <= tol, (body, p, dt, err)


def test_batch_box_sweep(gpu):
    """the record-free box sweep (every problem a uniform all-vacuum DG1 quad
    box) against the oracle's direct solve of L, problem by problem, and
    against the record sweep of the same batch, both precisions"""
    probs = P.batch_quads(n_problems=5, seed=13, side=(5, 90), sns=(2, 4, 8, 12), groups=(1, 1))
    ops = [p.operator() for p in probs]
    b = u.Batch(ops)
    b.set_group(0)
    assert b.sweep_info() == {"box_fp64": True, "box_fp32": True, "exact_fp64": True, "exact_fp32": True}
    x = np.random.default_rng(17).standard_normal(b.n_rows)
    out = {}
    for box in (True, False):
        b.set_sweep(box)
        assert b.sweep_info()["box_fp64"] == box
        for dt in (torch.float64, torch.float32):
            y = b.vec(u.FP64 if dt == torch.float64 else u.FP32)
            b.sweep_apply(to_dev(x, dt), y)
            torch.cuda.synchronize()
            out[(box, dt)] = to_host(y)
    b.set_sweep(True)
    for p, prob in enumerate(probs):
        sl = slice(b.offsets[p], b.offsets[p + 1])
        want = spl.spsolve(oracle_rows(prob).to_scipy().tocsc(), x[sl])
        e64 = rel_max(out[(True, torch.float64)][sl], want)
        e32 = rel_max(out[(True, torch.float32)][sl], want)
        print(f"[case] box sweep problem {p} ({prob.name}): fp64 {e64:.2e} fp32 {e32:.2e}", flush=True)
        assert e64 <= 1e-11 and e32 <= 1e-4, (p, e64, e32)
        assert rel_max(out[(True, torch.float64)][sl], out[(False, torch.float64)][sl]) <= 1e-12


def test_batch_pc_folds_to_exact_sweep(gpu):
    """On a batch whose sweep inverts every problem's L exactly the batch
    solver's preconditioner is the sweep alone: the composite it stands for -
    the batched element blocks, then the sweep on the L-residual - formed here
    from the batch's own stages, with the box sweep and with the record sweep,
    against the sweep of r; a batch holding a problem with reflective faces
    is reported inexact (it keeps both stages)"""
    probs = P.batch_quads(n_problems=4, seed=23, side=(4, 40), sns=(2, 4, 8), groups=(1, 1))
    ops = [p.operator() for p in probs]
    b = u.Batch(ops)
    b.set_group(0)
    x = np.random.default_rng(29).standard_normal(b.n_rows)
    for box in (True, False):
        b.set_sweep(box)
        info = b.sweep_info()
        assert info["exact_fp64"] and info["exact_fp32"] and info["box_fp64"] == box
        for dt, tol in ((torch.float64, 1e-11), (torch.float32, 1e-4)):
            prec = u.FP64 if dt == torch.float64 else u.FP32
            r = to_dev(x, dt)
            y = b.blocks_apply(r, b.vec(prec))
            t = b.vec(prec)
            for p, op in enumerate(ops):
                op.mult_unscattered(b.part(y, p), b.part(t, p))
            comp = y + b.sweep_apply(r - t, b.vec(prec))
            sr = b.sweep_apply(r, b.vec(prec))
            torch.cuda.synchronize()
            err = rel_max(to_host(sr), to_host(comp))
            print(f"[case] box sweep {box} {dt}: S r vs composite {err:.2e}", flush=True)
            assert err <= tol, (box, dt, err)
    b.set_sweep(True)
    refl = P.dg_box_2d(5, 6, 4, order=1, bcs={1: "reflect", 2: "vacuum", 3: "vacuum", 4: "vacuum"})
    b2 = u.Batch(ops[:2] + [refl.operator()])
    b2.set_group(0)
    info = b2.sweep_info()
    assert not info["exact_fp64"] and not info["exact_fp32"] and not info["box_fp64"]


@pytest.mark.parametrize("mixed", [False, True])
def test_batch_solver_quadrant_order(gpu, mixed):
    """On a batch of uniform all-vacuum boxes the batch solver keeps its Krylov
    vectors in the box sweep's quadrant-major order (the box apply reads and
    writes it, the sweep needs no reorder); with the record sweep it runs in
    the reference order. Both solves against the oracle's matrices and each
    other, iteration for iteration"""
    probs = P.batch_quads(n_problems=4, seed=31, side=(6, 40), sns=(4, 8, 12), groups=(1, 1))
    ops = [p.operator() for p in probs]
    b = u.Batch(ops)
    b.set_group(0)
    rhs = b.vec()
    b.rhs(rhs)
    out = {}
    for box in (True, False):
        b.set_sweep(box)
        sol = u.BatchSolver(b, rtol=1e-8 if mixed else 1e-11, max_it=600, restart=20, mixed=mixed, gs_passes=1)
        sol.refresh()
        x = b.vec()
        res = sol.solve(rhs, x)
        torch.cuda.synchronize()
        out[box] = (to_host(x), res)
        print(f"[case] box sweep {box} mixed {mixed}: iterations {list(res['iterations'])}", flush=True)
    b.set_sweep(True)
    bh = to_host(rhs)
    (xq, rq), (xn, rn) = out[True], out[False]
    assert list(rq["iterations"]) == list(rn["iterations"])
    for p, prob in enumerate(probs):
        sl = slice(b.offsets[p], b.offsets[p + 1])
        A = oracle_full(prob)
        tq = np.linalg.norm(bh[sl] - A @ xq[sl]) / np.linalg.norm(bh[sl])
        assert rq["reason"][p] == 1 and tq <= (1.05e-8 if mixed else 1e-10), (p, tq)
        assert rel_max(xq[sl], xn[sl]) <= (1e-6 if mixed else 1e-9), p


def test_batch_singular_block_names_problem(gpu):
    """the batch's singular-block check (one counter per problem, read back
    together) names the problem: on unit DG0 quads an S2 ordinate's outflow
    sum is |mu| + |eta|, so sigma_t = -(|mu| + |eta|) zeroes cell 0's blocks"""
    q = u.Quadrature(2, 2)
    coords, cells = P.box_quads(3, 3, 3.0, 3.0)
    ops = []
    for k in range(2):
        d = u.Disc.dg(q, u.Mesh(coords, cells), 0)
        st = np.full(9, 1.0)
        if k == 1:
            st[0] = -(abs(q.omega[0, 0]) + abs(q.omega[0, 1]))
        ops.append(P.Problem(f"p{k}", q, d, st, np.zeros(9), None).operator())
    b = u.Batch(ops)
    bs = u.BatchSolver(b, rtol=1e-8)
    with pytest.raises(u.UbtError) as e:
        bs.refresh()
    assert e.value.status == "ERR_MAT_LU_ZRPVT", e.value
    assert "problem 1" in str(e.value) and "of 36" in str(e.value), str(e.value)
