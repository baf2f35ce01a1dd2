"""GPU parity of the adjoint path: every transposed apply against the oracle's
assembled matrices transposed (FP64 <= 1e-12 relative max-norm, FP32 bounded),
the dot-product identity <A x, y> = <x, A^T y> at the full benchmark size, and
the adjoint solve against a direct solve with the transposed oracle matrix.

The identity the adjoint kernel is built on, A^T = V P A P V^{-1} on the
streaming + removal rows, is pinned on the oracle itself in
test_oracle.test_transpose_symmetry."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl
import torch

import ubolt_amd as u
from helpers import oracle_dsa_matrix, oracle_full, oracle_rows, rel_max, to_dev, to_host
from test_gpu_operator import CASES, FP32_TOL, FP64_TOL
from ubolt_amd import problems as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=sorted(CASES))
def built(request):
    prob = CASES[request.param]()
    op = prob.operator()
    print(f"[case] {request.param}: rows={prob.disc.n_rows}", flush=True)
    return request.param, prob, op


def test_mult_transpose(gpu, built):
    name, prob, op = built
    x = np.random.default_rng(31).standard_normal(prob.disc.n_rows)
    want = oracle_full(prob).T @ x
    want_l = oracle_rows(prob).to_scipy().T @ x
    y, yl = op.vec(), op.vec()
    op.mult_transpose(to_dev(x), y)
    op.mult_unscattered_transpose(to_dev(x), yl)
    y32 = op.vec(u.FP32)
    op.mult_transpose(to_dev(x, torch.float32), y32)
    torch.cuda.synchronize()
    e64, el, e32 = rel_max(to_host(y), want), rel_max(to_host(yl), want_l), rel_max(to_host(y32), want)
    print(f"{name}: A^T fp64 {e64:.2e}, L^T fp64 {el:.2e}, A^T fp32 {e32:.2e}")
    assert e64 <= FP64_TOL and el <= FP64_TOL and e32 <= FP32_TOL


def diagonal_blocks(prob):
    """the same-cell, same-angle entries of the streaming + removal rows"""
    d = prob.disc
    R = oracle_rows(prob).to_scipy().tocoo()
    per_cell = d.nb * d.n_angles
    keep = (R.row // per_cell == R.col // per_cell) & (R.row % d.n_angles == R.col % d.n_angles)
    return sp.csr_matrix((R.data[keep], (R.row[keep], R.col[keep])), shape=R.shape)


def test_blockinv_transpose(gpu, built):
    name, prob, op = built
    bi = u.BlockInverse(op)
    bi.setup()
    x = np.random.default_rng(32).standard_normal(prob.disc.n_rows)
    y = op.vec()
    bi.apply_transpose(to_dev(x), y)
    torch.cuda.synchronize()
    Dt = diagonal_blocks(prob).T.tocsr()
    err = rel_max(Dt @ to_host(y), x)
    print(f"{name}: D^T (D^-T x) vs x {err:.2e}")
    assert err <= 1e-12


def has_reflective_face(d):
    if d.kind == "fd":
        return bool(np.any(d.get("bc_face").reshape(6, 7)[:, 0] == 1))
    return bool(np.any(d.get("face_nb") <= -2))


def test_sweep_transpose(gpu, built):
    name, prob, op = built
    if has_reflective_face(prob.disc):
        pytest.skip("the sweep lags reflective couplings; it is exact only without them")
    sw = u.Sweep(op)
    r = np.random.default_rng(33).standard_normal(prob.disc.n_rows)
    want = spl.spsolve(oracle_rows(prob).to_scipy().T.tocsc(), r)
    psi, psi32 = op.vec(), op.vec(u.FP32)
    sw.apply_transpose(to_dev(r), psi)
    sw.apply_transpose(to_dev(r, torch.float32), psi32)
    torch.cuda.synchronize()
    e64, e32 = rel_max(to_host(psi), want), rel_max(to_host(psi32), want)
    print(f"{name}: adjoint sweep fp64 {e64:.2e}, fp32 {e32:.2e}")
    assert e64 <= 1e-11 and e32 <= 1e-4


def test_dsa_transpose(gpu, built):
    """(P D^{-1} W R)^T x = w_a W_node D^{-1} (sum_a x / sum_w), D symmetric"""
    name, prob, op = built
    dsa = u.DSA(op, inner_max_it=3000, inner_rtol=1e-14)
    dsa.set_group(0)
    M, wgt = oracle_dsa_matrix(prob)
    q = prob.quad
    x = np.random.default_rng(34).standard_normal(prob.disc.n_rows)
    s = spl.spsolve(M.tocsc(), x.reshape(-1, q.n_angles).sum(1) / q.sum_weights)
    want = ((wgt * s)[:, None] * q.w[None, :]).ravel()
    y = op.vec()
    dsa.apply_transpose(to_dev(x), y)
    torch.cuda.synchronize()
    err = rel_max(to_host(y), want)
    print(f"{name}: DSA^T {err:.2e}")
    assert err <= 1e-9


@pytest.mark.parametrize("mixed", [False, True])
def test_solve_transpose(gpu, built, mixed):
    name, prob, op = built
    dsa = u.DSA(op)
    dsa.set_group(0)
    sol = u.Solver(op, dsa, rtol=1e-8 if mixed else 1e-12, max_it=400, mixed=mixed, inner_rtol=1e-4)
    sol.refresh()
    b = np.random.default_rng(35).uniform(0.0, 1.0, prob.disc.n_rows)
    x = op.vec()
    res = sol.solve_transpose(to_dev(b), x)
    torch.cuda.synchronize()
    At = oracle_full(prob).T.tocsc()
    got = to_host(x)
    true_res = np.linalg.norm(b - At @ got) / np.linalg.norm(b)
    want = spl.spsolve(At, b)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"{name}: adjoint solve {'mixed' if mixed else 'fp64'} its {res['iterations']} rel res "
          f"{res['rel_residual']:.2e} (oracle {true_res:.2e}), rel L2 vs direct {err:.2e}")
    assert res["reason"] == 1
    assert true_res <= (1.05e-8 if mixed else 2e-12)
    if not mixed:
        assert err <= 1e-8


@pytest.mark.parametrize("which", ["dg1_delaunay_20k_s8", "fd3d_64_s8"])
def test_dot_product_identity_full_size(gpu, which):
    """<A x, y> = <x, A^T y> on the benchmark-size problems (no oracle needed)"""
    if which.startswith("dg1"):
        prob = P.dg1_delaunay(n_side=100, sn=8, seed=2)
    else:
        prob = P.fd_cube_multigroup(n=64, sn=8, G=2)
    op = prob.operator()
    g = torch.Generator(device="cuda").manual_seed(5)
    x = torch.randn(prob.disc.n_rows, dtype=torch.float64, device="cuda", generator=g)
    y = torch.randn(prob.disc.n_rows, dtype=torch.float64, device="cuda", generator=g)
    ax, aty = op.vec(), op.vec()
    op.mult(x, ax)
    op.mult_transpose(y, aty)
    lhs, rhs = float(torch.dot(ax, y)), float(torch.dot(x, aty))
    scale = float(ax.norm() * y.norm())
    print(f"{which}: <Ax,y> {lhs:.15e} <x,A^T y> {rhs:.15e} rel {abs(lhs - rhs) / scale:.2e}")
    assert abs(lhs - rhs) <= 1e-13 * scale
