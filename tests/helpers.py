"""Test helpers: hand a product discretisation's host tables to the oracle and
compare device results with it."""
import numpy as np
import torch

import oracle


def xs_arrays(prob):
    G, nc = prob.n_groups, prob.disc.n_cells
    st = np.broadcast_to(np.asarray(prob.sigma_t, dtype=np.float64).reshape(G, -1), (G, nc))
    ss = np.broadcast_to(np.asarray(prob.sigma_s, dtype=np.float64).reshape(G, G, -1), (G, G, nc))
    q = None
    if prob.source is not None:
        q = np.broadcast_to(np.asarray(prob.source, dtype=np.float64).reshape(G, -1), (G, nc))
    return st, ss, q


def oracle_rows(prob, g=0, sigt=None):
    d, q = prob.disc, prob.quad
    st, _, _ = xs_arrays(prob)
    sig = st[g] if sigt is None else sigt
    if d.kind == "fd":
        bcf = d.get("bc_face").reshape(6, 7)
        return oracle.fd_assemble(d.dim, d.get("n"), d.get("h"), q.omega, q.reflect, bcf[:, 0].astype(int), sig)
    return oracle.dg_assemble(d.nb, d.nf, q.omega, q.reflect, d.get("face_nb"), d.get("face_nA"), d.get("volume"),
                              d.get("face_own"), d.get("face_up"), d.get("basis_grad"), sig)


def oracle_mult(prob, x, g=0, scatter=True, rows=None):
    R = rows if rows is not None else oracle_rows(prob, g)
    _, ss, _ = xs_arrays(prob)
    return oracle.shell_mult(R, prob.disc.nb, prob.quad.w, prob.quad.sum_weights, ss[g, g], x, scatter)


def oracle_ghost_inflow(prob):
    d, q = prob.disc, prob.quad
    b = np.zeros(d.n_rows)
    if d.kind == "fd":
        return oracle.fd_ghost_inflow(d.dim, d.get("n"), d.get("h"), q.omega, d.get("bc_face"), b)
    return oracle.dg_ghost_inflow(d.nb, d.nf, q.omega, d.get("face_nb"), d.get("face_nA"), d.get("volume"),
                                  d.get("face_basis_value"), d.get("face_inflow"), b)


def oracle_rhs(prob, g, phi_cache):
    """ghost inflow + source/W on basis 0 + sum_{g' < g} sigma_s(g'->g)/W phi_g'"""
    d, q = prob.disc, prob.quad
    st, ss, src = xs_arrays(prob)
    b = oracle_ghost_inflow(prob)
    W = q.sum_weights
    if src is not None:
        ind = np.zeros(d.n_nodes)
        ind[::d.nb] = 1.0
        oracle.broadcast_add(d.nb, d.n_angles, src[g] / W, ind, b)
    for gf in range(g):
        oracle.broadcast_add(d.nb, d.n_angles, ss[gf, g] / W, phi_cache[gf], b)
    return b


def rel_max(a, b):
    """max_r |a - b| / max_r |b| - the reference's -check_matfree measure"""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def to_dev(x, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device="cuda")


def to_host(t):
    return t.detach().double().cpu().numpy()


def oracle_dsa_matrix(prob, g=0, cons=True, p=1.5, C=4.0, **void):
    """the reference DSA diffusion matrix of group g (oracle/dsa.py) and the
    restriction weight per node (zero on masked void nodes); void: the void
    policy's options (bridge, void_sigma_t, void_d)"""
    from oracle import dsa as odsa
    d, q = prob.disc, prob.quad
    st, ss, _ = xs_arrays(prob)
    if d.kind == "fd":
        ft = d.get("bc_face").reshape(3, 2, 7)[:, :, 0].astype(int)
        M = odsa.structured(d.dim, d.get("n"), d.get("h"), ft, q.omega, q.w, q.sum_weights, st[g], ss[g, g], cons, p,
                            **void)
        wgt = np.ones(d.n_nodes)
    else:
        args = (d.nf, d.get("face_nb"), d.get("face_nA"), d.get("face_area"), d.get("face_distance"), d.get("volume"))
        if d.nb == 1:
            M = odsa.dg0(*args, q.omega, q.w, q.sum_weights, st[g], ss[g, g], cons, p, **void)
        else:
            M = odsa.dg1(d.nb, *args, d.get("face_own"), d.get("face_up"), d.get("basis_grad"), st[g], ss[g, g], C,
                         **void)
        wgt = np.repeat(d.get("volume"), d.nb)
    return M, wgt * np.repeat(oracle_dsa_unmasked(prob, g, **void), d.nb)


def oracle_dsa_unmasked(prob, g=0, **void):
    """1 per cell, 0 on the cells the void policy masks"""
    from oracle import dsa as odsa
    d = prob.disc
    st, ss, _ = xs_arrays(prob)
    if d.kind == "fd":
        ft = d.get("bc_face").reshape(3, 2, 7)[:, :, 0].astype(int)
        vs = lambda v: odsa.structured_void_vs(d.dim, d.get("n"), d.get("h"), ft, v)
        any_vac = bool(np.any(ft[:d.dim] == 0))
    else:
        fnb = d.get("face_nb")
        vs = lambda v: odsa.plex_void_vs(d.nf, fnb, d.get("face_area"), d.get("volume"), v)
        any_vac = bool(np.any(fnb == -1))
    _, vd, _, masked = odsa.diffusion_coefficients(st[g], ss[g, g], vs, any_vac, **void)
    return np.where(vd & masked, 0.0, 1.0)


def oracle_dsa_apply(prob, x, g=0, **void):
    from oracle import dsa as odsa
    M, wgt = oracle_dsa_matrix(prob, g, **void)
    return odsa.apply(M, wgt, prob.quad.n_angles, prob.quad.w, prob.quad.sum_weights, x)


def oracle_full(prob, g=0):
    """the whole group operator A = streaming + removal - scatter as a sparse matrix"""
    import scipy.sparse as sp
    d, q = prob.disc, prob.quad
    _, ss, _ = xs_arrays(prob)
    R = oracle_rows(prob, g).to_scipy()
    A = d.n_angles
    S = sp.kron(sp.diags(np.repeat(ss[g, g], d.nb) / q.sum_weights), sp.csr_matrix(np.ones((A, 1)) @ q.w[None, :]))
    return (R - S).tocsc()


def oracle_solve(prob, b, g=0):
    import scipy.sparse.linalg as spl
    return spl.spsolve(oracle_full(prob, g), b)
