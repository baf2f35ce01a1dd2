"""Gradients of flux responses with respect to per-cell cross sections and
sources, by the adjoint transport solve.

The forward model is the multigroup downscatter solve the reference's driver
runs (tests/transportk.kokkos.cxx:1036-1156): groups in order, each

    A_g psi_g = inflow_g + q_g / W e_0 + sum_{g' < g} T_{g'g} psi_{g'},

A_g = L_g(sigma_t) - S_g(sigma_s(g -> g)), T_{g'g} psi = sigma_s(g' -> g) / W
times the scalar flux of psi (all on basis 0 for q, on every basis for T).
The output is the scalar flux phi_g (per node: cell x basis) of every group,
so any response J(phi) written in torch differentiates through it.

Backward: with c_g(node, a) = w_a dJ/dphi_g(node), the adjoint groups run in
reverse order,

    A_g^T lambda_g = c_g + sum_{g'' > g} w_a sigma_s(g -> g'') / W mu_{g''},

mu = sum_a lambda (unweighted), each solve with the adjoint operator kernel
(ubt_solver_solve_transpose). Then, per cell c,

    dJ/dsigma_t(g, c)      = - sum_{basis, a} lambda_g psi_g
    dJ/dsigma_s(g'->g, c)  = (1/W) sum_basis phi_{g'} mu_g      (g' <= g)
    dJ/dq(g, c)            = (1/W) mu_g(basis 0).

The reference has no adjoint; this wrapper is the sensitivity harness over the
library's C ABI. Upscatter (sigma_s(g' -> g), g' > g) is outside the forward
model and refused.
"""
import numpy as np
import torch

from . import core


class TransportModel:
    """Handles for repeated forward / adjoint multigroup solves on one discretisation."""

    def __init__(self, disc, n_groups, rtol=1e-12, max_it=2000, restart=30, mixed=False, dsa_inner_rtol=1e-6,
                 dsa_inner_max_it=200):
        self.disc, self.G = disc, n_groups
        self.quad = disc.quad
        self.op = core.Operator(disc, n_groups)
        self.dsa = None
        self._dsa_opts = dict(inner_rtol=dsa_inner_rtol, inner_max_it=dsa_inner_max_it)
        self._solver_opts = dict(rtol=rtol, max_it=max_it, restart=restart, mixed=mixed)
        self.solver = None
        self.w = torch.as_tensor(self.quad.w, dtype=torch.float64, device="cuda")
        self.W = float(self.quad.sum_weights)
        self.nb, self.A, self.nc = disc.nb, disc.n_angles, disc.n_cells
        self.last_results = []

    def _set(self, sigma_t, sigma_s, source):
        st = sigma_t.detach().double().cpu().numpy()
        ss = sigma_s.detach().double().cpu().numpy()
        q = source.detach().double().cpu().numpy()
        G = self.G
        if st.shape != (G, self.nc) or ss.shape != (G, G, self.nc) or q.shape != (G, self.nc):
            raise ValueError("expected sigma_t [G, cells], sigma_s [G, G, cells], source [G, cells] with G=%d, "
                             "cells=%d" % (G, self.nc))
        up = np.array([[gf > gt for gt in range(G)] for gf in range(G)])
        if np.any(ss[up] != 0.0):
            raise ValueError("upscatter sigma_s(g' -> g), g' > g, is outside the downscatter forward model")
        self.op.set_xs(st, ss, q)
        if self.dsa is None:
            self.dsa = core.DSA(self.op, **self._dsa_opts)
            self.solver = core.Solver(self.op, self.dsa, **self._solver_opts)
        return ss

    def _group(self, g):
        self.op.set_group(g)
        self.dsa.set_group(g)
        self.solver.refresh()

    def _check(self, res, what, g):
        self.last_results.append((what, g, res))
        if res["reason"] != 1:
            raise RuntimeError("%s solve of group %d did not converge: %r" % (what, g, res))

    def forward_solve(self, sigma_t, sigma_s, source):
        """psi [G, rows] and phi [G, nodes] (FP64, device)"""
        self._set(sigma_t, sigma_s, source)
        G = self.G
        # the library's group sweep: downscatter only, so one forward sweep
        psi, rep = self.solver.solve_groups()
        self.last_results = []
        for g in range(G):
            self._check({"iterations": rep["group_iterations"][g], "rel_residual": rep["group_rel_residual"][g],
                         "reason": rep["group_reason"][g]}, "forward", g)
        phi = (psi.view(G, -1, self.A) * self.w).sum(-1)
        return psi, phi

    def adjoint(self, psi, phi, grad_phi, sigma_s):
        """gradients (sigma_t, sigma_s, source) of J with dJ/dphi = grad_phi"""
        G, n, nb, A, nc, W = self.G, self.disc.n_rows, self.nb, self.A, self.nc, self.W
        ss = torch.as_tensor(sigma_s, dtype=torch.float64, device="cuda")
        lam = torch.zeros(n, dtype=torch.float64, device="cuda")
        mu = torch.zeros(G, nc * nb, dtype=torch.float64, device="cuda")
        g_st = torch.zeros(G, nc, dtype=torch.float64, device="cuda")
        g_ss = torch.zeros(G, G, nc, dtype=torch.float64, device="cuda")
        g_q = torch.zeros(G, nc, dtype=torch.float64, device="cuda")
        for g in reversed(range(G)):
            src = grad_phi[g].reshape(-1).clone()
            for gd in range(g + 1, G):  # transposed transfers g -> gd, from the solved adjoint groups
                src += (ss[g, gd].repeat_interleave(nb) / W) * mu[gd]
            c = (src[:, None] * self.w[None, :]).reshape(-1).contiguous()
            self._group(g)
            lam.zero_()
            if float(c.abs().max()) > 0.0:
                self._check(self.solver.solve_transpose(c, lam), "adjoint", g)
            lv = lam.view(nc, nb, A)
            mu[g] = lv.sum(-1).reshape(-1)
            g_st[g] = -(lv * psi[g].view(nc, nb, A)).sum((1, 2))
            mg = mu[g].view(nc, nb)
            for gf in range(g + 1):
                g_ss[gf, g] = (phi[gf].view(nc, nb) * mg).sum(1) / W
            g_q[g] = mg[:, 0] / W
        return g_st, g_ss, g_q


class _FluxSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sigma_t, sigma_s, source, model):
        psi, phi = model.forward_solve(sigma_t, sigma_s, source)
        ctx.model = model
        ctx.sigma_s = sigma_s.detach()
        ctx.like = [(x.device, x.dtype) for x in (sigma_t, sigma_s, source)]
        ctx.save_for_backward(psi, phi)
        return phi

    @staticmethod
    def backward(ctx, grad_phi):
        psi, phi = ctx.saved_tensors
        g = ctx.model.adjoint(psi, phi, grad_phi.double().contiguous(), ctx.sigma_s)
        # each gradient on the device and in the dtype of its input
        out = tuple(gi.to(device=dev, dtype=dt) for gi, (dev, dt) in zip(g, ctx.like))
        return out + (None,)


def scalar_flux(model, sigma_t, sigma_s, source):
    """phi [G, cells * n_basis] of the multigroup solve, differentiable in
    sigma_t [G, cells], sigma_s [G_from, G_to, cells] and source [G, cells]
    through the adjoint solves."""
    return _FluxSolve.apply(sigma_t, sigma_s, source, model)


def box_response_weights(disc, lo, hi):
    """weights r[cell] so that J = sum_g sum_c r_c phi_g(c, basis 0) is the
    volume integral of the group-summed scalar flux over the cells whose
    centroid lies in the box [lo, hi] (the basis is orthonormal in the cell
    mean, so basis 0 carries the cell average)"""
    cen = disc.get("centroid").reshape(-1, 3)[:, :disc.dim] if disc.kind == "dg" else _fd_centres(disc)
    lo, hi = np.asarray(lo)[:disc.dim], np.asarray(hi)[:disc.dim]
    inside = np.all((cen >= lo) & (cen <= hi), axis=1)
    vol = disc.get("volume") if disc.kind == "dg" else np.full(disc.n_cells, float(np.prod(disc.get("h")[:disc.dim])))
    return inside * vol


def _fd_centres(disc):
    n, h = disc.get("n"), disc.get("h")
    idx = np.indices([n[2], n[1], n[0]])[::-1].reshape(3, -1).T
    return ((idx + 0.5) * h)[:, :disc.dim]
