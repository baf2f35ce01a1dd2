"""The oracle's group sweep - TEST INFRASTRUCTURE ONLY.

A restatement of the reference test driver's multigroup loop
(tests/transportk.kokkos.cxx:925-1208) with exact inner solves: the
Gauss-Seidel sweep over the groups (a group's rhs takes every transfer out of
the cached scalar flux of the groups above it, and a thermal group g >= g_up
every other group's: zero before that group is first solved, :964-976,
:1058-1065), the scalar flux cached after each solve with the thermal block's
change summed (src/multigroupk.kokkos.cxx:266-299), and the outer test on the
error estimate change * rho / (1 - rho) (:1159-1207), stopping only on a
sweep whose thermal solves ran at the inner rtol. With an inexact inner solve
the thermal solves after the first sweep start from the group's last psi at
the adaptive rtol max(inner rtol, gamma * min(change, 1)) (:1108-1121); with
exact (LU) inner solves there is nothing to loosen, the gamma = 0 iteration.
"""
import math

import numpy as np
import scipy.sparse.linalg as spl


def g_up_of(sigma_s):
    """the first group anything upscatters into: the smallest g_to of a
    nonzero sigma_s[g_from][g_to], g_from > g_to (reference :953-962);
    sigma_s is [G][G][cells]"""
    G = sigma_s.shape[0]
    g_up = G
    for gf in range(G):
        for gt in range(gf):
            if np.any(sigma_s[gf, gt] != 0.0):
                g_up = min(g_up, gt)
    return g_up


def gmres_right(A, precond, b, x0, rtol, restart=30, max_it=2000):
    """x with ||b - A x|| <= rtol ||b|| (the Krylov stopping test with a
    nonzero guess, on the unpreconditioned residual) by right-preconditioned
    GMRES(restart) from x0 (None: zero): classical Gram-Schmidt, Givens
    rotations, the residual estimate |g_{j+1}| checked every step"""
    x = np.zeros_like(b) if x0 is None else x0.copy()
    nb = np.linalg.norm(b)
    its = 0
    while its < max_it:
        r = b - A @ x
        beta = np.linalg.norm(r)
        if beta <= rtol * nb:
            return x
        V = np.zeros((restart + 1, b.size))
        Z = np.zeros((restart, b.size))
        H = np.zeros((restart + 1, restart))
        cs, sn = np.zeros(restart), np.zeros(restart)
        g = np.zeros(restart + 1)
        V[0], g[0] = r / beta, beta
        k = 0
        for j in range(restart):
            Z[j] = precond(V[j])
            w = A @ Z[j]
            H[:j + 1, j] = V[:j + 1] @ w
            w = w - H[:j + 1, j] @ V[:j + 1]
            H[j + 1, j] = np.linalg.norm(w)
            if H[j + 1, j] > 0.0:
                V[j + 1] = w / H[j + 1, j]
            for i in range(j):
                h0, h1 = H[i, j], H[i + 1, j]
                H[i, j], H[i + 1, j] = cs[i] * h0 + sn[i] * h1, -sn[i] * h0 + cs[i] * h1
            den = np.hypot(H[j, j], H[j + 1, j])
            cs[j], sn[j] = H[j, j] / den, H[j + 1, j] / den
            H[j, j], H[j + 1, j] = den, 0.0
            g[j + 1], g[j] = -sn[j] * g[j], cs[j] * g[j]
            its += 1
            k = j + 1
            if abs(g[j + 1]) <= rtol * nb or its >= max_it:
                break
        y = np.linalg.solve(np.triu(H[:k, :k]), g[:k])
        x += y @ Z[:k]
    return x


def group_sweep(A, b0, transfer, moments, g_up, outer_rtol=1e-8, max_outer=1000, trace=None, inner=None,
                inner_gamma=0.0, inner_rtol=1e-10):
    """A[g]: group g's operator (scipy sparse); b0[g]: its inflow and source;
    transfer(gf, gt, phi): the rhs contribution of group gf's scalar flux to
    group gt; moments(psi): the scalar flux; inner(g, b, x0, rtol): an inexact
    solve of A[g] x = b from x0 (None: exact LU solves). Returns (psi, phi,
    sweeps, status) with status 1 converged, -2 out of sweeps; a trace list
    receives (change, rho, estimate, the sweep's thermal rtol) per sweep."""
    G = len(A)
    n_nodes = len(moments(np.zeros(A[0].shape[0])))
    phi = [np.zeros(n_nodes) for _ in range(G)]
    psi = [None] * G
    lu = [spl.splu(a.tocsc()) for a in A] if inner is None else None
    outer_change = 0.0
    upscatter = g_up < G
    final_sweep = False
    for outer in range(max_outer):
        if upscatter and inner_gamma > 0.0:
            sweep_rtol = inner_rtol if final_sweep else \
                max(inner_rtol, inner_gamma * (1.0 if outer == 0 else min(outer_change, 1.0)))
        else:
            sweep_rtol = inner_rtol
        change_sq = norm_sq = 0.0
        for g in (range(G) if outer == 0 else range(g_up, G)):
            b = b0[g].copy()
            for gf in range(G if g >= g_up else g):
                if gf != g:
                    b += transfer(gf, g, phi[gf])
            if inner is None:
                psi[g] = lu[g].solve(b)
            else:
                psi[g] = inner(g, b, psi[g] if outer >= 1 else None, sweep_rtol if g >= g_up else inner_rtol)
            new = moments(psi[g])
            if g >= g_up:
                change_sq += float(np.sum((new - phi[g]) ** 2))
                norm_sq += float(np.sum(new ** 2))
            phi[g] = new
        sweeps = outer + 1
        if g_up >= G:
            return psi, phi, sweeps, 1
        change = math.sqrt(change_sq / norm_sq) if norm_sq > 0.0 else math.sqrt(change_sq)
        have_rho = outer >= 2 and outer_change > 0.0
        rho = change / outer_change if have_rho else 0.0
        outer_change = change
        est = change * rho / (1.0 - rho) if (have_rho and rho < 1.0) else math.inf
        if trace is not None:
            trace.append((change, rho, est, sweep_rtol))
        if change == 0.0 or est < outer_rtol:
            if sweep_rtol == inner_rtol:
                return psi, phi, sweeps, 1
            final_sweep = True
    return psi, phi, max_outer, -2
