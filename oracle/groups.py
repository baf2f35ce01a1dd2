"""The oracle's group sweep - TEST INFRASTRUCTURE ONLY.

A restatement of the reference test driver's multigroup loop
(tests/transportk.kokkos.cxx:925-1208) with exact inner solves: the
Gauss-Seidel sweep over the groups (a group's rhs takes every transfer out of
the cached scalar flux of the groups above it, and a thermal group g >= g_up
every other group's: zero before that group is first solved, :964-976,
:1058-1065), the scalar flux cached after each solve with the thermal block's
change summed (src/multigroupk.kokkos.cxx:266-299), and the outer test on the
error estimate change * rho / (1 - rho) (:1159-1207). With exact inner solves
the adaptive inner tolerance (-outer_inner_gamma) has nothing to loosen: this
is the gamma = 0 iteration.
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


def group_sweep(A, b0, transfer, moments, g_up, outer_rtol=1e-8, max_outer=1000, trace=None):
    """A[g]: group g's operator (scipy sparse); b0[g]: its inflow and source;
    transfer(gf, gt, phi): the rhs contribution of group gf's scalar flux to
    group gt; moments(psi): the scalar flux. Returns (psi, phi, sweeps,
    status) with status 1 converged, -2 out of sweeps; a trace list receives
    (change, rho, estimate) per sweep."""
    G = len(A)
    n_nodes = len(moments(np.zeros(A[0].shape[0])))
    phi = [np.zeros(n_nodes) for _ in range(G)]
    psi = [None] * G
    lu = [spl.splu(a.tocsc()) for a in A]
    outer_change = 0.0
    for outer in range(max_outer):
        change_sq = norm_sq = 0.0
        for g in (range(G) if outer == 0 else range(g_up, G)):
            b = b0[g].copy()
            for gf in range(G if g >= g_up else g):
                if gf != g:
                    b += transfer(gf, g, phi[gf])
            psi[g] = lu[g].solve(b)
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
            trace.append((change, rho, est))
        if change == 0.0 or est < outer_rtol:
            return psi, phi, sweeps, 1
    return psi, phi, max_outer, -2
