"""Closed-form pins of the DSA restatement's boundary terms (oracle/dsa.py;
reference src/dsa_structuredk.kokkos.cxx:183-285, src/dsa_plexk.kokkos.cxx:
288-475): on a slab - a box with reflective sides along y, so the solution
depends on x only - with uniform D = 1/(3 sigma_t), sigma_a and source S,

* the structured star with Marshak vacuum faces converges at second order to
  phi = S/sa (1 - cosh(k (x - L/2)) / (cosh(k L/2) + 2 D k sinh(k L/2))),
  k = sqrt(sa / D), the solution of -D phi'' + sa phi = S with
  phi -/+ 2 D phi' = 0 at x = 0, L (a wrong Marshak factor converges to
  another function);
* the consistent-D blend moves the face D by O(h^p): its difference from the
  plain operator's solution shrinks by 2^p per halving of h, pinning p = 1.5
  (p = 1 would halve it, p = 2 quarter it);
* the DG1 MIP operator, whose vacuum penalty max(C D / h, 1/4) is C D / h on
  these cells, converges to the weak-Dirichlet solution
  phi = S/sa (1 - cosh(k (x - L/2)) / cosh(k L/2)), its cell averages at
  first order in h (the boundary's consistency terms carry the factor 1/2 of
  the interior face average), and stays away from the Marshak solution (the
  limit a Robin boundary of the 1/4 floor alone would approach).
"""
import numpy as np
import scipy.sparse.linalg as spl

import ubolt_amd as u
from helpers import oracle_dsa_matrix
from oracle import dsa as odsa
from ubolt_amd import problems as P

L, ST, SA, S = 1.0, 1.0, 0.5, 1.0
D = 1.0 / (3.0 * ST)
K = np.sqrt(SA / D)


def marshak(x):
    den = np.cosh(K * L / 2) + 2 * D * K * np.sinh(K * L / 2)
    return S / SA * (1 - np.cosh(K * (x - L / 2)) / den)


def dirichlet_cell_average(a, b):
    # (1/h) int_a^b of S/SA (1 - cosh(K (x - L/2)) / cosh(K L/2))
    integ = (np.sinh(K * (b - L / 2)) - np.sinh(K * (a - L / 2))) / (K * np.cosh(K * L / 2))
    return S / SA * (1 - integ / (b - a))


def slab_structured(n, cons, p=1.5):
    q = u.Quadrature(2, 4)
    face_type = np.array([[0, 0], [1, 1], [1, 1]])  # x vacuum, y reflective
    st, ss = np.full(n, ST), np.full(n, ST - SA)
    M = odsa.structured(2, [n, 1, 1], [L / n, 1.0, 1.0], face_type, q.omega, q.w, q.sum_weights, st, ss, cons, p)
    phi = spl.spsolve(M.tocsc(), np.full(n, S))  # the star is per unit volume
    x = (np.arange(n) + 0.5) * L / n
    return x, phi


def test_structured_marshak_slab_converges_second_order():
    errs = []
    for n in (20, 40, 80, 160):
        x, phi = slab_structured(n, cons=False)
        errs.append(np.abs(phi - marshak(x)).max() / marshak(x).max())
    rates = [errs[i] / errs[i + 1] for i in range(len(errs) - 1)]
    print("Marshak slab errors", errs, "ratios", rates)
    assert errs[-1] < 2e-5
    assert all(3.5 < r < 4.5 for r in rates), rates


def test_structured_marshak_factor_is_pinned():
    """the same slab with a Robin coefficient of 1/4 instead of the Marshak
    1/2 (phi + 4 D phi' = 0): the restatement's solution stays away from it"""
    n = 160
    x, phi = slab_structured(n, cons=False)
    den = np.cosh(K * L / 2) + 4 * D * K * np.sinh(K * L / 2)
    other = S / SA * (1 - np.cosh(K * (x - L / 2)) / den)
    assert np.abs(phi - other).max() / other.max() > 1e-2


def test_consistent_blend_exponent():
    diffs = []
    for n in (40, 80, 160, 320):
        _, a = slab_structured(n, cons=True)
        _, b = slab_structured(n, cons=False)
        diffs.append(np.abs(a - b).max())
    rates = [diffs[i] / diffs[i + 1] for i in range(len(diffs) - 1)]
    print("blend differences", diffs, "ratios", rates)
    # p = 1.5: 2^1.5 = 2.83 per halving
    assert all(2.5 < r < 3.2 for r in rates), rates
    # and the blended solution converges to the same Marshak solution
    x, phi = slab_structured(320, cons=True)
    assert np.abs(phi - marshak(x)).max() / marshak(x).max() < 2e-3


def test_mip_slab_converges_to_weak_dirichlet():
    errs = []
    for n in (8, 16, 32, 64):
        coords, cells = P.box_quads(n, 1, L, L / n)
        q = u.Quadrature(2, 4)
        d = u.Disc.dg(q, u.Mesh(coords, cells), 1, {1: "reflect", 3: "reflect"})
        prob = P.Problem("slab", q, d, np.full(n, ST), np.full(n, ST - SA), None)
        M, _ = oracle_dsa_matrix(prob)
        nb = d.nb
        rhs = np.zeros(n * nb)
        rhs[::nb] = S * d.get("volume")  # int S phi_0 (phi_0 = 1; the others have zero mean)
        sol = spl.spsolve(M.tocsc(), rhs)
        cen = d.get("centroid").reshape(-1, 3)[:, 0]
        h = L / n
        want = dirichlet_cell_average(cen - h / 2, cen + h / 2)
        errs.append(np.abs(sol[::nb] - want).max() / want.max())
        away = np.abs(sol[::nb] - marshak(cen)).max() / marshak(cen).max()
    rates = [errs[i] / errs[i + 1] for i in range(len(errs) - 1)]
    print("MIP slab errors", errs, "ratios", rates, "distance from the Marshak solution", away)
    assert errs[-1] < 1e-2
    assert all(1.8 < r < 2.2 for r in rates), rates
    assert away > 0.1
