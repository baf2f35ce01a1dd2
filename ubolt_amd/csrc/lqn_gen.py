#!/usr/bin/env python3
"""Generate lqn_table.inc: the level-symmetric (LQn) SN constants, S2-S18.

    python3 ubolt_amd/csrc/lqn_gen.py           # write lqn_table.inc beside this file
    python3 ubolt_amd/csrc/lqn_gen.py --check   # regenerate and compare, write nothing

The construction is the classical one (Lewis & Miller, Computational Methods of
Neutron Transport, ch. 4; Carlson's LQn sets), solved at 60 digits with mpmath
and rounded to double on output. It follows the same defining conditions as the
reference's offline generator (src/sn_lqn_table.py), so the emitted numbers are
the same constants:

* order N, n = N/2 levels, squared cosines in arithmetic progression
  x_l = x_1 + (l-1) d with d = (1 - 3 x_1)/(n - 1);
* one octant holds the ordered level triples (i, j, k), i + j + k = n + 2, in
  lexicographic order; a point's weight depends only on its sorted triple;
* the level sums y_l (weight of the points whose first index is l) satisfy the
  even axis moments  sum_l x_l^m y_l = 1/(8(2m+1)), m = 0, 2, 3, ..., n
  (m = 1 is implied by m = 0), and x_1 is the root of
  sum_l (3l - (n+2)) y_l = 0, the condition for y to come from class weights;
* class weights are recovered from the level sums; where that map has a null
  space (S14 and up) the member minimising the relative residuals of the mixed
  even moments <mu^2a eta^2b xi^2c>, a >= b >= c, b > 0, a+b+c <= n, in least
  squares, is taken.

Published first cosines seed the root finder and are checked against.
"""
import itertools
import os
import sys
from fractions import Fraction

from mpmath import matrix, mp, mpf, findroot, lu_solve, sqrt

mp.dps = 60
ORDERS = range(2, 20, 2)
# Lewis & Miller Table 4-1 first cosines: root-finder seeds and a cross-check
SEED_MU1 = {4: "0.3500212", 6: "0.2666355", 8: "0.2182179", 10: "0.1893213",
            12: "0.1672126", 14: "0.1519859", 16: "0.1389568", 18: "0.1293445"}
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lqn_table.inc")


def octant(n):
    return [(i, j, n + 2 - i - j) for i in range(1, n + 1) for j in range(1, n + 1)
            if 1 <= n + 2 - i - j <= n]


def squares(n, x1):
    if n == 1:
        return [mpf(1) / 3]
    d = (1 - 3 * x1) / (n - 1)
    return [x1 + l * d for l in range(n)]


def level_sums(n, x):
    powers = [0] + list(range(2, n + 1))
    V = matrix([[xl ** m for xl in x] for m in powers])
    r = matrix([mpf(1) / (8 * (2 * m + 1)) for m in powers])
    return lu_solve(V, r)


def first_level(n):
    if n == 1:
        return mpf(1) / 3
    f = lambda mu1: sum((3 * (l + 1) - (n + 2)) * y for l, y in
                        enumerate(level_sums(n, squares(n, mu1 * mu1))))
    s = mpf(SEED_MU1[2 * n])
    mu1 = findroot(f, (s * mpf("0.999"), s * mpf("1.001")), solver="secant", tol=mpf("1e-80"))
    return mu1 * mu1


def dfact(k):
    return 1 if k <= 1 else k * dfact(k - 2)


def sphere(a, b, c):
    return Fraction(dfact(2 * a - 1) * dfact(2 * b - 1) * dfact(2 * c - 1), dfact(2 * (a + b + c) + 1))


def exact_nullspace(rows, ncols):
    """Particular-solution map and null basis of an integer matrix, exactly."""
    A = [[Fraction(v) for v in r] + [Fraction(int(i == k)) for k in range(len(rows))]
         for i, r in enumerate(rows)]
    piv, r = [], 0
    for c in range(ncols):
        p = next((k for k in range(r, len(A)) if A[k][c] != 0), None)
        if p is None:
            continue
        A[r], A[p] = A[p], A[r]
        A[r] = [v / A[r][c] for v in A[r]]
        for k in range(len(A)):
            if k != r and A[k][c] != 0:
                A[k] = [u - A[k][c] * v for u, v in zip(A[k], A[r])]
        piv.append(c)
        r += 1
    free = [c for c in range(ncols) if c not in piv]
    return A, piv, free


def build(N):
    n = N // 2
    x = squares(n, first_level(n))
    pts = octant(n)
    classes = sorted({tuple(sorted(t)) for t in pts})
    cid = {c: i for i, c in enumerate(classes)}
    y = level_sums(n, x)
    m2l = [[0] * len(classes) for _ in range(n)]
    for t in pts:
        m2l[t[0] - 1][cid[tuple(sorted(t))]] += 1
    A, piv, free = exact_nullspace(m2l, len(classes))
    nc = len(classes)
    base = [mpf(0)] * nc
    for row, p in enumerate(piv):
        base[p] = sum(mpf(A[row][nc + k].numerator) / A[row][nc + k].denominator * y[k] for k in range(n))
    null = []
    for f in free:
        v = [mpf(0)] * nc
        v[f] = mpf(1)
        for row, p in enumerate(piv):
            v[p] = -mpf(A[row][f].numerator) / A[row][f].denominator
        null.append(v)

    def mom(wc, a, b, c):
        return sum(wc[cid[tuple(sorted(t))]] * x[t[0] - 1] ** a * x[t[1] - 1] ** b * x[t[2] - 1] ** c
                   for t in pts)

    keys = [(a, b, c) for a in range(n + 1) for b in range(min(a, n - a) + 1)
            for c in range(min(b, n - a - b) + 1) if b > 0]
    wc = base
    if null:
        want = [mpf(sphere(*k).numerator) / sphere(*k).denominator / 8 for k in keys]
        G = matrix(len(keys), len(null))
        r0 = matrix(len(keys), 1)
        for i, k in enumerate(keys):
            r0[i] = (mom(base, *k) - want[i]) / want[i]
            for j, v in enumerate(null):
                G[i, j] = mom(v, *k) / want[i]
        t = lu_solve(G.T * G, -(G.T * r0))
        wc = [base[c] + sum(t[j] * null[j][c] for j in range(len(null))) for c in range(nc)]
    for m in range(n + 1):  # every axis moment, m = 1 included
        got = sum(wc[cid[tuple(sorted(tp))]] * x[tp[0] - 1] ** m for tp in pts)
        assert abs(got - mpf(1) / (8 * (2 * m + 1))) < mpf("1e-50"), (N, m)
    assert min(wc) > 0, N
    mu = [sqrt(v) for v in x]
    if N in SEED_MU1:
        assert abs(mu[0] - mpf(SEED_MU1[N])) < mpf("6e-7"), (N, mu[0])
    return mu, pts, [wc[cid[tuple(sorted(t))]] for t in pts]


def render():
    out = ["// GENERATED by lqn_gen.py - level-symmetric SN constants at 60 digits,",
           "// rounded to double. Regenerate, do not edit.",
           "// Per order: sn_order, n_levels, n_points, level cosines, then per point",
           "// the 1-based level triple and the weight as a fraction of the sphere."]
    for N in ORDERS:
        mu, pts, w = build(N)
        out.append("{%d, %d, %d, {%s}," % (N, len(mu), len(pts), ", ".join(repr(float(v)) for v in mu)))
        out.append(" {%s}," % ", ".join("%d, %d, %d" % t for t in pts))
        out.append(" {%s}}," % ", ".join(repr(float(v)) for v in w))
    return "\n".join(out) + "\n"


if __name__ == "__main__":
    text = render()
    if "--check" in sys.argv:
        with open(OUT) as f:
            same = f.read() == text
        print("lqn_table.inc is", "current" if same else "STALE")
        sys.exit(0 if same else 1)
    with open(OUT, "w") as f:
        f.write(text)
    print("wrote", OUT)
