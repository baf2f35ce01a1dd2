"""The oracle's group sweep (oracle/groups.py) pinned on the CPU:

* g_up on hand-made transfer tables (no upscatter, one upscatter pair, the
  deepest pair winning);
* the infinite-medium file converges to the closed-form per-group constants
  sigma_t,g phi_g - sum_g' sigma_s(g' -> g) phi_g' = S_g (the reference's
  -check_inf_medium oracle) within its pinned sweeps (tests/Makefile:
  -outer_max_it 27);
* box_up3_layers at gamma 0 converges within the reference's pin for that
  setting (tests/Makefile: -outer_inner_gamma 0 ... -outer_max_it 15) to a
  flux with a small multigroup residual, and at the default gamma 0.01 with
  warm-started inexact inner solves within that recipe's pin (16);
* a planted mistake - the thermal rhs without the upscatter transfers -
  misses the closed form.
"""
import os

import numpy as np

import oracle
from oracle import groups as og
from helpers import oracle_full, oracle_rhs, xs_arrays
from ubolt_amd import problem_file as pfm

PROBLEMS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "problems")


def group_system(prob, drop_upscatter=False):
    """the oracle's per-group matrices, fixed rhs and transfers of a problem"""
    d, q = prob.disc, prob.quad
    G, W = prob.n_groups, q.sum_weights
    _, ss, _ = xs_arrays(prob)
    A = [oracle_full(prob, g) for g in range(G)]
    b0 = [oracle_rhs(prob, g, [np.zeros(d.n_nodes)] * g) for g in range(G)]

    def transfer(gf, gt, phi):
        if drop_upscatter and gf > gt:
            return np.zeros(d.n_rows)
        return oracle.broadcast_add(d.nb, d.n_angles, ss[gf, gt] / W, phi, np.zeros(d.n_rows))

    return A, b0, transfer, (lambda psi: oracle.moments(psi, q.w)), og.g_up_of(ss)


def test_g_up_tables():
    z = np.zeros((3, 3, 4))
    assert og.g_up_of(z) == 3
    down = z.copy()
    down[0, 1] = down[1, 2] = 0.5
    assert og.g_up_of(down) == 3
    up = down.copy()
    up[2, 1, 3] = 0.1          # one cell upscatters 2 -> 1
    assert og.g_up_of(up) == 1
    up[1, 0, 0] = 0.1          # and 1 -> 0: the smallest target wins
    assert og.g_up_of(up) == 0


def inf_medium_constants(prob):
    st, ss, q = xs_arrays(prob)
    return np.linalg.solve(np.diag(st[:, 0]) - ss[:, :, 0].T, q[:, 0])


def test_infinite_medium_closed_form():
    prob = pfm.load(os.path.join(PROBLEMS, "box_up2_inf_medium.json")).problem
    A, b0, tr, mom, g_up = group_system(prob)
    assert g_up == 0
    psi, phi, sweeps, status = og.group_sweep(A, b0, tr, mom, g_up, outer_rtol=1e-11, max_outer=27)
    want = inf_medium_constants(prob)
    print(f"[case] oracle infinite medium: {sweeps} sweeps, phi {[float(p.mean()) for p in phi]} want {want}")
    assert status == 1 and sweeps <= 27
    for g in range(prob.n_groups):
        assert np.abs(phi[g] - want[g]).max() <= 1e-9 * abs(want[g]), g


def test_planted_mistake_misses_the_closed_form():
    prob = pfm.load(os.path.join(PROBLEMS, "box_up2_inf_medium.json")).problem
    A, b0, tr, mom, g_up = group_system(prob, drop_upscatter=True)
    psi, phi, sweeps, status = og.group_sweep(A, b0, tr, mom, g_up, outer_rtol=1e-11, max_outer=27)
    want = inf_medium_constants(prob)
    assert max(np.abs(phi[g] - want[g]).max() / abs(want[g]) for g in range(prob.n_groups)) > 1e-3


def test_up3_layers_within_gamma0_pin():
    prob = pfm.load(os.path.join(PROBLEMS, "box_up3_layers.json")).problem
    A, b0, tr, mom, g_up = group_system(prob)
    psi, phi, sweeps, status = og.group_sweep(A, b0, tr, mom, g_up)
    print(f"[case] oracle box_up3_layers: g_up {g_up}, {sweeps} sweeps")
    assert status == 1 and sweeps <= 15
    # every group's equation with every other group's final flux
    for g in range(prob.n_groups):
        b = b0[g].copy()
        for gf in range(prob.n_groups):
            if gf != g:
                b += tr(gf, g, phi[gf])
        assert np.linalg.norm(b - A[g] @ psi[g]) <= 1e-7 * np.linalg.norm(b), g


def test_up3_layers_adaptive_inner_within_pin():
    """the adaptive inner tolerance (gamma 0.01, inner rtol 1e-10, the
    reference test driver's defaults) with warm-started inexact inner solves
    (right-preconditioned GMRES(30), the exact unscattered inverse as the
    preconditioner - the sweep on this all-vacuum box):
    within the reference's pin for that recipe (tests/Makefile: box_up3_layers
    -outer_max_it 16), its last sweep at the inner rtol, the thermal rtols
    loosened to gamma on the first sweep and tightening with the change"""
    import scipy.sparse.linalg as spl
    from helpers import oracle_rows
    prob = pfm.load(os.path.join(PROBLEMS, "box_up3_layers.json")).problem
    A, b0, tr, mom, g_up = group_system(prob)
    L = [spl.splu(oracle_rows(prob, g).to_scipy().tocsc()) for g in range(prob.n_groups)]
    inner = lambda g, b, x0, rtol: og.gmres_right(A[g], L[g].solve, b, x0, rtol)
    trace = []
    psi, phi, sweeps, status = og.group_sweep(A, b0, tr, mom, g_up, inner=inner, inner_gamma=0.01,
                                              inner_rtol=1e-10, trace=trace)
    rtols = [t[3] for t in trace]
    print(f"[case] oracle box_up3_layers gamma 0.01: {sweeps} sweeps, thermal rtols {['%.1e' % r for r in rtols]}")
    assert status == 1 and sweeps <= 16
    assert rtols[0] == 0.01 and rtols[-1] == 1e-10
    assert all(r1 <= r0 for r0, r1 in zip(rtols, rtols[1:]))
    for g in range(prob.n_groups):
        b = b0[g].copy()
        for gf in range(prob.n_groups):
            if gf != g:
                b += tr(gf, g, phi[gf])
        assert np.linalg.norm(b - A[g] @ psi[g]) <= 1e-7 * np.linalg.norm(b), g


def test_gmres_right_matches_direct_solve():
    import scipy.sparse.linalg as spl
    from helpers import oracle_rows
    prob = pfm.load(os.path.join(PROBLEMS, "box_up3_layers.json")).problem
    A, b0, _, _, _ = group_system(prob)
    g = 2
    L = spl.splu(oracle_rows(prob, g).to_scipy().tocsc())
    want = spl.splu(A[g].tocsc()).solve(b0[g])
    for x0 in (None, 0.5 * want):
        x = og.gmres_right(A[g], L.solve, b0[g], x0, 1e-12, restart=5)
        assert np.linalg.norm(b0[g] - A[g] @ x) <= 1e-12 * np.linalg.norm(b0[g])
        assert np.linalg.norm(x - want) <= 1e-9 * np.linalg.norm(want)


def test_out_of_sweeps_status():
    prob = pfm.load(os.path.join(PROBLEMS, "box_up3_layers.json")).problem
    A, b0, tr, mom, g_up = group_system(prob)
    _, _, sweeps, status = og.group_sweep(A, b0, tr, mom, g_up, max_outer=3)
    assert (sweeps, status) == (3, -2)
