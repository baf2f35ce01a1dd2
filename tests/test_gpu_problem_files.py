"""The reference's own problem files (tests/golden/problems) through the GPU
group sweep: the scalar flux of every group against the oracle's direct
solves (downscatter: group by group with the oracle's transfers), and with
upscatter the converged multigroup residual with the oracle's matrices."""
import os

import numpy as np
import pytest
import torch

import oracle
from helpers import oracle_full, oracle_rhs, oracle_solve, to_host, xs_arrays
from ubolt_amd import problem_file as pfm

pytestmark = pytest.mark.gpu

PROBLEMS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "problems")

DOWNSCATTER = ["box_50_st2", "box_crooked_pipe", "box_50_reflect_lb", "box_layers", "cube_10_st2",
               "cube_10_mg4_t05", "cube_10_reflect_3faces", "plex_box_50_st2_dg1", "plex_tri_30_st2_dg1",
               "plex_square_msh_dg1", "plex_square_msh", "plex_tet_6_st2_dg1", "plex_cube_10_st2_dg1",
               "plex_box_50_reflect_lb_dg1", "plex_tri_30_st2"]


@pytest.mark.parametrize("name", DOWNSCATTER)
def test_problem_file_solve(gpu, name):
    pf = pfm.load(os.path.join(PROBLEMS, name + ".json"))
    assert not pf.upscatter
    prob = pf.problem
    print(f"[case] {name}: {prob.disc.n_cells} cells, nb {prob.disc.nb}, {prob.disc.n_angles} angles, "
          f"{prob.n_groups} groups", flush=True)
    psi, rep = pfm.solve(pf, rtol=1e-12, max_it=800)
    torch.cuda.synchronize()
    assert rep["status"] == 1 and rep["sweeps"] == 1 and rep["g_up"] == prob.n_groups, rep
    cache = []
    for g in range(prob.n_groups):
        want = oracle_solve(prob, oracle_rhs(prob, g, cache), g)
        phw = oracle.moments(want, prob.quad.w)
        cache.append(phw)
        ph = oracle.moments(to_host(psi[g]), prob.quad.w)
        err = np.linalg.norm(ph - phw) / np.linalg.norm(phw)
        its = rep["group_iterations"][g]
        print(f"[case] {name} group {g}: its {its}, scalar flux rel L2 vs oracle {err:.2e}", flush=True)
        assert rep["group_reason"][g] == 1 and err <= 1e-8, (g, err)


@pytest.mark.parametrize("name", ["box_up3_layers", "plex_box_up3_layers_dg1"])
def test_problem_file_upscatter(gpu, name):
    """the outer iteration over the thermal groups: at convergence every
    group's transport equation holds with every other group's final flux"""
    pf = pfm.load(os.path.join(PROBLEMS, name + ".json"))
    assert pf.upscatter
    prob = pf.problem
    psi, rep = pfm.solve(pf, rtol=1e-12, max_it=800, outer_rtol=1e-11, max_outer=200)
    torch.cuda.synchronize()
    assert rep["status"] == 1, rep
    G, W = prob.n_groups, prob.quad.sum_weights
    _, ss, _ = xs_arrays(prob)
    phis = [oracle.moments(to_host(p), prob.quad.w) for p in psi]
    outers = rep["sweeps"]
    for g in range(G):
        b = oracle_rhs(prob, g, phis[:g])
        for gf in range(g + 1, G):
            if np.any(ss[gf, g]):
                oracle.broadcast_add(prob.disc.nb, prob.disc.n_angles, ss[gf, g] / W, phis[gf], b)
        r = b - oracle_full(prob, g) @ to_host(psi[g])
        rel = np.linalg.norm(r) / np.linalg.norm(b)
        print(f"[case] {name} group {g}: multigroup residual {rel:.2e} after {outers} outer passes", flush=True)
        assert rel <= 1e-9, (g, rel)


@pytest.mark.parametrize("name,max_dsa", [("box_diffusive", 7), ("cube_diffusive", 7), ("plex_box_diffusive_dg1", 9),
                                          ("plex_cube_diffusive_dg1", 9)])
def test_dsa_in_the_diffusive_regime(gpu, name, max_dsa):
    """the reference's diffusive problem files (sigma_t 100 over 0.1-wide cells,
    scattering ratio 0.99): the DSA stage after the sweep cuts the Krylov
    iterations the way the reference's DSA pins do (docs/dev/iteration_counts.md:
    box_diffusive 29 -> 6, plex_box_diffusive_dg1 34 -> 7, plex_cube_diffusive_dg1
    8 with DSA; those counts are with PCAIR, this build's stage is the sweep)"""
    import ubolt_amd as u
    pf = pfm.load(os.path.join(PROBLEMS, name + ".json"))
    op = pf.problem.operator()
    its = {}
    for use in (False, True):
        d = None
        if use:
            d = u.DSA(op, inner_max_it=200, inner_rtol=1e-6)
            d.set_group(0)
        sol = u.Solver(op, d, rtol=1e-8, max_it=500, pc_dsa=use)
        sol.refresh()
        b = op.vec()
        op.rhs(0, b)
        x = op.vec()
        res = sol.solve(b, x)
        assert res["reason"] == 1
        its[use] = res["iterations"]
    print(f"[case] {name}: FGMRES iterations without DSA {its[False]}, with DSA {its[True]}", flush=True)
    assert its[True] <= max_dsa and 3 * its[True] <= its[False]
