"""The DSA stage with voids on the GPU (reference src/dsak.kokkos.cxx:233-345,
bridged by default, masked with -dsa_void_bridge 0; the reference's void
files, tests/Makefile "DSA with VOIDS"):

* the bridged and the masked diffusion operators of the structured, DG0 and
  DG1 void files equal the oracle's (oracle/dsa.py, pinned by the
  hand-computed chords and the masked strip = vacuum face identity) to 1e-12,
  and their applies with a tight inner solve the oracle's exact apply;
* on those files the DSA stage cuts the Krylov iterations the way the
  reference's pins do (bridged / masked: box_void_channel 31 -> 6 / 8,
  cube_void_duct 22 -> 5 / 7, plex_box_void_channel 31 -> 6 / 8, its DG1 twin
  37 -> 10 / 11; those counts are with PCAIR, this build's stage is the
  sweep) and the solve keeps the oracle's scalar flux;
* the four-group decades files with a void channel in every group and a
  block void in the top group only run the group sweep with DSA to the
  oracle's scalar flux, DSA cutting the thick groups' iterations (reference
  pin with DSA: 7 per group, tests/Makefile);
* a group void everywhere is masked whole: a zero correction, the solve the
  same as without DSA; voids that cannot be bridged fall back to the mask.
"""
import os

import numpy as np
import pytest
import torch

import oracle
import ubolt_amd as u
from helpers import oracle_dsa_apply, oracle_dsa_matrix, oracle_rhs, oracle_solve, rel_max, to_dev, to_host
from ubolt_amd import problem_file as pfm
from ubolt_amd import problems as P

pytestmark = pytest.mark.gpu

PROBLEMS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "problems")
VOID_FILES = ["box_void_channel", "plex_box_void_channel", "plex_box_void_channel_dg1", "cube_void_duct"]
# the reference's FGMRES counts without DSA, with bridged and with masked DSA
REF_COUNTS = {"box_void_channel": (31, 6, 8), "plex_box_void_channel": (31, 6, 8),
              "plex_box_void_channel_dg1": (37, 10, 11), "cube_void_duct": (22, 5, 7)}


def load(name):
    return pfm.load(os.path.join(PROBLEMS, name + ".json")).problem


@pytest.mark.parametrize("bridge", [True, False])
@pytest.mark.parametrize("name", VOID_FILES)
def test_void_matvec_and_apply(gpu, name, bridge):
    prob = load(name)
    op = prob.operator()
    dsa = u.DSA(op, inner_max_it=3000, inner_rtol=1e-14, void_bridge=bridge)
    dsa.set_group(0)
    M, _ = oracle_dsa_matrix(prob, bridge=bridge)
    x = np.random.default_rng(3).standard_normal(prob.disc.n_nodes)
    y = torch.zeros(prob.disc.n_nodes, dtype=torch.float64, device="cuda")
    dsa.matvec(to_dev(x), y)
    xr = np.random.default_rng(4).standard_normal(prob.disc.n_rows)
    ya = op.vec()
    dsa.apply(to_dev(xr), ya)
    torch.cuda.synchronize()
    em, ea = rel_max(to_host(y), M @ x), rel_max(to_host(ya), oracle_dsa_apply(prob, xr, bridge=bridge))
    print(f"[case] {name} {'bridged' if bridge else 'masked'}: matvec {em:.2e}, apply {ea:.2e}", flush=True)
    assert em <= 1e-12 and ea <= 1e-9


@pytest.mark.parametrize("name,g,opts", [("box_void_channel", 0, dict(void_d=0.05)),
                                         ("plex_box_void_channel_dg1", 0, dict(void_d=0.05)),
                                         ("box_decades4_void", 0, dict(void_sigma_t=0.2)),
                                         ("plex_decades4_void", 0, dict(void_sigma_t=0.2, bridge=False))])
def test_void_threshold_and_fixed_d(gpu, name, g, opts):
    """-dsa_void_d fixes a bridged void's D; -dsa_void_sigma_t 0.2 makes the
    decades files' thin top group (sigma_t 0.1) void outside its dense block"""
    prob = load(name)
    op = prob.operator()
    dsa = u.DSA(op, void_bridge=opts.get("bridge", True), void_sigma_t=opts.get("void_sigma_t", 0.0),
                void_d=opts.get("void_d", 0.0))
    dsa.set_group(g)
    M, _ = oracle_dsa_matrix(prob, g, **opts)
    x = np.random.default_rng(8).standard_normal(prob.disc.n_nodes)
    y = torch.zeros(prob.disc.n_nodes, dtype=torch.float64, device="cuda")
    dsa.matvec(to_dev(x), y)
    torch.cuda.synchronize()
    em = rel_max(to_host(y), M @ x)
    print(f"[case] {name} group {g} {opts}: matvec {em:.2e}", flush=True)
    assert em <= 1e-12


@pytest.mark.parametrize("bridge", [True, False])
@pytest.mark.parametrize("name", VOID_FILES)
def test_void_dsa_cuts_iterations(gpu, name, bridge):
    ref_plain, ref_b, ref_m = REF_COUNTS[name]
    prob = load(name)
    op = prob.operator()
    its, flux = {}, {}
    for use in (False, True):
        d = None
        if use:
            d = u.DSA(op, inner_max_it=200, inner_rtol=1e-6, void_bridge=bridge)
            d.set_group(0)
        sol = u.Solver(op, d, rtol=1e-10, max_it=500, pc_dsa=use)
        sol.refresh()
        b = op.vec()
        op.rhs(0, b)
        x = op.vec()
        res = sol.solve(b, x)
        torch.cuda.synchronize()
        assert res["reason"] == 1
        its[use] = res["iterations"]
        flux[use] = oracle.moments(to_host(x), prob.quad.w)
    want = oracle.moments(oracle_solve(prob, oracle_rhs(prob, 0, [])), prob.quad.w)
    err = max(np.linalg.norm(flux[k] - want) / np.linalg.norm(want) for k in flux)
    kind = "bridged" if bridge else "masked"
    print(f"[case] {name}: FGMRES iterations without DSA {its[False]}, with {kind} DSA {its[True]} (reference "
          f"{ref_plain} -> {ref_b if bridge else ref_m}); scalar flux vs oracle {err:.2e}", flush=True)
    assert 2 * its[True] <= its[False]
    assert err <= 1e-8


@pytest.mark.parametrize("bridge", [True, False])
@pytest.mark.parametrize("name", ["box_decades4_void", "plex_decades4_void"])
def test_decades_voids_group_sweep_with_dsa(gpu, name, bridge):
    pf = pfm.load(os.path.join(PROBLEMS, name + ".json"))
    prob = pf.problem
    psi0, r0 = pfm.solve(pf, rtol=1e-10, max_it=800)
    psi1, r1 = pfm.solve(pf, rtol=1e-10, max_it=800, dsa=True, dsa_void_bridge=bridge)
    torch.cuda.synchronize()
    cache, errs = [], []
    for g in range(prob.n_groups):
        want = oracle.moments(oracle_solve(prob, oracle_rhs(prob, g, cache), g), prob.quad.w)
        cache.append(want)
        errs.append(np.linalg.norm(oracle.moments(to_host(psi1[g]), prob.quad.w) - want) / np.linalg.norm(want))
    kind = "bridged" if bridge else "masked"
    print(f"[case] {name}: per-group iterations without DSA {r0['group_iterations']}, with {kind} DSA "
          f"{r1['group_iterations']} (reference pin with bridged DSA 7); scalar flux vs oracle {max(errs):.2e}",
          flush=True)
    assert r1["status"] == 1 and max(errs) <= 1e-8
    assert max(r1["group_iterations"]) < max(r0["group_iterations"])


def test_all_void_group_is_masked(gpu):
    prob = P.fd_box_2d_central_source(n=6, sn=2)
    op = prob.operator()
    n = prob.disc.n_cells
    op.set_xs(np.zeros(n), np.zeros(n), np.ones(n))
    dsa = u.DSA(op)
    dsa.set_group(0)
    x = to_dev(np.random.default_rng(5).standard_normal(prob.disc.n_rows))
    y = op.vec()
    y.fill_(1.0)
    dsa.apply(x, y)
    torch.cuda.synchronize()
    assert float(y.abs().max()) == 0.0
    with pytest.raises(u.UbtError) as e:
        dsa.matvec(torch.zeros(n, dtype=torch.float64, device="cuda"), torch.zeros(n, dtype=torch.float64,
                                                                                      device="cuda"))
    assert e.value.status == "ERR_ARG_WRONGSTATE"
    got = {}
    for use in (False, True):
        sol = u.Solver(op, dsa if use else None, rtol=1e-10, max_it=200, pc_dsa=use)
        sol.refresh()
        b = op.vec()
        op.rhs(0, b)
        xs = op.vec()
        res = sol.solve(b, xs)
        torch.cuda.synchronize()
        assert res["reason"] == 1
        got[use] = (res["iterations"], to_host(xs))
    assert got[True][0] == got[False][0]
    assert rel_max(got[True][1], got[False][1]) <= 1e-12


def test_unbridgeable_void_falls_back_to_mask(gpu):
    pf = load("box_50_inf_medium")  # every face reflective
    op = pf.operator()
    st = np.ones(pf.disc.n_cells)
    st[7] = 0.0
    op.set_xs(st, st.copy(), np.ones_like(st))  # and no absorption
    dsa = u.DSA(op)
    dsa.set_group(0)
    prob = P.Problem("refl", pf.quad, pf.disc, st, st.copy(), None)
    M, _ = oracle_dsa_matrix(prob)
    x = np.random.default_rng(6).standard_normal(pf.disc.n_nodes)
    y = torch.zeros(pf.disc.n_nodes, dtype=torch.float64, device="cuda")
    dsa.matvec(to_dev(x), y)
    torch.cuda.synchronize()
    assert rel_max(to_host(y), M @ x) <= 1e-12
    assert float(M[7, 7]) == 1.0
