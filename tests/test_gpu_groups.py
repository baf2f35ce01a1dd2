"""The library's multigroup driver (ubt_solver_solve_groups and
ubt_batch_solver_solve_groups; reference tests/transportk.kokkos.cxx:925-1208):
the Gauss-Seidel group sweep and the upscatter outer iteration.

* the infinite-medium files (all faces reflective, one material): the exact
  discrete solution is the per-group constant of the G x G group system
  sigma_t,g phi_g - sum_g' sigma_s(g' -> g) phi_g' = S_g, which the outer
  iteration must reach (the reference's -check_inf_medium oracle) within the
  reference's pinned sweep count (tests/Makefile: -outer_max_it 27);
* the *_up3_layers regressions at the reference test driver's defaults
  (outer_rtol 1e-8, inner rtol 1e-10, inner_gamma 0.01) converge within the
  reference's pinned sweeps (16, 15 at inner_gamma 0, 17 for DG1) to a flux
  whose multigroup residual with the oracle's matrices is at the outer
  tolerance;
* the outer_rtol below the inner rtol is refused like the reference's;
* a batch of problems of different group counts, with and without
  upscatter, gives every problem the flux and sweep count of its own
  single-problem sweep;
* at gamma 0 with tight inner solves the library's sweep takes the oracle's
  sweep count (oracle/groups.py, exact inner solves) with its change and
  scalar flux.
"""
import os

import numpy as np
import pytest
import torch

import oracle
import ubolt_amd as u
from helpers import oracle_full, oracle_rhs, to_host, xs_arrays
from ubolt_amd import problem_file as pfm
from ubolt_amd import problems as P

pytestmark = pytest.mark.gpu

PROBLEMS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "problems")


def load(name):
    return pfm.load(os.path.join(PROBLEMS, name + ".json"))


def multigroup_residual(prob, psi):
    """max over groups of ||b_g - A_g psi_g|| / ||b_g||, b_g with every other
    group's final scalar flux (the oracle's matrices and transfers)"""
    G, W = prob.n_groups, prob.quad.sum_weights
    _, ss, _ = xs_arrays(prob)
    phis = [oracle.moments(to_host(p), prob.quad.w) for p in psi]
    worst = 0.0
    for g in range(G):
        b = oracle_rhs(prob, g, phis[:g])
        for gf in range(g + 1, G):
            if np.any(ss[gf, g]):
                oracle.broadcast_add(prob.disc.nb, prob.disc.n_angles, ss[gf, g] / W, phis[gf], b)
        r = b - oracle_full(prob, g) @ to_host(psi[g])
        worst = max(worst, np.linalg.norm(r) / np.linalg.norm(b))
    return worst


@pytest.mark.parametrize("name,gamma", [("box_up2_inf_medium", 0.01), ("box_up2_inf_medium", 0.0),
                                        ("plex_box_up2_inf_medium", 0.01), ("plex_box_up2_inf_medium_dg1", 0.01)])
def test_infinite_medium_constants(gpu, name, gamma):
    pf = load(name)
    prob = pf.problem
    st, ss, q = xs_arrays(prob)
    G, nb = prob.n_groups, prob.disc.nb
    # sigma_t,g phi_g - sum_g' sigma_s(g' -> g) phi_g' = S_g, uniform in space
    want = np.linalg.solve(np.diag(st[:, 0]) - ss[:, :, 0].T, q[:, 0])
    psi, rep = pfm.solve(pf, rtol=1e-12, max_it=800, outer_rtol=1e-11, inner_gamma=gamma, max_outer=27)
    torch.cuda.synchronize()
    print(f"[case] {name} gamma {gamma}: {rep['sweeps']} sweeps, change {rep['change']:.2e}, rho {rep['rho']:.3f}, "
          f"{rep['iterations']} inner iterations", flush=True)
    assert rep["status"] == 1 and rep["g_up"] == 0, rep
    for g in range(G):
        phi = oracle.moments(to_host(psi[g]), prob.quad.w).reshape(-1, nb)
        err = max(np.abs(phi[:, 0] - want[g]).max(), np.abs(phi[:, 1:]).max() if nb > 1 else 0.0) / abs(want[g])
        print(f"[case] {name} group {g}: phi {want[g]:.6f}, max error {err:.2e}", flush=True)
        assert err <= 1e-9, (g, err)


@pytest.mark.parametrize("name,gamma,pin", [("box_up3_layers", 0.01, 16), ("box_up3_layers", 0.0, 15),
                                            ("plex_box_up3_layers", 0.01, 16), ("plex_box_up3_layers_dg1", 0.01, 17)])
def test_outer_sweeps_within_reference_pins(gpu, name, gamma, pin):
    pf = load(name)
    prob = pf.problem
    psi, rep = pfm.solve(pf, max_it=800, inner_gamma=gamma, max_outer=pin)
    torch.cuda.synchronize()
    res = multigroup_residual(prob, psi)
    print(f"[case] {name} gamma {gamma}: {rep['sweeps']} sweeps (reference pin {pin}), g_up {rep['g_up']}, "
          f"error estimate {rep['err_est']:.2e}, per-group iterations {rep['group_iterations']}, "
          f"multigroup residual {res:.2e}", flush=True)
    assert rep["status"] == 1 and rep["sweeps"] <= pin, rep
    assert rep["err_est"] < 1e-8 or rep["change"] == 0.0
    assert res <= 1e-7, res


def test_outer_iteration_with_dsa_and_mixed(gpu):
    """the same sweep with the DSA stage, and with FP32 inner solves inside
    FP64 refinement: converged within the DG1 pin, the same flux"""
    pf = load("plex_box_up3_layers_dg1")
    prob = pf.problem
    base, rb = pfm.solve(pf, max_it=800)
    for kw in (dict(dsa=True), dict(mixed=True)):
        psi, rep = pfm.solve(pf, max_it=800, **kw)
        torch.cuda.synchronize()
        diff = max(float(torch.linalg.norm(psi[g] - base[g]) / torch.linalg.norm(base[g])) for g in range(3))
        print(f"[case] {kw}: {rep['sweeps']} sweeps (plain {rb['sweeps']}), iterations {rep['iterations']} "
              f"(plain {rb['iterations']}), psi difference {diff:.2e}", flush=True)
        assert rep["status"] == 1 and rep["sweeps"] <= 17, rep
        assert diff <= 1e-6, diff


@pytest.mark.parametrize("gamma", [0.0, 0.01])
@pytest.mark.parametrize("name", ["box_up3_layers", "plex_box_up3_layers", "plex_box_up3_layers_dg1"])
def test_outer_iteration_matches_oracle(gpu, name, gamma):
    """gamma 0: tight inner solves (rtol 1e-12) against the oracle's exact
    ones; gamma 0.01: the reference test driver's defaults (inner rtol 1e-10)
    against the oracle's warm-started right-preconditioned GMRES(30) with the
    exact unscattered inverse, the library's FGMRES(30) with its sweep (exact
    on these all-vacuum boxes)"""
    import scipy.sparse.linalg as spl
    from oracle import groups as og
    from test_oracle_groups import group_system
    from helpers import oracle_rows
    pf = load(name)
    prob = pf.problem
    A, b0, tr, mom, g_up = group_system(prob)
    trace = []
    if gamma == 0.0:
        _, phi_o, sweeps_o, status_o = og.group_sweep(A, b0, tr, mom, g_up, outer_rtol=1e-8, trace=trace)
        psi, rep = pfm.solve(pf, rtol=1e-12, max_it=800, outer_rtol=1e-8, inner_gamma=0.0)
    else:
        L = [spl.splu(oracle_rows(prob, g).to_scipy().tocsc()) for g in range(prob.n_groups)]
        inner = lambda g, b, x0, rtol: og.gmres_right(A[g], L[g].solve, b, x0, rtol)
        _, phi_o, sweeps_o, status_o = og.group_sweep(A, b0, tr, mom, g_up, outer_rtol=1e-8, trace=trace,
                                                      inner=inner, inner_gamma=gamma, inner_rtol=1e-10)
        psi, rep = pfm.solve(pf, max_it=800, outer_rtol=1e-8, inner_gamma=gamma)
    assert status_o == 1
    torch.cuda.synchronize()
    diff = max(np.abs(to_host(rep["operator"].scalar_flux(g)) - phi_o[g]).max() / np.abs(phi_o[g]).max()
               for g in range(prob.n_groups))
    change_o, rho_o = trace[-1][:2]
    print(f"[case] {name} gamma {gamma}: library {rep['sweeps']} sweeps (change {rep['change']:.4e}, rho {rep['rho']:.4f}), "
          f"oracle {sweeps_o} (change {change_o:.4e}, rho {rho_o:.4f}); scalar flux difference {diff:.2e}",
          flush=True)
    assert rep["status"] == 1 and rep["g_up"] == g_up and rep["sweeps"] == sweeps_o
    assert abs(rep["change"] - change_o) <= 1e-2 * change_o
    assert abs(rep["rho"] - rho_o) <= 1e-2 * rho_o
    assert diff <= (1e-9 if gamma == 0.0 else 1e-8), diff


def test_outer_rtol_below_inner_is_refused(gpu):
    pf = load("box_up3_layers")
    with pytest.raises(u.UbtError) as e:
        pfm.solve(pf, rtol=1e-10, outer_rtol=1e-11)
    assert e.value.status == "ERR_ARG_INCOMP"
    assert "below the inner rtol" in str(e.value)
    # without upscatter there is no outer test to protect
    psi, rep = pfm.solve(load("box_50_st2"), rtol=1e-10, outer_rtol=1e-11)
    assert rep["status"] == 1 and rep["sweeps"] == 1


def test_downscatter_sweep_is_one_pass_per_group(gpu):
    """downscatter only: one sweep, every group solved once, the result equal
    to the groups solved one by one through the C ABI"""
    pf = load("cube_10_mg4_t05")
    prob = pf.problem
    psi, rep = pfm.solve(pf, rtol=1e-10, max_it=800)
    assert rep["sweeps"] == 1 and rep["g_up"] == prob.n_groups and rep["status"] == 1
    op = prob.operator()
    sol = u.Solver(op, None, rtol=1e-10, max_it=800)
    b = op.vec()
    for g in range(prob.n_groups):
        op.set_group(g)
        sol.refresh()
        op.rhs(g, b)
        x = op.vec()
        r = sol.solve(b, x)
        op.set_scalar_flux(g, x)
        assert r["iterations"] == rep["group_iterations"][g]
        diff = float(torch.linalg.norm(x - psi[g]) / torch.linalg.norm(x))
        print(f"[case] group {g}: {r['iterations']} iterations, difference {diff:.1e}", flush=True)
        assert diff <= 1e-14, diff


def downscatter_twin(prob):
    st, ss, q = xs_arrays(prob)
    ss = ss.copy()
    for gf in range(prob.n_groups):
        ss[gf, :gf] = 0.0
    return P.Problem(prob.name + "_down", prob.quad, prob.disc, st.copy(), ss, q.copy(), n_groups=prob.n_groups)


def test_batch_group_sweep_matches_single(gpu):
    up3 = load("plex_box_up3_layers_dg1").problem
    probs = [up3, load("plex_box_up2_inf_medium_dg1").problem, load("plex_box_50_st2_dg1").problem,
             downscatter_twin(up3)]
    ops = [p.operator() for p in probs]
    batch = u.Batch(ops)
    bs = u.BatchSolver(batch, rtol=1e-10, max_it=800)
    psi_b, reps = bs.solve_groups()
    torch.cuda.synchronize()
    G = psi_b.shape[0]
    for p, prob in enumerate(probs):
        sol = u.Solver(prob.operator(), None, rtol=1e-10, max_it=800)
        psi_s, rs = sol.solve_groups()
        rb = reps[p]
        span = slice(int(batch.offsets[p]), int(batch.offsets[p + 1]))
        diffs = []
        for g in range(prob.n_groups):
            pb = oracle.moments(to_host(psi_b[g, span]), prob.quad.w)
            ps = oracle.moments(to_host(psi_s[g]), prob.quad.w)
            diffs.append(np.linalg.norm(pb - ps) / np.linalg.norm(ps))
        for g in range(prob.n_groups, G):
            assert float(psi_b[g, span].abs().max()) == 0.0
        print(f"[case] problem {p} ({prob.name}, {prob.n_groups} groups, g_up {rb['g_up']}): batch {rb['sweeps']} "
              f"sweeps / single {rs['sweeps']}, scalar flux difference {max(diffs):.2e}", flush=True)
        assert rb["status"] == 1 and rs["status"] == 1
        assert rb["g_up"] == rs["g_up"] and abs(rb["sweeps"] - rs["sweeps"]) <= 1
        assert max(diffs) <= 1e-7, diffs


# the moments store (-outer_store moments): one psi, each thermal group's
# solve after the first sweep warm-started from the P1 expansion of its
# scalar flux and current (recon_steps 0, -outer_recon none) or that
# expansion refined by Richardson steps on L_g (-outer_recon tol). Every solve
# still stops on its own true residual, so the answer is the angular store's
# (reference docs: the infinite-medium files stay the exact oracle and the
# sweep counts match the angular recipes)
@pytest.mark.parametrize("recon", [0, 10])
def test_moments_store_infinite_medium(gpu, recon):
    pf = load("box_up2_inf_medium")
    prob = pf.problem
    st, ss, q = xs_arrays(prob)
    want = np.linalg.solve(np.diag(st[:, 0]) - ss[:, :, 0].T, q[:, 0])
    psi, rep = pfm.solve(pf, rtol=1e-12, max_it=800, outer_rtol=1e-11, max_outer=27, store_moments=True,
                         recon_steps=recon)
    torch.cuda.synchronize()
    print(f"[case] moments store, recon {recon}: {rep['sweeps']} sweeps, {rep['iterations']} inner iterations, "
          f"{rep['recon_steps']} reconstruction steps", flush=True)
    assert rep["status"] == 1 and rep["sweeps"] <= 27, rep
    assert len(psi) == 1 and psi[0].numel() == prob.disc.n_rows
    assert (rep["recon_steps"] > 0) == (recon > 0)
    for g in range(prob.n_groups):
        phi = to_host(rep["operator"].scalar_flux(g))
        err = np.abs(phi - want[g]).max() / abs(want[g])
        print(f"[case] group {g}: phi {want[g]:.6f}, max error {err:.2e}", flush=True)
        assert err <= 1e-9, (g, err)


@pytest.mark.parametrize("name,pin", [("box_up3_layers", 16), ("plex_box_up3_layers_dg1", 17)])
def test_moments_store_matches_angular_store(gpu, name, pin):
    pf = load(name)
    G = pf.problem.n_groups
    _, ra = pfm.solve(pf, max_it=800)
    phi_a = [to_host(ra["operator"].scalar_flux(g)) for g in range(G)]
    for recon in (0, 10):
        _, rm = pfm.solve(pf, max_it=800, store_moments=True, recon_steps=recon)
        torch.cuda.synchronize()
        diff = max(np.linalg.norm(to_host(rm["operator"].scalar_flux(g)) - phi_a[g]) / np.linalg.norm(phi_a[g])
                   for g in range(G))
        print(f"[case] {name} recon {recon}: moments store {rm['sweeps']} sweeps / {rm['iterations']} inner "
              f"iterations ({rm['recon_steps']} reconstruction steps), angular store {ra['sweeps']} / "
              f"{ra['iterations']}; scalar flux difference {diff:.2e}", flush=True)
        assert rm["status"] == 1 and rm["sweeps"] <= pin, rm
        assert diff <= 1e-7, diff


def test_moments_store_refused_on_a_batch(gpu):
    import ctypes as C
    from ubolt_amd import capi
    ops = [load("plex_box_up3_layers_dg1").problem.operator()]
    bs = u.BatchSolver(u.Batch(ops), rtol=1e-10)
    G, n = ops[0].n_groups, bs.batch.n_rows
    o = capi.GroupOptions(1e-8, 100, 0.01, 0.0, 1, 10)
    psi = torch.zeros((G, n), dtype=torch.float64, device="cuda")
    rc = capi.lib().ubt_batch_solver_solve_groups(bs.h, C.byref(o), C.c_void_p(psi.data_ptr()), None)
    assert capi.STATUS.get(rc) == "ERR_SUP"
