"""The allocation-free DSA inner solve (ubt_dsa_solve_noalloc) through the C
ABI against the allocating one (ubt_dsa_solve) on the same handle: two calls
in a row, the second on a smaller operator, each equal to the allocating
solve's result; and the size limit refused with a status code."""
import ctypes as C

import numpy as np
import pytest
import torch

import ubolt_amd as u
from helpers import rel_max, to_dev, to_host
from ubolt_amd import capi
from ubolt_amd import core
from ubolt_amd import problems as P

pytestmark = pytest.mark.gpu


def _dsa(prob):
    op = prob.operator()
    dsa = u.DSA(op, inner_max_it=300, inner_rtol=1e-10)
    dsa.set_group(0)
    return op, dsa


def _noalloc(dsa, b, x):
    rc = capi.lib().ubt_dsa_solve_noalloc(dsa.h, C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()),
                                          capi.stream_ptr())
    return rc


def test_noalloc_matches_allocating_solve(gpu):
    cases = [("dg1_delaunay_40", P.dg1_delaunay(n_side=40, sn=4, seed=3)),   # ~3.2k triangles, 9.6k nodes
             ("fd2d_12", P.fd_box_2d_central_source(n=12, sn=4))]            # then a smaller one: 144 nodes
    for name, prob in cases:
        op, dsa = _dsa(prob)
        _, nodes = dsa.info()
        b = to_dev(np.random.default_rng(nodes).standard_normal(nodes))
        want = torch.zeros(nodes, dtype=torch.float64, device="cuda")
        dsa.solve(b, want)
        got = torch.full((nodes,), float("nan"), dtype=torch.float64, device="cuda")
        print(f"[case] noalloc {name}: {nodes} nodes", flush=True)
        assert _noalloc(dsa, b, got) == 0, capi.last_error()
        torch.cuda.synchronize()
        wh, gh = to_host(want), to_host(got)
        err = rel_max(gh, wh)
        print(f"[case] noalloc {name}: rel err vs ubt_dsa_solve {err:.2e}, bitwise {np.array_equal(gh, wh)}",
              flush=True)
        assert np.all(np.isfinite(gh)) and err <= 1e-14


def test_noalloc_size_limit(gpu):
    """a structured box of 129 x 128 x 128 cells has 2^21 + 16384 diffusion nodes"""
    q = core.Quadrature(3, 2)
    d = core.Disc.structured(q, [129, 128, 128], [1.0, 1.0, 1.0], P.VACUUM_3D)
    nc = d.n_cells
    prob = P.Problem("big", q, d, np.ones(nc), np.full(nc, 0.5), None)
    op, dsa = _dsa(prob)
    _, nodes = dsa.info()
    assert nodes == (1 << 21) + 16384
    b = torch.ones(nodes, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(b)
    rc = _noalloc(dsa, b, x)
    code, msg = capi.last_error()
    print(f"[case] noalloc size limit: rc {rc} {msg}", flush=True)
    assert rc == 60 and "at most 2097152" in msg
