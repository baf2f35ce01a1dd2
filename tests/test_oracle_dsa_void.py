"""Pins the oracle's DSA void policy (oracle/dsa.py; reference
src/dsak.kokkos.cxx:233-345) on the reference's void files:

* the voids' mean chord L = 4 V / S by hand: box_void_channel's void is
  3.5 x 0.6 with its left end on a vacuum face, so V = 2.1 and S = 0.6 + 0.6
  + 3.5 + 3.5 = 8.2; cube_void_duct's is 0.7 x 0.2 x 0.2, V = 0.028, S =
  0.04 + 0.04 + 4 x 0.14 = 0.64; a reflective face does not end a flight;
* D in a void (sigma_t = 0) is L / 3;
* the DG0 plex twin is V times the structured star with its voids too;
* a masked void column equals a box one column narrower with a vacuum face
  there (FD, DG0 and DG1: a face into a masked void is a vacuum face);
* the DG1 operator with its harmonic void faces stays symmetric positive
  definite, and a planted mistake (the void faces without the harmonic D)
  changes it, so the GPU matvec comparison pins that rule;
* a group void everywhere is masked whole, and voids that only masking
  could treat are refused.
"""
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spl

from helpers import oracle_dsa_matrix, xs_arrays
from oracle import dsa as odsa
from ubolt_amd import problem_file as pfm
from ubolt_amd import problems as P

PROBLEMS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "problems")


def load(name):
    return pfm.load(os.path.join(PROBLEMS, name + ".json")).problem


def coefficients(prob):
    d = prob.disc
    st, ss, _ = xs_arrays(prob)
    if d.kind == "fd":
        ft = d.get("bc_face").reshape(3, 2, 7)[:, :, 0].astype(int)
        vs = lambda v: odsa.structured_void_vs(d.dim, d.get("n"), d.get("h"), ft, v)
        any_vac = bool(np.any(ft[:d.dim] == 0))
    else:
        fnb = d.get("face_nb")
        vs = lambda v: odsa.plex_void_vs(d.nf, fnb, d.get("face_area"), d.get("volume"), v)
        any_vac = bool(np.any(fnb == -1))
    return odsa.diffusion_coefficients(st[0], ss[0, 0], vs, any_vac)[:3]


@pytest.mark.parametrize("name,V,S", [("box_void_channel", 2.1, 8.2), ("plex_box_void_channel", 2.1, 8.2),
                                      ("plex_box_void_channel_dg1", 2.1, 8.2), ("cube_void_duct", 0.028, 0.64)])
def test_chord_by_hand(name, V, S):
    prob = load(name)
    D, void, L = coefficients(prob)
    n_void = {"cube_void_duct": 7 * 2 * 2}.get(name, 35 * 6)
    assert int(void.sum()) == n_void
    assert abs(L - 4.0 * V / S) <= 1e-12 * L
    assert np.allclose(D[void], L / 3.0, rtol=1e-14, atol=0.0)
    assert np.allclose(D[~void], 1.0 / 300.0, rtol=1e-14, atol=0.0)


def test_reflective_face_does_not_end_a_flight():
    prob = load("box_void_channel")
    d = prob.disc
    ft = d.get("bc_face").reshape(3, 2, 7)[:, :, 0].astype(int).copy()
    void = ~(xs_arrays(prob)[0][0] > 0.0)
    V, S = odsa.structured_void_vs(d.dim, d.get("n"), d.get("h"), ft, void)
    ft[0, 0] = 1  # the left face, where the channel starts, reflective
    V2, S2 = odsa.structured_void_vs(d.dim, d.get("n"), d.get("h"), ft, void)
    assert V2 == V and abs(S - S2 - 0.6) <= 1e-12


def test_dg0_twin_is_volume_times_structured():
    Ms, _ = oracle_dsa_matrix(load("box_void_channel"))
    Md, _ = oracle_dsa_matrix(load("plex_box_void_channel"))
    V = 0.1 ** 2
    assert abs(Md - V * Ms).max() <= 1e-12 * abs(Md).max()


def test_dg1_void_faces_symmetric_positive_definite():
    prob = load("plex_box_void_channel_dg1")
    M, _ = oracle_dsa_matrix(prob)
    assert abs(M - M.T).max() <= 1e-12 * abs(M).max()
    lo = spl.eigsh(M.tocsc(), k=1, sigma=0.0, which="LM", return_eigenvectors=False)[0]
    assert lo > 0.0
    # planted: the void faces keep each side's own D
    st, ss, _ = xs_arrays(prob)
    d = prob.disc
    real = odsa.diffusion_coefficients

    def no_face_rule(*a, **k):
        D, void, L, masked = real(*a, **k)
        return D, void, None, masked  # the face rule sees no bridged voids
    odsa.diffusion_coefficients = no_face_rule
    try:
        Mw = odsa.dg1(d.nb, d.nf, d.get("face_nb"), d.get("face_nA"), d.get("face_area"), d.get("face_distance"),
                      d.get("volume"), d.get("face_own"), d.get("face_up"), d.get("basis_grad"), st[0], ss[0, 0])
    finally:
        odsa.diffusion_coefficients = real
    assert abs(Mw - M).max() > 1e-6 * abs(M).max()


def test_all_void_and_unbridgeable():
    prob = P.fd_box_2d_central_source(n=6, sn=2)
    n = prob.disc.n_cells
    with pytest.raises(ValueError):
        oracle_dsa_matrix(P.Problem("void", prob.quad, prob.disc, np.zeros(n), np.zeros(n), None))
    # every face reflective, no absorption: the void falls back to the mask
    pf = load("box_50_inf_medium")
    d = pf.disc
    st = np.ones(d.n_cells)
    st[7] = 0.0
    M, wgt = oracle_dsa_matrix(P.Problem("refl", pf.quad, d, st, st.copy(), None))
    assert wgt[7] == 0.0 and wgt.sum() == d.n_cells - 1
    assert M[7, 7] == 1.0 and M.getrow(7).nnz == 1 and M.getcol(7).nnz == 1


def masked_strip_vs_narrower_box(kind):
    """a box whose first column is a masked void against the same box one
    column narrower with a vacuum face there: the non-void block must equal
    the narrower box's operator (a face into a masked void is a vacuum face)"""
    from ubolt_amd import core as uc
    n, q = 6, uc.Quadrature(2, 4)
    h = 1.0 / n
    rng = np.random.default_rng(7)
    st_in = rng.uniform(0.5, 3.0, (n, n - 1))
    st = np.concatenate([np.zeros((n, 1)), st_in], axis=1).ravel()
    def disc(nx):
        if kind == "fd":
            return uc.Disc.structured(q, [nx, n], [nx * h, 1.0], None)
        coords, cells = P.box_quads(nx, n, lx=nx * h, ly=1.0)
        return uc.Disc.dg(q, uc.Mesh(coords, cells), 0 if kind == "dg0" else 1, None)
    Mw, wgt = oracle_dsa_matrix(P.Problem("w", q, disc(n), st, 0.5 * st, None), bridge=False)
    Mn, _ = oracle_dsa_matrix(P.Problem("n", q, disc(n - 1), st_in.ravel(), 0.5 * st_in.ravel(), None))
    return Mw, wgt, Mn, n


@pytest.mark.parametrize("kind", ["fd", "dg0", "dg1"])
def test_masked_strip_is_a_vacuum_face(kind):
    Mw, wgt, Mn, n = masked_strip_vs_narrower_box(kind)
    keep = np.flatnonzero(wgt > 0)
    nb = 3 if kind == "dg1" else 1
    assert keep.size == Mn.shape[0] == n * (n - 1) * nb
    sub = Mw.tocsr()[keep][:, keep]
    assert abs(sub - Mn).max() <= 1e-12 * abs(Mn).max()
    drop = np.flatnonzero(wgt == 0)
    assert abs(Mw.tocsr()[drop][:, keep]).max() == 0.0
