"""Host DG geometry against the numpy restatement (oracle/geometry.py, built by
different exact decompositions) and geometric invariants (reference
tests/verify_plexk.kokkos.cxx:579 geometry invariants; DG1 tables
src/unstructured_dgk.kokkos.cxx:706-821)."""
import numpy as np
import pytest

import ubolt_amd as u
from oracle import geometry as G
from ubolt_amd import problems as P


def skewed_quads(n, seed=3):
    coords, cells = P.box_quads(n, n)
    rng = np.random.default_rng(seed)
    inner = (coords[:, 0] > 1e-12) & (coords[:, 0] < 1 - 1e-12) & (coords[:, 1] > 1e-12) & (coords[:, 1] < 1 - 1e-12)
    coords = coords.copy()
    coords[inner] += rng.uniform(-0.2, 0.2, (inner.sum(), 2)) / n
    return coords, cells


def tets():
    coords, hexes = P.box_hexes(2, 3, 2)
    return coords, np.concatenate([hexes[:, list(t)] for t in G.HEX_TETS])


MESHES = {
    "tri_box": lambda: P.box_tris(4, 3),
    "quad_box": lambda: P.box_quads(3, 5, 2.0, 1.0),
    "quad_skew": lambda: skewed_quads(5),
    "delaunay": lambda: P.delaunay_square(6, seed=9),
    "hex_box": lambda: P.box_hexes(3, 2, 2, 1.0, 2.0, 1.5),
    "tet_box": tets,
}


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("order", [0, 1])
def test_tables_match_restatement(name, order):
    coords, cells = MESHES[name]()
    dim = coords.shape[1]
    q = u.Quadrature(dim, 4)
    d = u.Disc.dg(q, u.Mesh(coords, cells), order)
    ref = G.dg_geometry(coords, cells, order)
    nf = d.nf
    assert np.allclose(d.get("volume"), ref["volume"], rtol=1e-12, atol=0)
    assert np.allclose(d.get("centroid").reshape(-1, 3), ref["centroid"], rtol=0, atol=1e-13)
    assert np.allclose(d.get("face_nA").reshape(-1, 3), ref["face_nA"], rtol=0, atol=1e-13)
    assert np.allclose(d.get("face_area"), ref["face_area"], rtol=1e-12, atol=0)
    assert np.allclose(d.get("face_centroid").reshape(-1, 3), ref["face_centroid"], rtol=0, atol=1e-13)
    fnb = d.get("face_nb")
    assert np.array_equal(np.where(fnb >= 0, fnb, -1), ref["face_neighbour"])
    if order == 1:
        nb = dim + 1
        gscale = np.abs(ref["basis_grad"]).max()
        assert np.allclose(d.get("basis_grad").reshape(-1, nb, 3), ref["basis_grad"], rtol=0, atol=1e-11 * gscale)
        scale = np.abs(ref["face_own"]).max()
        assert np.allclose(d.get("face_own").reshape(-1, nb, nb), ref["face_own"], rtol=0, atol=1e-12 * scale)
        up = d.get("face_up").reshape(-1, nb, nb)
        interior = fnb >= 0
        assert np.allclose(up[interior], ref["face_up"][interior], rtol=0, atol=1e-12 * scale)
        assert np.allclose(d.get("face_basis_value").reshape(-1, nb), ref["face_basis_value"], rtol=0, atol=1e-12)


@pytest.mark.parametrize("name", sorted(MESHES))
def test_invariants(name):
    coords, cells = MESHES[name]()
    dim = coords.shape[1]
    d = u.Disc.dg(u.Quadrature(dim, 2), u.Mesh(coords, cells), 1)
    nA = d.get("face_nA").reshape(d.n_cells, d.nf, 3)
    # closed cells: the outward area normals sum to zero
    assert np.abs(nA.sum(axis=1)).max() < 1e-13
    # the cells tile the bounding box
    ext = coords.max(axis=0) - coords.min(axis=0)
    assert abs(d.get("volume").sum() - np.prod(ext)) < 1e-12 * np.prod(ext)
    # interior faces are seen oppositely from their two cells
    fnb = d.get("face_nb")
    cen = d.get("face_centroid").reshape(-1, 3)
    for k in np.nonzero(fnb >= 0)[0][:200]:
        c, nbr = k // d.nf, fnb[k]
        kk = [nbr * d.nf + f for f in range(d.nf) if np.allclose(cen[nbr * d.nf + f], cen[k], atol=1e-13)]
        assert len(kk) == 1 and np.allclose(nA[c, k % d.nf], -nA.reshape(-1, 3)[kk[0]], atol=1e-13)
    # the DG1 face matrices: F_own(0,0) = 1/V (phi_0 = 1), F_own(i,0) = (1/(V A)) int_f phi_i
    own = d.get("face_own").reshape(d.n_cells, d.nf, dim + 1, dim + 1)
    assert np.allclose(own[:, :, 0, 0], 1.0 / d.get("volume")[:, None], rtol=1e-12)
    assert np.allclose(own, np.swapaxes(own, 2, 3), atol=1e-12 * np.abs(own).max())
    # distances: centroid to face plane and face plane to the neighbour's centroid, both positive
    dist = d.get("face_distance").reshape(-1, 2)
    assert (dist[:, 0] > 0).all() and (dist[fnb >= 0, 1] > 0).all()


def test_box_labels_and_reflect_check():
    coords, cells = P.box_quads(3, 3)
    m = u.Mesh(coords, cells)
    d = u.Disc.dg(u.Quadrature(2, 2), m, 0, {1: "reflect", 3: {"type": "vacuum", "inflow": 2.0, "window": [0.0, 0.5]}})
    lab = d.get("face_label")
    fnb = d.get("face_nb")
    assert set(lab[fnb < 0]) == {1, 2, 3, 4}
    # reflective bottom: the y axis (1) is encoded as -2 - 1
    assert (fnb[lab == 1] == -3).all()
    # windowed inflow on the top face: only face centroids with x in [0, 0.5]
    inflow = d.get("face_inflow")
    fc = d.get("face_centroid").reshape(-1, 3)
    top = lab == 3
    assert np.allclose(inflow[top], np.where(fc[top, 0] <= 0.5, 2.0 / (4 * np.pi), 0.0))


def test_errors():
    coords, cells = P.box_quads(2, 2)
    with pytest.raises(u.UbtError) as e:
        u.Mesh(coords, cells[:, :2])
    assert e.value.status == "ERR_SUP"
    m = u.Mesh(coords, cells)
    with pytest.raises(u.UbtError) as e:
        u.Disc.dg(u.Quadrature(2, 2), m, 2)
    assert e.value.status == "ERR_ARG_OUTOFRANGE"
    with pytest.raises(u.UbtError) as e:
        u.Disc.dg(u.Quadrature(3, 2), m, 1)
    assert e.value.status == "ERR_ARG_INCOMP"
    with pytest.raises(u.UbtError) as e:
        u.Disc.dg(u.Quadrature(2, 2), m, 1, {9: "vacuum"})
    assert e.value.status == "ERR_ARG_WRONG"
    with pytest.raises(u.UbtError) as e:  # a skewed reflective face
        c2 = coords.copy()
        c2[2, 1] += 0.1  # the bottom-right corner lifted: face (1, 2) is no longer axis-aligned
        m2 = u.Mesh(c2, cells)
        m2.set_boundary_labels([[1, 2]], [1])
        u.Disc.dg(u.Quadrature(2, 2), m2, 1, {1: "reflect"})
    assert e.value.status == "ERR_ARG_WRONG"
    with pytest.raises(u.UbtError) as e:
        u.Disc.structured(u.Quadrature(2, 2), [4, 4], [1, 1], {7: "vacuum"})
    assert e.value.status == "ERR_ARG_WRONG"
