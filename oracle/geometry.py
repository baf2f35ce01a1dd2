"""DG0/DG1 geometry restated in numpy - TEST INFRASTRUCTURE ONLY.

Restates the reference's cell and face geometry and DG1 basis
(src/unstructured_dgk.kokkos.cxx:29-218 moments and orthonormal gradients,
:706-821 face matrices) from vertex coordinates, deliberately by different
decompositions than the fans the product uses, so the two only agree if both
are exact:

* 2D cells: polygon area, centroid and second moments by Green's theorem
  (the shoelace sums);
* 3D cells: the hexahedron split into 6 tetrahedra along its 0-6 diagonal,
  tetrahedra as themselves;
* faces: a 2D edge's second moment about its midpoint is (L/12) e e^T; a 3D
  quad face is split along its 0-2 diagonal into two triangles.

Basis: phi_0 = 1, phi_{1+k} = beta_k . (x - x_c) with (1/V) int phi_i phi_j =
delta_ij, beta = L^{-1} for C / V = L L^T. Face matrices
F_own(i,j) = (A phi_i(x_f) phi_j(x_f) + b_i^T C_f b_j) / (V A).
"""
import numpy as np

TRI_FACES = [(0, 1), (1, 2), (2, 0)]
QUAD_FACES = [(0, 1), (1, 2), (2, 3), (3, 0)]
TET_FACES = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]
HEX_FACES = [(0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)]
HEX_TETS = [(0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)]


def simplex_moments(P):
    """Volume (area), first moment about the origin and second moment
    int x x^T of a simplex with vertex rows P (d+1 x 3), dimension d = rows-1."""
    d = P.shape[0] - 1
    E = (P[1:] - P[0])
    if d == 1:
        meas = np.linalg.norm(E[0])
    elif d == 2:
        meas = 0.5 * np.linalg.norm(np.cross(E[0], E[1]))
    else:
        meas = abs(np.linalg.det(E)) / 6.0
    s = P.sum(axis=0)
    m1 = meas * s / (d + 1)
    m2 = meas / ((d + 1) * (d + 2)) * (P.T @ P + np.outer(s, s))
    return meas, m1, m2


def polygon_moments_2d(X):
    """Area, centroid and central second moment of a simple polygon (Green),
    about the vertex average so the central moment does not cancel."""
    o = X.mean(axis=0)
    X = X - o
    x, y = X[:, 0], X[:, 1]
    xn, yn = np.roll(x, -1), np.roll(y, -1)
    cr = x * yn - xn * y
    A = 0.5 * cr.sum()
    sgn = np.sign(A)
    A = abs(A)
    cx = sgn * ((x + xn) * cr).sum() / 6.0 / A
    cy = sgn * ((y + yn) * cr).sum() / 6.0 / A
    Ixx = sgn * ((x * x + x * xn + xn * xn) * cr).sum() / 12.0
    Iyy = sgn * ((y * y + y * yn + yn * yn) * cr).sum() / 12.0
    Ixy = sgn * ((x * yn + 2 * x * y + 2 * xn * yn + xn * y) * cr).sum() / 24.0
    C = np.zeros((3, 3))
    C[0, 0] = Ixx - A * cx * cx
    C[1, 1] = Iyy - A * cy * cy
    C[0, 1] = C[1, 0] = Ixy - A * cx * cy
    return A, np.array([cx + o[0], cy + o[1], 0.0]), C


def cell_moments(X, dim):
    if dim == 2:
        return polygon_moments_2d(X)
    tets = HEX_TETS if X.shape[0] == 8 else [(0, 1, 2, 3)]
    o = X.mean(axis=0)
    X = X - o
    V, m1, m2 = 0.0, np.zeros(3), np.zeros((3, 3))
    for t in tets:
        v, a, b = simplex_moments(X[list(t)])
        V, m1, m2 = V + v, m1 + a, m2 + b
    c = m1 / V
    return V, c + o, m2 - V * np.outer(c, c)


def face_moments(X, dim):
    """Area, centroid, central second moment and (unoriented) unit normal."""
    if dim == 2:
        e = X[1] - X[0]
        L = np.linalg.norm(e)
        n = np.array([e[1], -e[0], 0.0]) / L
        return L, 0.5 * (X[0] + X[1]), np.outer(e, e) * L / 12.0, n
    tris = [(0, 1, 2)] if X.shape[0] == 3 else [(0, 1, 2), (0, 2, 3)]
    o = X.mean(axis=0)
    Y = X - o
    A, m1, m2 = 0.0, np.zeros(3), np.zeros((3, 3))
    for t in tris:
        a, b, c = simplex_moments(Y[list(t)])
        A, m1, m2 = A + a, m1 + b, m2 + c
    xf = m1 / A
    n = np.cross(X[1] - X[0], X[2] - X[0])
    return A, xf + o, m2 - A * np.outer(xf, xf), n / np.linalg.norm(n)


def local_faces(dim, nv):
    return {(2, 3): TRI_FACES, (2, 4): QUAD_FACES, (3, 4): TET_FACES, (3, 8): HEX_FACES}[(dim, nv)]


def dg_geometry(coords, cells, order):
    """Per-cell and per-(cell, face) tables in the product's slot order
    (cell * nf + local face, local faces as in local_faces). Returns a dict of
    volume, centroid, face_nA, face_area, face_centroid, face_neighbour and, at
    order 1, basis_grad, face_basis_value, face_own, face_up (face_up for a
    neighbour only; boundary faces zero)."""
    coords = np.asarray(coords, dtype=np.float64)
    dim = coords.shape[1]
    X3 = np.zeros((coords.shape[0], 3))
    X3[:, :dim] = coords
    nc, nv = cells.shape
    F = local_faces(dim, nv)
    nf = len(F)
    keys = {}
    nbr = -np.ones(nc * nf, dtype=np.int64)
    for c in range(nc):
        for f, fv in enumerate(F):
            key = tuple(sorted(int(cells[c, i]) for i in fv))
            if key in keys:
                o = keys[key]
                nbr[c * nf + f] = o // nf
                nbr[o] = c
            else:
                keys[key] = c * nf + f
    vol = np.zeros(nc)
    cen = np.zeros((nc, 3))
    cmom = np.zeros((nc, 3, 3))
    for c in range(nc):
        vol[c], cen[c], cmom[c] = cell_moments(X3[cells[c]], dim)
    out = {"volume": vol, "centroid": cen, "face_neighbour": nbr}
    nA = np.zeros((nc * nf, 3))
    area = np.zeros(nc * nf)
    fcen = np.zeros((nc * nf, 3))
    fC = np.zeros((nc * nf, 3, 3))
    for c in range(nc):
        for f, fv in enumerate(F):
            k = c * nf + f
            A, xf, C, n = face_moments(X3[cells[c, list(fv)]], dim)
            if np.dot(n, xf - cen[c]) < 0:
                n = -n
            nA[k], area[k], fcen[k], fC[k] = A * n, A, xf, C
    out.update(face_nA=nA, face_area=area, face_centroid=fcen)
    if order == 0:
        return out
    nb = dim + 1
    beta = np.zeros((nc, 3, 3))
    for c in range(nc):
        L = np.linalg.cholesky(cmom[c, :dim, :dim] / vol[c])
        beta[c, :dim, :dim] = np.linalg.inv(L)

    def val(c, i, x):
        return 1.0 if i == 0 else float(beta[c, i - 1] @ (x - cen[c]))

    def grad(c, i):
        return np.zeros(3) if i == 0 else beta[c, i - 1]

    bg = np.zeros((nc, nb, 3))
    fbv = np.zeros((nc * nf, nb))
    own = np.zeros((nc * nf, nb, nb))
    up = np.zeros((nc * nf, nb, nb))
    for c in range(nc):
        for i in range(nb):
            bg[c, i] = grad(c, i)
        for f in range(nf):
            k = c * nf + f
            s = 1.0 / (vol[c] * area[k])
            for i in range(nb):
                fbv[k, i] = val(c, i, fcen[k])
                for j in range(nb):
                    own[k, i, j] = s * (area[k] * val(c, i, fcen[k]) * val(c, j, fcen[k]) +
                                        grad(c, i) @ fC[k] @ grad(c, j))
                    if nbr[k] >= 0:
                        n = nbr[k]
                        up[k, i, j] = s * (area[k] * val(c, i, fcen[k]) * val(n, j, fcen[k]) +
                                           grad(c, i) @ fC[k] @ grad(n, j))
    out.update(basis_grad=bg, face_basis_value=fbv, face_own=own, face_up=up)
    return out
