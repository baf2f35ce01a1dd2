"""DSA diffusion operators restated with scipy - TEST INFRASTRUCTURE ONLY.

The reference's three DSA operators, assembled as sparse matrices cell by cell
from the same formulas, and the DSA apply y = P D^{-1} R x with an exact inner
solve:

* structured star, per unit volume (src/dsa_structuredk.kokkos.cxx:183-285):
  harmonic face D, consistent-D blend with the axis half-range current,
  Marshak vacuum faces, zero Neumann reflective faces;
* DG0 two-point flux, volume-weighted (src/dsa_plexk.kokkos.cxx:288-348);
* DG1 MIP interior penalty, volume-weighted (src/dsa_plexk.kokkos.cxx:365-475);
helpers: src/dsa_operatork.hpp:110-147 (half-range current, blend).
D = 1/(3 sigma_t), sigma_a = sigma_t - sigma_s(g,g). Voids (sigma_t <=
void_sigma_t) follow src/dsak.kokkos.cxx:233-345: bridged (-dsa_void_bridge,
the default) with D = void_d or 1/(3 (max(sigma_t, 0) + 1/L)), L = 4 V / S, V
the voids' volume and S their faces onto a non-void cell or a vacuum face
(src/dsa_structuredk.kokkos.cxx:127-170, src/dsa_plexk.kokkos.cxx:243-262),
when something is not void, a vacuum face or some absorption holds the
operator and S > 0, a DG1 face touching a void taking the harmonic face D on
both sides (src/dsa_plexk.kokkos.cxx:433-441); masked otherwise: identity
rows (V times identity blocks on the plex), a face into one a vacuum face
(src/dsa_structuredk.kokkos.cxx:212-268, src/dsa_plexk.kokkos.cxx:296-317,
381-417), and a zero restriction / prolongation weight there.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl


def structured_void_vs(dim, n, h, face_type, void):
    """(V, S) of the void cells of a structured box"""
    n = list(n[:dim])
    vol = float(np.prod(h[:dim]))
    strides = [1, n[0], n[0] * (n[1] if dim > 1 else 1)]
    V = S = 0.0
    for c in np.flatnonzero(void):
        idx, cc = [], int(c)
        for d in range(dim):
            idx.append(cc % n[d])
            cc //= n[d]
        V += vol
        for d in range(dim):
            for side, step in ((0, -1), (1, 1)):
                m = idx[d] + step
                ends = (not void[c + step * strides[d]]) if 0 <= m < n[d] else face_type[d][side] == 0
                if ends:
                    S += vol / h[d]
    return V, S


def plex_void_vs(nf, face_nb, area, volume, void):
    """(V, S) of the void cells of a mesh (face_nb: >= 0 a cell, -1 vacuum, < -1 reflective)"""
    V = S = 0.0
    for c in np.flatnonzero(void):
        V += volume[c]
        for f in range(nf):
            k = c * nf + f
            n = face_nb[k]
            if (n >= 0 and not void[n]) or n == -1:
                S += area[k]
    return V, S


def diffusion_coefficients(sigt, sigs, vs, any_vacuum, bridge=True, void_sigma_t=0.0, void_d=0.0):
    """D per cell, the void cells, the mean chord L (None unless bridged) and
    whether the voids are masked"""
    sigt = np.asarray(sigt, dtype=np.float64)
    void = ~(sigt > void_sigma_t)
    D = np.full(len(sigt), np.nan)
    D[~void] = 1.0 / (3.0 * sigt[~void])
    if not void.any():
        return D, void, None, False
    if void.all():
        raise ValueError("void everywhere: the operator is masked whole")
    if bridge and (any_vacuum or np.any(sigt - sigs > 0.0)):
        V, S = vs(void)
        if S > 0.0:
            L = 4.0 * V / S
            D[void] = void_d if void_d > 0.0 else 1.0 / (3.0 * (np.maximum(sigt[void], 0.0) + 1.0 / L))
            return D, void, L, False
    return D, void, None, True


def half_range(omega, w, W, n):
    c = omega @ n
    return float((w * c)[c > 0].sum() / W)


def blend(d, m, p):
    return d + m if p == 1.0 else (d ** p + m ** p) ** (1.0 / p)


def structured(dim, n, h, face_type, omega, w, W, sigt, sigs, cons=True, p=1.5, **void):
    """face_type[d][side]: 0 vacuum, 1 reflect; void: the void policy's options"""
    n = list(n[:dim])
    cells = int(np.prod(n))
    hr = [half_range(omega, w, W, np.eye(3)[d]) for d in range(dim)]
    rows, cols, vals = [], [], []
    strides = [1, n[0], n[0] * (n[1] if dim > 1 else 1)]
    any_vac = any(face_type[d][s] == 0 for d in range(dim) for s in range(2))
    D, vd, _, masked = diffusion_coefficients(sigt, sigs, lambda v: structured_void_vs(dim, n, h, face_type, v),
                                              any_vac, **void)
    mk = vd if masked else np.zeros(cells, dtype=bool)
    for c in range(cells):
        if mk[c]:
            rows.append(c), cols.append(c), vals.append(1.0)
            continue
        idx, cc = [], c
        for d in range(dim):
            idx.append(cc % n[d])
            cc //= n[d]
        dc = D[c]
        diag = sigt[c] - sigs[c]
        for d in range(dim):
            hd = h[d]
            for side, step in ((0, -1), (1, 1)):
                m = idx[d] + step
                if 0 <= m < n[d] and not mk[c + step * strides[d]]:
                    nb = c + step * strides[d]
                    dn = D[nb]
                    df = 2.0 * dc * dn / (dc + dn)
                    if cons:
                        df = blend(df, hr[d] * hd, p)
                    rows.append(c), cols.append(nb), vals.append(-df / hd ** 2)
                    diag += df / hd ** 2
                elif 0 <= m < n[d] or face_type[d][side] == 0:
                    db = blend(dc, hr[d] * hd, p) if cons else dc
                    diag += 1.0 / (hd * (2.0 + hd / (2.0 * db)))
        rows.append(c), cols.append(c), vals.append(diag)
    return sp.csr_matrix((vals, (rows, cols)), shape=(cells, cells))


def dg0(nf, face_nb, face_nA, area, dist, volume, omega, w, W, sigt, sigs, cons=True, p=1.5, **void):
    cells = volume.shape[0]
    nA = face_nA.reshape(-1, 3)
    dist = dist.reshape(-1, 2)
    rows, cols, vals = [], [], []
    D, vd, _, masked = diffusion_coefficients(sigt, sigs, lambda v: plex_void_vs(nf, face_nb, area, volume, v),
                                              bool(np.any(np.asarray(face_nb) == -1)), **void)
    mk = vd if masked else np.zeros(cells, dtype=bool)
    for c in range(cells):
        if mk[c]:
            rows.append(c), cols.append(c), vals.append(volume[c])
            continue
        dc = D[c]
        diag = volume[c] * (sigt[c] - sigs[c])
        for f in range(nf):
            k = c * nf + f
            A = area[k]
            m = half_range(omega, w, W, nA[k] / A)
            d_own = dist[k, 0]
            if face_nb[k] >= 0 and not mk[face_nb[k]]:
                nb = face_nb[k]
                dn = D[nb]
                tf = A / (d_own / dc + dist[k, 1] / dn)
                if cons:
                    L = d_own + dist[k, 1]
                    tf = A / L * blend(tf * L / A, m * L, p)
                rows.append(c), cols.append(nb), vals.append(-tf)
                diag += tf
            elif face_nb[k] >= -1:
                db = blend(dc, 2.0 * d_own * m, p) if cons else dc
                diag += A / (2.0 + d_own / db)
        rows.append(c), cols.append(c), vals.append(diag)
    return sp.csr_matrix((vals, (rows, cols)), shape=(cells, cells))


def dg1(nb, nf, face_nb, face_nA, area, dist, volume, face_own, face_up, basis_grad, sigt, sigs, C=4.0, **void):
    cells = volume.shape[0]
    nA = face_nA.reshape(-1, 3)
    dist = dist.reshape(-1, 2)
    Fo = face_own.reshape(-1, nb, nb)
    Fu = face_up.reshape(-1, nb, nb)
    G = basis_grad.reshape(cells, nb, 3)
    M = sp.lil_matrix((cells * nb, cells * nb))
    D, vd, L, masked = diffusion_coefficients(sigt, sigs, lambda v: plex_void_vs(nf, face_nb, area, volume, v),
                                              bool(np.any(np.asarray(face_nb) == -1)), **void)
    mk = vd if masked else np.zeros(cells, dtype=bool)
    for c in range(cells):
        if mk[c]:
            M[c * nb:(c + 1) * nb, c * nb:(c + 1) * nb] = volume[c] * np.eye(nb)
            continue
        dc = D[c]
        sa = sigt[c] - sigs[c]
        V = volume[c]
        own = V * (dc * (G[c] @ G[c].T) + sa * np.eye(nb))
        for f in range(nf):
            k = c * nf + f
            n = face_nb[k]
            if n < -1:
                continue
            A = area[k]
            s = V * A
            nrm = nA[k] / A
            gc = G[c] @ nrm
            hc = 2.0 * dist[k, 0]
            m = s * Fo[k][:, 0]  # int_f phi_i
            if n >= 0 and not mk[n]:
                dcf, dn = dc, D[n]
                if L is not None and (vd[c] or vd[n]):
                    dcf = dn = 2.0 * dc * D[n] / (dc + D[n])
                hn = 2.0 * dist[k, 1]
                gn = G[n] @ nrm
                kap = max(0.5 * C * (dcf / hc + dn / hn), 0.25)
                own += kap * s * Fo[k] - 0.5 * dcf * np.outer(m, gc) - 0.5 * dcf * np.outer(gc, s * Fo[k][0, :])
                blk = -kap * s * Fu[k] - 0.5 * dn * np.outer(m, gn) + 0.5 * dcf * np.outer(gc, s * Fu[k][0, :])
                M[c * nb:(c + 1) * nb, n * nb:(n + 1) * nb] = M[c * nb:(c + 1) * nb, n * nb:(n + 1) * nb] + blk
            else:
                kap = max(C * dc / hc, 0.25)
                own += kap * s * Fo[k] - 0.5 * dc * np.outer(m, gc) - 0.5 * dc * np.outer(gc, s * Fo[k][0, :])
        M[c * nb:(c + 1) * nb, c * nb:(c + 1) * nb] = own
    return M.tocsr()


def apply(Dmat, weight, A, w, W, x):
    """y = P D^{-1} R x, exact inner solve"""
    phi = x.reshape(-1, A) @ w
    sol = spl.spsolve(Dmat.tocsc(), phi * weight)
    return np.repeat(sol / W, A)
