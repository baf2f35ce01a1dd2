"""The oracle - TEST INFRASTRUCTURE ONLY.

A CPU restatement of the reference (stevendargaville/UBOLT) algorithms on the
transport operator path, used by tests/, __graft_entry__.smoke() and the
bench's CPU-baseline leg as the checker. Never imported by the product package.

* ubt_oracle.c (built to liboracle.so by oracle/Makefile): FP64 C with OpenMP -
  explicit assembly of the streaming + removal rows in the reference's slot
  layout, the shell mult (SpMV + matrix-free scatter), moments, transfer /
  source / ghost-inflow rhs, element blocks read off the assembled rows and
  inverted by Gauss-Jordan.
* geometry.py: numpy restatement of the DG0/DG1 geometry (fan moments,
  orthonormal basis, face matrices).
* dsa.py / krylov.py: the DSA diffusion operators assembled as scipy sparse
  matrices and the right-preconditioned GMRES.

Every function takes plain numpy arrays in the reference layouts; the caller
extracts them from whatever it is checking.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

i64 = C.c_int64
PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)
PL = C.POINTER(C.c_int64)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError("oracle/liboracle.so is not built; run `make -C oracle`")
        L = C.CDLL(_LIB)
        L.oracle_threads.restype = C.c_int
        L.oracle_block_setup.restype = C.c_int64
        _lib = L
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _pd(a):
    return a.ctypes.data_as(PD)


def _pi(a):
    return a.ctypes.data_as(PI)


def _pl(a):
    return a.ctypes.data_as(PL)


def threads():
    return lib().oracle_threads()


def set_threads(n):
    lib().oracle_set_threads(int(n))


class Rows:
    """Assembled streaming + removal rows: S slots per row, col -1 = nulled."""

    def __init__(self, n_rows, S):
        self.n_rows, self.S = n_rows, S
        self.col = np.full(n_rows * S, -1, dtype=np.int64)
        self.val = np.zeros(n_rows * S)

    def to_scipy(self):
        import scipy.sparse as sp
        r = np.repeat(np.arange(self.n_rows), self.S)
        m = self.col >= 0
        return sp.csr_matrix((self.val[m], (r[m], self.col[m])), shape=(self.n_rows, self.n_rows))


def fd_assemble(dim, n, h, omega, reflect, face_type, sigt):
    """face_type: [dim][2] (0 vacuum, 1 reflect); omega [A,3]; reflect [A,3]"""
    A = omega.shape[0]
    cells = int(np.prod(n[:dim]))
    R = Rows(cells * A, dim + 1)
    ft = _i(np.asarray(face_type).reshape(-1)[: 2 * dim].tolist() + [0] * (6 - 2 * dim))
    lib().oracle_fd_assemble(dim, _pi(_i(n)), _pd(_d(h)), A, _pd(_d(omega)), _pi(_i(reflect)), _pi(ft), _pd(_d(sigt)),
                             _pl(R.col), _pd(R.val))
    return R


def fd_ghost_inflow(dim, n, h, omega, face_bc, b):
    A = omega.shape[0]
    lib().oracle_fd_ghost_inflow(dim, _pi(_i(n)), _pd(_d(h)), A, _pd(_d(omega)), _pd(_d(face_bc)), _pd(b))
    return b


def dg_assemble(nb, nf, omega, reflect, face_nb, face_nA, volume, face_own, face_up, basis_grad, sigt):
    A = omega.shape[0]
    cells = volume.shape[0]
    R = Rows(cells * nb * A, (nf + 1) * nb)
    empty = np.zeros(1)
    lib().oracle_dg_assemble(nb, nf, i64(cells), A, _pd(_d(omega)), _pi(_i(reflect)), _pi(_i(face_nb)),
                             _pd(_d(face_nA)), _pd(_d(volume)),
                             _pd(_d(face_own) if face_own is not None and len(face_own) else empty),
                             _pd(_d(face_up) if face_up is not None and len(face_up) else empty),
                             _pd(_d(basis_grad) if basis_grad is not None and len(basis_grad) else empty),
                             _pd(_d(sigt)), _pl(R.col), _pd(R.val))
    return R


def dg_ghost_inflow(nb, nf, omega, face_nb, face_nA, volume, face_basis_value, face_inflow, b):
    A = omega.shape[0]
    lib().oracle_dg_ghost_inflow(nb, nf, i64(volume.shape[0]), A, _pd(_d(omega)), _pi(_i(face_nb)), _pd(_d(face_nA)),
                                 _pd(_d(volume)), _pd(_d(face_basis_value)), _pd(_d(face_inflow)), _pd(b))
    return b


def rows_mult(R, x):
    y = np.zeros(R.n_rows)
    lib().oracle_rows_mult(i64(R.n_rows), R.S, _pl(R.col), _pd(R.val), _pd(_d(x)), _pd(y))
    return y


def moments(psi, w):
    A = w.shape[0]
    nodes = psi.size // A
    phi = np.zeros(nodes)
    lib().oracle_moments(i64(nodes), A, _pd(_d(w)), _pd(_d(psi)), _pd(phi))
    return phi


def scatter_add(nb, w, sum_w, sigs, x, y):
    A = w.shape[0]
    lib().oracle_scatter_add(i64(x.size // A), nb, A, _pd(_d(w)), C.c_double(sum_w), _pd(_d(sigs)), _pd(_d(x)), _pd(y))
    return y


def broadcast_add(nb, A, scale, phi, b):
    lib().oracle_broadcast_add(i64(phi.size), nb, A, _pd(_d(scale)), _pd(_d(phi)), _pd(b))
    return b


def shell_mult(R, nb, w, sum_w, sigs, x, scatter=True):
    """y = (assembled rows) x - sigma_s/W phi, the reference ShellMatMultApply"""
    y = rows_mult(R, x)
    if scatter:
        scatter_add(nb, w, sum_w, sigs, _d(x), y)
    return y


def block_setup(R, nb, cells, A):
    inv = np.zeros(cells * A * nb * nb)
    n_sing = lib().oracle_block_setup(nb, i64(cells), A, R.S, _pl(R.col), _pd(R.val), _pd(inv))
    return inv, int(n_sing)


def block_apply(inv, nb, cells, A, x):
    y = np.zeros(cells * nb * A)
    lib().oracle_block_apply(nb, i64(cells), A, _pd(inv), _pd(_d(x)), _pd(y))
    return y
