"""ctypes binding of libubolt_amd's C ABI (include/ubt.h).

The shared library is built in place by __graft_entry__.build() (or `make -C
ubolt_amd/csrc`). Loading it imports torch first, so the library binds to the
HIP runtime torch already loaded; a missing library raises instead of falling
back to anything.
"""
import ctypes as C
import os

import torch  # noqa: F401  (loads the HIP runtime the library links against)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libubolt_amd.so")

FP64, FP32 = 0, 1
BC_VACUUM, BC_REFLECT = 0, 1

STATUS = {0: "OK", 55: "ERR_MEM", 56: "ERR_SUP", 60: "ERR_ARG_SIZ", 61: "ERR_ARG_IDN", 62: "ERR_ARG_WRONG",
          63: "ERR_ARG_OUTOFRANGE", 71: "ERR_MAT_LU_ZRPVT", 73: "ERR_ARG_WRONGSTATE", 75: "ERR_ARG_INCOMP",
          77: "ERR_PLIB", 85: "ERR_ARG_NULL", 91: "ERR_NOT_CONVERGED", 97: "ERR_GPU"}


class UbtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code
        self.status = STATUS.get(code, str(code))


class SolverOptions(C.Structure):
    _fields_ = [("rtol", C.c_double), ("max_it", C.c_int), ("restart", C.c_int), ("mixed", C.c_int),
                ("inner_rtol", C.c_double), ("max_outer", C.c_int), ("pc_blocks", C.c_int), ("pc_sweep", C.c_int),
                ("pc_dsa", C.c_int), ("use_amat", C.c_int), ("gs_passes", C.c_int)]


class GroupOptions(C.Structure):
    _fields_ = [("outer_rtol", C.c_double), ("max_outer", C.c_int), ("inner_gamma", C.c_double),
                ("inner_rtol", C.c_double), ("store_moments", C.c_int), ("recon_steps", C.c_int)]


class GroupResult(C.Structure):
    _fields_ = [("g_up", C.c_int), ("sweeps", C.c_int), ("status", C.c_int), ("change", C.c_double),
                ("rho", C.c_double), ("err_est", C.c_double), ("iterations", C.c_int64),
                ("recon_steps", C.c_int64)]


class Bc(C.Structure):
    _fields_ = [("label", C.c_int), ("type", C.c_int), ("inflow", C.c_double), ("n_window_pairs", C.c_int),
                ("window", C.c_double * 4)]


P = C.c_void_p
I = C.c_int
D = C.c_double
PI = C.POINTER(C.c_int)
PD = C.POINTER(C.c_double)
PP = C.POINTER(C.c_void_p)
I64 = C.c_int64
PI64 = C.POINTER(C.c_int64)

# name -> argtypes (every function returns int status)
SIGNATURES = {
    "ubt_last_error": [C.c_char_p, C.c_size_t],
    "ubt_device_count": [PI],
    "ubt_quad_create": [I, I, PP],
    "ubt_quad_size": [P, PI, PD],
    "ubt_quad_get": [P, PD, PD, PI],
    "ubt_quad_destroy": [P],
    "ubt_lqn_octant": [I, PI, PD, PI, PI, PD],
    "ubt_mesh_create": [I, I, PD, I, I, PI, PP],
    "ubt_mesh_label_box": [P],
    "ubt_mesh_set_boundary_labels": [P, I, I, PI, PI],
    "ubt_mesh_size": [P, PI, PI, PI],
    "ubt_mesh_destroy": [P],
    "ubt_disc_create_structured": [P, I, PI, PD, C.POINTER(Bc), I, PP],
    "ubt_disc_create_dg_device": [P, P, I, P, I, P, PP],
    "ubt_disc_create_dg": [P, P, I, C.POINTER(Bc), I, PP],
    "ubt_disc_size": [P, PI, PI, PI, PI, PI],
    "ubt_disc_get": [P, C.c_char_p, P, C.POINTER(C.c_size_t)],
    "ubt_disc_destroy": [P],
    "ubt_op_create": [P, I, PP],
    "ubt_op_set_xs": [P, PD, PD, PD],
    "ubt_op_set_group": [P, I],
    "ubt_op_size": [P, PI64, PI64, PI],
    "ubt_op_mult": [P, I, P, P, P],
    "ubt_op_mult_unscattered": [P, I, P, P, P],
    "ubt_op_mult_transpose": [P, I, P, P, P],
    "ubt_op_geometry_classes": [P, P],
    "ubt_op_mult_unscattered_transpose": [P, I, P, P, P],
    "ubt_op_diagonal": [P, I, P, P],
    "ubt_op_set_variant": [P, I, I],
    "ubt_op_tuning": [P, PI, C.POINTER(C.c_float)],
    "ubt_op_destroy": [P],
    "ubt_moments": [P, I, P, P, I, P],
    "ubt_set_scalar_flux": [P, I, I, P, P],
    "ubt_scatter_source": [P, I, I, I, P, P],
    "ubt_rhs": [P, I, I, P, P],
    "ubt_blockinv_create": [P, PP],
    "ubt_blockinv_setup": [P, P],
    "ubt_blockinv_apply": [P, I, P, P, P],
    "ubt_blockinv_apply_transpose": [P, I, P, P, P],
    "ubt_blockinv_destroy": [P],
    "ubt_dsa_create": [P, PP],
    "ubt_dsa_set_options": [P, I, D, D, I, D, I],
    "ubt_dsa_set_void_policy": [P, I, D, D],
    "ubt_dsa_set_group": [P, I, P],
    "ubt_dsa_apply": [P, I, P, P, P],
    "ubt_dsa_apply_transpose": [P, I, P, P, P],
    "ubt_dsa_matvec": [P, P, P, P],
    "ubt_dsa_solve": [P, P, P, P],
    "ubt_dsa_solve_noalloc": [P, P, P, P],
    "ubt_dsa_info": [P, PI, PI],
    "ubt_dsa_destroy": [P],
    "ubt_sweep_create": [P, PP],
    "ubt_sweep_create_opts": [P, I, PP],
    "ubt_sweep_info": [P, PI, PI64],
    "ubt_sweep_trace_slots": [P, PI],
    "ubt_sweep_apply": [P, I, P, P, P],
    "ubt_sweep_apply_transpose": [P, I, P, P, P],
    "ubt_sweep_destroy": [P],
    "ubt_solver_create": [P, P, C.POINTER(SolverOptions), PP],
    "ubt_solver_refresh": [P, P],
    "ubt_solver_solve": [P, P, P, P],
    "ubt_solver_solve_transpose": [P, P, P, P],
    "ubt_solver_apply_pc": [P, I, P, P, P],
    "ubt_solver_result": [P, PI, PI, PD, PI, PD, I, PI],
    "ubt_solver_destroy": [P],
    "ubt_solver_solve_groups": [P, C.POINTER(GroupOptions), P, P],
    "ubt_op_scalar_flux": [P, I, P, P],
    "ubt_solver_groups_result": [P, C.POINTER(GroupResult), PI, PD, PI],
    "ubt_batch_create": [PP, I, PP],
    "ubt_batch_size": [P, PI64, PI],
    "ubt_batch_offsets": [P, PI64],
    "ubt_batch_set_group": [P, I],
    "ubt_batch_mult": [P, I, P, P, P],
    "ubt_batch_tuning": [P, PI, PI, C.POINTER(C.c_float)],
    "ubt_batch_set_sweep": [P, I],
    "ubt_batch_sweep_info": [P, PI, PI],
    "ubt_batch_set_body": [P, I, I],
    "ubt_batch_rhs": [P, I, P, P],
    "ubt_batch_set_scalar_flux": [P, I, P, P],
    "ubt_batch_blocks_apply": [P, I, P, P, P],
    "ubt_batch_sweep_apply": [P, I, P, P, P],
    "ubt_batch_destroy": [P],
    "ubt_batch_solver_create": [P, C.POINTER(SolverOptions), PP],
    "ubt_batch_solver_refresh": [P, P],
    "ubt_batch_solver_solve": [P, P, P, P],
    "ubt_batch_solver_result": [P, PI, PD, PI, PI, PI],
    "ubt_batch_solver_set_graph": [P, I],
    "ubt_batch_solver_destroy": [P],
    "ubt_batch_solver_solve_groups": [P, C.POINTER(GroupOptions), P, P],
    "ubt_batch_solver_groups_result": [P, C.POINTER(GroupResult), PI, PD, PI],
}

_lib = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libubolt_amd.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        _lib = L
    return _lib


def last_error():
    buf = C.create_string_buffer(1024)
    code = lib().ubt_last_error(buf, 1024)
    return code, buf.value.decode()


def call(name, *args):
    """Call an entry point and raise UbtError on a non-zero status."""
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        _, msg = last_error()
        raise UbtError(rc, msg)
    return rc


def dptr(t):
    """Device (or host) address of a torch tensor / numpy array as c_void_p."""
    if hasattr(t, "data_ptr"):
        return C.c_void_p(t.data_ptr())
    return t.ctypes.data_as(C.c_void_p)


def as_pd(a):
    return a.ctypes.data_as(PD)


def as_pi(a):
    return a.ctypes.data_as(PI)


def stream_ptr(stream=None):
    """The raw hipStream_t of a torch stream (the current one by default)."""
    if stream is None:
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)
