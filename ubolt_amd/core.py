"""Python handles over the C ABI: quadrature, mesh, discretisation, operator,
block inverse. Thin owners of the C handles; every compute call goes through
libubolt_amd and fails loudly (UbtError) on a non-zero status."""
import ctypes as C

import numpy as np
import torch

from . import capi
from .capi import call, dptr, stream_ptr, as_pd, as_pi

FP64, FP32 = capi.FP64, capi.FP32


def torch_dtype(prec):
    return torch.float64 if prec == FP64 else torch.float32


class _Handle:
    _destroy = None

    def __init__(self):
        self.h = C.c_void_p()

    def __del__(self):
        try:
            if self.h and self._destroy:
                getattr(capi.lib(), self._destroy)(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass


class Quadrature(_Handle):
    """SN quadrature (reference AngularQuadrature / SNQuadrature{,2D,3D})."""
    _destroy = "ubt_quad_destroy"

    def __init__(self, dim, sn_order):
        super().__init__()
        call("ubt_quad_create", dim, sn_order, C.byref(self.h))
        n = C.c_int()
        sw = C.c_double()
        call("ubt_quad_size", self.h, C.byref(n), C.byref(sw))
        self.dim, self.sn_order, self.n_angles, self.sum_weights = dim, sn_order, n.value, sw.value
        self.omega = np.zeros((self.n_angles, 3))
        self.w = np.zeros(self.n_angles)
        self.reflect = np.zeros((self.n_angles, 3), dtype=np.int32)
        call("ubt_quad_get", self.h, as_pd(self.omega), as_pd(self.w), as_pi(self.reflect))


def lqn_octant(sn_order):
    nl, npt = C.c_int(), C.c_int()
    mu = np.zeros(9)
    tr = np.zeros(3 * 45, dtype=np.int32)
    wt = np.zeros(45)
    call("ubt_lqn_octant", sn_order, C.byref(nl), as_pd(mu), C.byref(npt), as_pi(tr), as_pd(wt))
    return mu[:nl.value].copy(), tr[:3 * npt.value].reshape(-1, 3).copy(), wt[:npt.value].copy()


class Mesh(_Handle):
    """An unstructured mesh of one cell type (triangles, quads, tets, hexes)."""
    _destroy = "ubt_mesh_destroy"

    def __init__(self, coords, cells, label_box=True):
        super().__init__()
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        self.dim = coords.shape[1]
        self.coords, self.cells = coords, cells
        call("ubt_mesh_create", self.dim, coords.shape[0], as_pd(coords), cells.shape[0], cells.shape[1],
             as_pi(cells), C.byref(self.h))
        if label_box:
            call("ubt_mesh_label_box", self.h)

    def set_boundary_labels(self, face_verts, labels):
        fv = np.ascontiguousarray(face_verts, dtype=np.int32)
        lb = np.ascontiguousarray(labels, dtype=np.int32)
        call("ubt_mesh_set_boundary_labels", self.h, fv.shape[0], fv.shape[1], as_pi(fv), as_pi(lb))

    def sizes(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        call("ubt_mesh_size", self.h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value


def bc_array(bcs):
    """bcs: {label: "vacuum" | "reflect" | dict(type=..., inflow=..., window=[...])}"""
    items = list((bcs or {}).items())
    arr = (capi.Bc * max(1, len(items)))()
    for i, (label, spec) in enumerate(items):
        if isinstance(spec, str):
            spec = {"type": spec}
        arr[i].label = int(label)
        arr[i].type = capi.BC_REFLECT if spec["type"] == "reflect" else capi.BC_VACUUM
        arr[i].inflow = float(spec.get("inflow", 0.0))
        win = list(spec.get("window", []))
        arr[i].n_window_pairs = len(win) // 2
        for k, v in enumerate(win):
            arr[i].window[k] = v
    return arr, len(items)


class Disc(_Handle):
    """A discretisation: structured FD box or unstructured DG0/DG1."""
    _destroy = "ubt_disc_destroy"

    @classmethod
    def structured(cls, quad, n_cells, lengths, bcs=None):
        self = cls()
        n = np.ascontiguousarray(n_cells, dtype=np.int32)
        L = np.ascontiguousarray(lengths, dtype=np.float64)
        arr, nb = bc_array(bcs)
        call("ubt_disc_create_structured", quad.h, len(n), as_pi(n), as_pd(L), arr, nb, C.byref(self.h))
        self.kind, self.quad = "fd", quad
        self._sizes()
        return self

    @classmethod
    def dg(cls, quad, mesh, order, bcs=None, device=False):
        """device: the per-cell and per-face geometry computed by GPU kernels"""
        self = cls()
        arr, nb = bc_array(bcs)
        if device:
            call("ubt_disc_create_dg_device", quad.h, mesh.h, order, arr, nb, None, C.byref(self.h))
        else:
            call("ubt_disc_create_dg", quad.h, mesh.h, order, arr, nb, C.byref(self.h))
        self.kind, self.quad, self.mesh, self.order = "dg", quad, mesh, order
        self._sizes()
        return self

    def _sizes(self):
        v = [C.c_int() for _ in range(5)]
        call("ubt_disc_size", self.h, *[C.byref(x) for x in v])
        self.dim, self.n_cells, self.nb, self.n_angles, self.nf = [x.value for x in v]
        self.n_nodes = self.n_cells * self.nb
        self.n_rows = self.n_nodes * self.n_angles

    INT_TABLES = {"face_nb", "face_label", "n"}

    def get(self, name):
        nbytes = C.c_size_t(0)
        call("ubt_disc_get", self.h, name.encode(), None, C.byref(nbytes))
        dt = np.int32 if name in self.INT_TABLES else np.float64
        out = np.zeros(nbytes.value // np.dtype(dt).itemsize, dtype=dt)
        call("ubt_disc_get", self.h, name.encode(), dptr(out), C.byref(nbytes))
        return out


class Operator(_Handle):
    """The device transport operator (reference TransportOperator + terms)."""
    _destroy = "ubt_op_destroy"

    def __init__(self, disc, n_groups=1):
        super().__init__()
        self.disc, self.n_groups = disc, n_groups
        call("ubt_op_create", disc.h, n_groups, C.byref(self.h))
        self.n_rows, self.n_nodes = disc.n_rows, disc.n_nodes

    def set_xs(self, sigma_t, sigma_s, source=None):
        """sigma_t [G, cells], sigma_s [G_from, G_to, cells], source [G, cells]"""
        G, nc = self.n_groups, self.disc.n_cells
        st = np.ascontiguousarray(np.broadcast_to(np.asarray(sigma_t, dtype=np.float64).reshape(G, -1), (G, nc)))
        ss = np.ascontiguousarray(np.broadcast_to(np.asarray(sigma_s, dtype=np.float64).reshape(G, G, -1), (G, G, nc)))
        q = None
        if source is not None:
            q = np.ascontiguousarray(np.broadcast_to(np.asarray(source, dtype=np.float64).reshape(G, -1), (G, nc)))
        self.sigma_t, self.sigma_s, self.source = st, ss, q
        call("ubt_op_set_xs", self.h, as_pd(st), as_pd(ss), as_pd(q) if q is not None else None)

    def set_group(self, g):
        call("ubt_op_set_group", self.h, g)
        self.group = g

    def scalar_flux(self, g, stream=None):
        """the scalar flux cached for group g (FP64, n_nodes)"""
        phi = torch.empty(self.n_nodes, dtype=torch.float64, device="cuda")
        call("ubt_op_scalar_flux", self.h, g, dptr(phi), stream_ptr(stream))
        return phi

    def set_variant(self, prec, variant):
        """apply kernel: 0 tile, 1 wavefront per cell, 2/3 multi-cell tile (2/4 cells per thread),
        4/5/6 pencil on a structured box (8/16/32 cells per workgroup), 7 four-angle tile, 8/9 box tile on a
        uniform DG1 box (2/4 rows per thread), 10 strip march on a uniform DG1 box, 11/12 z-march on a 3D FD box
        (4/8 cells per lane column), -1 re-tune"""
        call("ubt_op_set_variant", self.h, prec, variant)

    def tuning(self):
        v = (C.c_int * 2)()
        us = (C.c_float * 26)()
        call("ubt_op_tuning", self.h, v, us)
        return {"variant_fp64": v[0], "variant_fp32": v[1], "us": [round(x, 2) for x in us]}

    def vec(self, prec=FP64, fill=None):
        t = torch.zeros(self.n_rows, dtype=torch_dtype(prec), device="cuda")
        if fill is not None:
            t.copy_(torch.as_tensor(fill))
        return t

    @staticmethod
    def _prec(t):
        return FP64 if t.dtype == torch.float64 else FP32

    def mult(self, x, y, stream=None):
        call("ubt_op_mult", self.h, self._prec(x), dptr(x), dptr(y), stream_ptr(stream))
        return y

    def mult_unscattered(self, x, y, stream=None):
        call("ubt_op_mult_unscattered", self.h, self._prec(x), dptr(x), dptr(y), stream_ptr(stream))
        return y

    def geometry_classes(self):
        n = C.c_int()
        call("ubt_op_geometry_classes", self.h, C.byref(n))
        return n.value

    def mult_transpose(self, x, y, stream=None):
        """y = A^T x, the adjoint operator"""
        call("ubt_op_mult_transpose", self.h, self._prec(x), dptr(x), dptr(y), stream_ptr(stream))
        return y

    def mult_unscattered_transpose(self, x, y, stream=None):
        call("ubt_op_mult_unscattered_transpose", self.h, self._prec(x), dptr(x), dptr(y), stream_ptr(stream))
        return y

    def diagonal(self, d, stream=None):
        call("ubt_op_diagonal", self.h, self._prec(d), dptr(d), stream_ptr(stream))
        return d

    def moments(self, psi, phi, n_moments=1, stream=None):
        call("ubt_moments", self.h, self._prec(psi), dptr(psi), dptr(phi), n_moments, stream_ptr(stream))
        return phi

    def set_scalar_flux(self, g, psi, stream=None):
        call("ubt_set_scalar_flux", self.h, g, self._prec(psi), dptr(psi), stream_ptr(stream))

    def scatter_source(self, g_from, g_to, b, stream=None):
        call("ubt_scatter_source", self.h, g_from, g_to, self._prec(b), dptr(b), stream_ptr(stream))

    def rhs(self, g, b, stream=None):
        call("ubt_rhs", self.h, g, self._prec(b), dptr(b), stream_ptr(stream))
        return b


class BlockInverse(_Handle):
    """Element block inverse (reference ElementBlockInverse)."""
    _destroy = "ubt_blockinv_destroy"

    def __init__(self, op):
        super().__init__()
        self.op = op
        call("ubt_blockinv_create", op.h, C.byref(self.h))

    def setup(self, stream=None):
        call("ubt_blockinv_setup", self.h, stream_ptr(stream))

    def apply(self, x, y, stream=None):
        call("ubt_blockinv_apply", self.h, Operator._prec(x), dptr(x), dptr(y), stream_ptr(stream))
        return y

    def apply_transpose(self, x, y, stream=None):
        call("ubt_blockinv_apply_transpose", self.h, Operator._prec(x), dptr(x), dptr(y), stream_ptr(stream))
        return y


class DSA(_Handle):
    """DSA preconditioner stage (reference DSAPrecon): caller-owned."""
    _destroy = "ubt_dsa_destroy"

    def __init__(self, op, consistent_d=True, consistent_power=1.5, mip_penalty=4.0, inner_max_it=200,
                 inner_rtol=1e-6, inner_check_every=8, void_bridge=True, void_sigma_t=0.0, void_d=0.0):
        super().__init__()
        self.op = op
        call("ubt_dsa_create", op.h, C.byref(self.h))
        self.set_options(consistent_d, consistent_power, mip_penalty, inner_max_it, inner_rtol, inner_check_every)
        self.set_void_policy(void_bridge, void_sigma_t, void_d)

    def set_void_policy(self, bridge=True, void_sigma_t=0.0, void_d=0.0):
        """-dsa_void_bridge, -dsa_void_sigma_t, -dsa_void_d; effective at the next set_group"""
        call("ubt_dsa_set_void_policy", self.h, int(bridge), void_sigma_t, void_d)

    def set_options(self, consistent_d=True, consistent_power=1.5, mip_penalty=4.0, inner_max_it=200, inner_rtol=1e-6,
                    inner_check_every=8):
        call("ubt_dsa_set_options", self.h, int(consistent_d), consistent_power, mip_penalty, inner_max_it, inner_rtol,
             inner_check_every)

    def set_group(self, g, stream=None):
        call("ubt_dsa_set_group", self.h, g, stream_ptr(stream))

    def apply(self, x, y, stream=None):
        call("ubt_dsa_apply", self.h, Operator._prec(x), dptr(x), dptr(y), stream_ptr(stream))
        return y

    def apply_transpose(self, x, y, stream=None):
        call("ubt_dsa_apply_transpose", self.h, Operator._prec(x), dptr(x), dptr(y), stream_ptr(stream))
        return y

    def matvec(self, x, y, stream=None):
        call("ubt_dsa_matvec", self.h, dptr(x), dptr(y), stream_ptr(stream))
        return y

    def solve(self, b, x, stream=None):
        call("ubt_dsa_solve", self.h, dptr(b), dptr(x), stream_ptr(stream))
        return x

    def info(self):
        it, nn = C.c_int(), C.c_int()
        call("ubt_dsa_info", self.h, C.byref(it), C.byref(nn))
        return it.value, nn.value


class Sweep(_Handle):
    """Transport sweep psi = L^{-1} r (the streaming stage)."""
    _destroy = "ubt_sweep_destroy"

    def __init__(self, op, lds_traces=True):
        super().__init__()
        self.op = op
        call("ubt_sweep_create_opts", op.h, int(lds_traces), C.byref(self.h))
        nl, nlag, ts = C.c_int(), C.c_int64(), C.c_int()
        call("ubt_sweep_info", self.h, C.byref(nl), C.byref(nlag))
        call("ubt_sweep_trace_slots", self.h, C.byref(ts))
        self.n_levels, self.n_lagged, self.trace_slots = nl.value, nlag.value, ts.value

    def apply(self, r, psi, stream=None):
        call("ubt_sweep_apply", self.h, Operator._prec(r), dptr(r), dptr(psi), stream_ptr(stream))
        return psi

    def apply_transpose(self, r, psi, stream=None):
        call("ubt_sweep_apply_transpose", self.h, Operator._prec(r), dptr(r), dptr(psi), stream_ptr(stream))
        return psi


class Solver(_Handle):
    """Right-preconditioned FGMRES with the composite preconditioner
    (reference TransportSolver)."""
    _destroy = "ubt_solver_destroy"

    def __init__(self, op, dsa=None, rtol=1e-8, max_it=500, restart=30, mixed=False, inner_rtol=1e-4, max_outer=20,
                 pc_blocks=True, pc_sweep=True, pc_dsa=True, use_amat=True, gs_passes=2):
        super().__init__()
        self.op, self.dsa = op, dsa
        o = capi.SolverOptions(rtol, max_it, restart, int(mixed), inner_rtol, max_outer, int(pc_blocks), int(pc_sweep),
                               int(pc_dsa), int(use_amat), int(gs_passes))
        call("ubt_solver_create", op.h, dsa.h if dsa is not None else None, C.byref(o), C.byref(self.h))

    def refresh(self, stream=None):
        call("ubt_solver_refresh", self.h, stream_ptr(stream))

    def solve(self, b, x, stream=None):
        call("ubt_solver_solve", self.h, dptr(b), dptr(x), stream_ptr(stream))
        return self.result()

    def solve_transpose(self, b, x, stream=None):
        """x = A^{-T} b, the adjoint solve"""
        call("ubt_solver_solve_transpose", self.h, dptr(b), dptr(x), stream_ptr(stream))
        return self.result()

    def apply_pc(self, r, y, stream=None):
        call("ubt_solver_apply_pc", self.h, Operator._prec(r), dptr(r), dptr(y), stream_ptr(stream))
        return y

    def solve_groups(self, psi=None, outer_rtol=1e-8, max_outer=1000, inner_gamma=0.01, inner_rtol=0.0,
                     store_moments=False, recon_steps=10, stream=None):
        """every group by the reference driver's group sweep (and, with
        upscatter, its outer iteration): psi (n_groups x n_rows FP64, made if
        None) receives every group's angular flux; with store_moments psi is
        one vector (n_rows) and every group's scalar flux is read with
        Operator.scalar_flux; returns (psi, report)"""
        G, n = self.op.n_groups, self.op.n_rows
        shape = (n,) if store_moments else (G, n)
        if psi is None:
            psi = torch.zeros(shape, dtype=torch.float64, device="cuda")
        assert psi.dtype == torch.float64 and psi.numel() == (n if store_moments else G * n) and psi.is_contiguous()
        o = capi.GroupOptions(outer_rtol, max_outer, inner_gamma, inner_rtol, int(store_moments), recon_steps)
        call("ubt_solver_solve_groups", self.h, C.byref(o), dptr(psi), stream_ptr(stream))
        return psi, self.groups_result()

    def groups_result(self):
        G = self.op.n_groups
        r = capi.GroupResult()
        it, rs = (C.c_int * G)(), (C.c_int * G)()
        rr = (C.c_double * G)()
        call("ubt_solver_groups_result", self.h, C.byref(r), it, rr, rs)
        return _group_report(r, list(it), list(rr), list(rs))

    def result(self):
        it, outer, reason, nh = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        rr = C.c_double()
        hist = np.zeros(4096)
        call("ubt_solver_result", self.h, C.byref(it), C.byref(outer), C.byref(rr), C.byref(reason), as_pd(hist), 4096,
             C.byref(nh))
        return {"iterations": it.value, "outer": outer.value, "rel_residual": rr.value, "reason": reason.value,
                "history": hist[:min(nh.value, 4096)].copy()}


def _group_report(r, its, res, reason):
    return {"g_up": r.g_up, "sweeps": r.sweeps, "status": r.status, "change": r.change, "rho": r.rho,
            "err_est": r.err_est, "iterations": r.iterations, "recon_steps": r.recon_steps,
            "group_iterations": its, "group_rel_residual": res, "group_reason": reason}


class Batch(_Handle):
    """Many independent DG problems driven as one: one persistent launch per
    apply, merged sweep levels, batched block inverse. The operators must
    outlive the batch."""
    _destroy = "ubt_batch_destroy"

    def __init__(self, ops):
        super().__init__()
        self.ops = list(ops)
        arr = (C.c_void_p * len(self.ops))(*[o.h.value for o in self.ops])
        call("ubt_batch_create", C.cast(arr, capi.PP), len(self.ops), C.byref(self.h))
        nr, npb = C.c_int64(), C.c_int()
        call("ubt_batch_size", self.h, C.byref(nr), C.byref(npb))
        self.n_rows, self.n_problems = nr.value, npb.value
        off = (C.c_int64 * (self.n_problems + 1))()
        call("ubt_batch_offsets", self.h, off)
        self.offsets = np.array(off[:], dtype=np.int64)

    def set_group(self, g):
        call("ubt_batch_set_group", self.h, g)

    def vec(self, prec=FP64):
        return torch.zeros(self.n_rows, dtype=torch_dtype(prec), device="cuda")

    def part(self, v, p):
        return v[self.offsets[p]:self.offsets[p + 1]]

    def mult(self, x, y, stream=None):
        call("ubt_batch_mult", self.h, Operator._prec(x), dptr(x), dptr(y), stream_ptr(stream))
        return y

    def blocks_apply(self, x, y, stream=None):
        call("ubt_batch_blocks_apply", self.h, Operator._prec(x), dptr(x), dptr(y), stream_ptr(stream))
        return y

    def sweep_apply(self, r, psi, stream=None):
        call("ubt_batch_sweep_apply", self.h, Operator._prec(r), dptr(r), dptr(psi), stream_ptr(stream))
        return psi

    def rhs(self, b, stream=None):
        call("ubt_batch_rhs", self.h, Operator._prec(b), dptr(b), stream_ptr(stream))
        return b

    def set_scalar_flux(self, psi, stream=None):
        call("ubt_batch_set_scalar_flux", self.h, Operator._prec(psi), dptr(psi), stream_ptr(stream))

    def set_body(self, prec, body):
        """apply body of one precision: 0 tile, 1 four-angle tile, 2 box tile, 3 strip march (uniform DG1 boxes)"""
        call("ubt_batch_set_body", self.h, prec, body)

    def set_sweep(self, box):
        """the record-free box sweep where the batch allows it (True), or the record sweep"""
        call("ubt_batch_set_sweep", self.h, int(box))

    def sweep_info(self):
        box, exact = (C.c_int * 2)(), (C.c_int * 2)()
        call("ubt_batch_sweep_info", self.h, box, exact)
        return {"box_fp64": bool(box[0]), "box_fp32": bool(box[1]), "exact_fp64": bool(exact[0]),
                "exact_fp32": bool(exact[1])}

    def tuning(self):
        body, persist = (C.c_int * 2)(), (C.c_int * 2)()
        us = (C.c_float * 8)()
        call("ubt_batch_tuning", self.h, body, persist, us)
        return {"body_fp64": body[0], "body_fp32": body[1], "persist_fp64": persist[0], "persist_fp32": persist[1],
                "us": [round(x, 2) for x in us]}


class BatchSolver(_Handle):
    """FGMRES over every problem of a batch at once (per-problem Hessenberg on
    the device, restart cycles replayed as a HIP graph); element blocks then
    the sweep as the preconditioner."""
    _destroy = "ubt_batch_solver_destroy"

    def __init__(self, batch, rtol=1e-8, max_it=500, restart=30, mixed=False, inner_rtol=1e-4, max_outer=20,
                 pc_blocks=True, pc_sweep=True, gs_passes=2):
        super().__init__()
        self.batch = batch
        o = capi.SolverOptions(rtol, max_it, restart, int(mixed), inner_rtol, max_outer, int(pc_blocks), int(pc_sweep),
                               0, 0, int(gs_passes))
        call("ubt_batch_solver_create", batch.h, C.byref(o), C.byref(self.h))

    def refresh(self, stream=None):
        call("ubt_batch_solver_refresh", self.h, stream_ptr(stream))

    def set_graph(self, on):
        call("ubt_batch_solver_set_graph", self.h, int(on))

    def solve(self, b, x, stream=None):
        call("ubt_batch_solver_solve", self.h, dptr(b), dptr(x), stream_ptr(stream))
        return self.result()

    def solve_groups(self, psi=None, outer_rtol=1e-8, max_outer=1000, inner_gamma=0.01, inner_rtol=0.0, stream=None):
        """every group of every problem (the group sweep per problem, one
        group's solves batched): psi (G_max x n_rows FP64, made if None);
        returns (psi, per-problem reports)"""
        G = max(o.n_groups for o in self.batch.ops)
        n = self.batch.n_rows
        if psi is None:
            psi = torch.zeros((G, n), dtype=torch.float64, device="cuda")
        assert psi.dtype == torch.float64 and psi.numel() == G * n and psi.is_contiguous()
        o = capi.GroupOptions(outer_rtol, max_outer, inner_gamma, inner_rtol, 0, 0)
        call("ubt_batch_solver_solve_groups", self.h, C.byref(o), dptr(psi), stream_ptr(stream))
        return psi, self.groups_result()

    def groups_result(self):
        G = max(o.n_groups for o in self.batch.ops)
        n = self.batch.n_problems
        r = (capi.GroupResult * n)()
        it, rs = (C.c_int * (n * G))(), (C.c_int * (n * G))()
        rr = (C.c_double * (n * G))()
        call("ubt_batch_solver_groups_result", self.h, r, it, rr, rs)
        return [_group_report(r[p], list(it[p * G:(p + 1) * G]), list(rr[p * G:(p + 1) * G]), list(rs[p * G:(p + 1) * G]))
                for p in range(n)]

    def result(self):
        n = self.batch.n_problems
        it = (C.c_int * n)()
        rr = (C.c_double * n)()
        rs = (C.c_int * n)()
        outer, cycles = C.c_int(), C.c_int()
        call("ubt_batch_solver_result", self.h, it, rr, rs, C.byref(outer), C.byref(cycles))
        return {"iterations": list(it), "rel_residual": list(rr), "reason": list(rs), "outer": outer.value,
                "cycles": cycles.value}
