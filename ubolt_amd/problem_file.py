"""Reference problem files: the JSON problem schema and Gmsh 2.2 meshes.

Reads the problem files of the reference driver (include/ubolt/problem_spec.hpp:
38-102, src/problem_speck.kokkos.cxx, schema in docs/problem_files.md) into
this build's objects - quadrature, structured box or unstructured DG mesh with
its "Face Sets" boundary conditions, per-cell cross sections painted from the
materials table - and runs the group sweep of the reference driver
(tests/transportk.kokkos.cxx:1019-1208) on the GPU.

The problem file is strict (an unknown key is an error, "_comment" ignored);
the materials schema is tolerant (unknown keys accepted). What the operator
path of this build does not implement is refused with UnsupportedProblem
rather than approximated: 1D (the slab backend), the CG-SUPG discretisation
and the Dirichlet-cell vacuum treatment. output.flux_vtk is accepted and not
written (no VTK output here).
"""
import json
import math
import os

import numpy as np

from . import core


class ProblemError(ValueError):
    """the file breaks the schema (the reference driver's error cases)"""


class UnsupportedProblem(ProblemError):
    """a valid problem whose discretisation is outside this build's operator path"""


FOUR_PI = 4.0 * math.pi
# "Face Sets" ids of PETSc's box mesher, the names the structured backends use
FACE_IDS = {
    2: {"bottom": 1, "right": 2, "top": 3, "left": 4},
    3: {"bottom": 1, "top": 2, "front": 3, "back": 4, "right": 5, "left": 6},
}

TOP_KEYS = {"_comment", "dimension", "mesh", "sn_order", "materials", "regions", "boundary_conditions",
            "vacuum_treatment", "output"}
MESH_KEYS = {"_comment", "type", "n_cells", "lengths", "order", "discretisation", "supg_zeta", "simplex", "file"}
REGION_KEYS = {"_comment", "background", "paint", "cell_sets"}
PAINT_KEYS = {"_comment", "material", "box", "interval"}
BC_KEYS = {"_comment", "type", "inflow", "window"}
OUTPUT_KEYS = {"_comment", "flux_vtk"}


def _strict(obj, allowed, where):
    if not isinstance(obj, dict):
        raise ProblemError(f"{where} must be an object")
    bad = sorted(set(obj) - allowed)
    if bad:
        raise ProblemError(f"unknown key(s) {bad} in {where}")


# ---------------------------------------------------------------------------
# Gmsh 2.2 ASCII
# ---------------------------------------------------------------------------
# element type -> (topological dimension, vertices)
GMSH_TYPES = {1: (1, 2), 2: (2, 3), 3: (2, 4), 4: (3, 4), 5: (3, 8), 15: (0, 1)}


class GmshMesh:
    """nodes (n, 3), elements by topological dimension: vertex index arrays
    (0-based) and the physical tag of each"""

    def __init__(self, nodes, elements, physical_names):
        self.nodes = nodes
        self.elements = elements          # {tdim: [(vertices [m, k], phys [m]) per element type]}
        self.physical_names = physical_names

    def cells(self, dim):
        blocks = self.elements.get(dim, [])
        if len(blocks) != 1:
            raise ProblemError(f"the mesh has {len(blocks)} cell types of dimension {dim}; one is supported")
        return blocks[0]

    def faces(self, dim):
        return self.elements.get(dim - 1, [])


def read_gmsh(path):
    """Gmsh MSH 2.2 ASCII: $MeshFormat, optional $PhysicalNames, $Nodes,
    $Elements (number, type, number of tags, tags - the first the physical
    group, the second the elementary entity -, nodes)."""
    with open(path) as f:
        lines = [ln.strip() for ln in f]
    i, nodes, elements, names = 0, None, {}, {}
    ids = {}
    while i < len(lines):
        sec = lines[i]
        if sec == "$MeshFormat":
            ver, ftype = lines[i + 1].split()[:2]
            if not ver.startswith("2.") or ftype != "0":
                raise ProblemError(f"{path}: only ASCII MSH 2.x is read, got version {ver} file-type {ftype}")
            i += 3
        elif sec == "$PhysicalNames":
            n = int(lines[i + 1])
            for ln in lines[i + 2:i + 2 + n]:
                d, tag, name = ln.split(maxsplit=2)
                names[(int(d), int(tag))] = name.strip('"')
            i += n + 3
        elif sec == "$Nodes":
            n = int(lines[i + 1])
            nodes = np.zeros((n, 3))
            for k, ln in enumerate(lines[i + 2:i + 2 + n]):
                p = ln.split()
                ids[int(p[0])] = k
                nodes[k] = [float(v) for v in p[1:4]]
            i += n + 3
        elif sec == "$Elements":
            n = int(lines[i + 1])
            raw = {}
            for ln in lines[i + 2:i + 2 + n]:
                p = [int(v) for v in ln.split()]
                et, nt = p[1], p[2]
                if et not in GMSH_TYPES:
                    raise ProblemError(f"{path}: element type {et} is not supported")
                tags = p[3:3 + nt]
                verts = [ids[v] for v in p[3 + nt:]]
                if len(verts) != GMSH_TYPES[et][1]:
                    raise ProblemError(f"{path}: element {p[0]} has {len(verts)} nodes, type {et} wants "
                                       f"{GMSH_TYPES[et][1]}")
                raw.setdefault(et, ([], []))
                raw[et][0].append(verts)
                raw[et][1].append(tags[0] if tags else 0)
            for et, (v, ph) in raw.items():
                elements.setdefault(GMSH_TYPES[et][0], []).append(
                    (np.asarray(v, dtype=np.int64), np.asarray(ph, dtype=np.int64)))
            i += n + 3
        else:
            i += 1
    if nodes is None:
        raise ProblemError(f"{path}: no $Nodes section")
    return GmshMesh(nodes, elements, names)


# Gmsh node order -> this build's cell vertex order (the box builders'):
# quads and hexes are counter-clockwise / bottom-then-top in both, simplices
# are any order
def _cells_for(kind_vertices):
    return np.ascontiguousarray(kind_vertices, dtype=np.int32)


# ---------------------------------------------------------------------------
# materials
# ---------------------------------------------------------------------------
class Materials:
    def __init__(self, spec, where):
        if not isinstance(spec, dict):
            raise ProblemError(f"{where}: the materials table must be an object")
        rep = spec.get("representation", "multigroup")
        if rep != "multigroup":
            raise ProblemError(f"{where}: representation {rep!r} is not supported (multigroup only)")
        G = spec.get("n_groups")
        if not isinstance(G, int) or G < 1:
            raise ProblemError(f"{where}: n_groups must be a positive integer")
        mats = spec.get("materials")
        if not isinstance(mats, list) or not mats:
            raise ProblemError(f"{where}: materials must be a non-empty array")
        self.G = G
        self.sigma_t = np.zeros((len(mats), G))
        self.sigma_s = np.zeros((len(mats), G, G))
        self.source = np.zeros((len(mats), G))
        self.names = {}
        seen = set()
        for m in mats:
            mid = m.get("id")
            if not isinstance(mid, int) or not 0 <= mid < len(mats) or mid in seen:
                raise ProblemError(f"{where}: material ids must be dense 0..{len(mats) - 1}, got {mid!r}")
            seen.add(mid)
            st = np.asarray(m.get("Sigma_t"), dtype=float)
            ss = np.asarray(m.get("Sigma_s"), dtype=float)
            if st.shape != (G,):
                raise ProblemError(f"{where}: material {mid} Sigma_t must have {G} entries")
            if ss.shape != (G, G):
                raise ProblemError(f"{where}: material {mid} Sigma_s must be {G} x {G} ([from][to])")
            self.sigma_t[mid], self.sigma_s[mid] = st, ss
            if "Source" in m:
                q = np.asarray(m["Source"], dtype=float)
                if q.shape != (G,):
                    raise ProblemError(f"{where}: material {mid} Source must have {G} entries")
                self.source[mid] = q
            if "name" in m:
                self.names[m["name"]] = mid

    def resolve(self, ref, where):
        if isinstance(ref, bool):
            raise ProblemError(f"{where}: material {ref!r} is neither an id nor a name")
        if isinstance(ref, int):
            if not 0 <= ref < len(self.sigma_t):
                raise ProblemError(f"{where}: no material {ref}")
            return ref
        if isinstance(ref, str) and ref in self.names:
            return self.names[ref]
        raise ProblemError(f"{where}: no material {ref!r}")


# ---------------------------------------------------------------------------
# the problem
# ---------------------------------------------------------------------------
class ProblemFile:
    """A parsed problem file, built into this build's objects: `problem`
    (quadrature, discretisation, per-cell sigma_t [G, cells], sigma_s
    [G, G, cells], source [G, cells]), the material of every cell, and
    whether any material upscatters."""

    def __init__(self, path):
        self.path = os.path.abspath(path)
        self.name = os.path.splitext(os.path.basename(path))[0]
        base = os.path.dirname(self.path)
        with open(self.path) as f:
            spec = json.load(f)
        if isinstance(spec, dict) and "inflow" in spec:
            raise ProblemError("the top-level 'inflow' key was removed: give the inflow per face in "
                               "boundary_conditions")
        _strict(spec, TOP_KEYS, "the problem file")
        dim = spec.get("dimension")
        if dim not in (1, 2, 3):
            raise ProblemError("dimension must be 1, 2 or 3")
        mesh = spec.get("mesh")
        _strict(mesh, MESH_KEYS, "mesh")
        mtype = mesh.get("type", "structured")
        if mtype not in ("structured", "unstructured"):
            raise ProblemError(f"mesh.type must be 'structured' or 'unstructured', got {mtype!r}")
        disc_kind = mesh.get("discretisation", "dg")
        if disc_kind not in ("dg", "cg_supg"):
            raise ProblemError(f"mesh.discretisation must be 'dg' or 'cg_supg', got {disc_kind!r}")
        order = mesh.get("order", 0)
        if order not in (0, 1):
            raise ProblemError("mesh.order must be 0 or 1")
        if mtype == "structured":
            for k in ("order", "discretisation", "supg_zeta", "simplex", "file"):
                if k in mesh:
                    raise ProblemError(f"mesh.{k} is for unstructured meshes only")
        has_file = "file" in mesh
        if has_file:
            for k in ("n_cells", "lengths", "simplex"):
                if k in mesh:
                    raise ProblemError(f"mesh.{k} is an error alongside mesh.file")
        else:
            n = mesh.get("n_cells")
            L = mesh.get("lengths")
            if not (isinstance(n, list) and len(n) == dim and all(isinstance(v, int) and v > 0 for v in n)):
                raise ProblemError(f"mesh.n_cells must be {dim} positive integers")
            if not (isinstance(L, list) and len(L) == dim and all(isinstance(v, (int, float)) and v > 0 for v in L)):
                raise ProblemError(f"mesh.lengths must be {dim} positive numbers")
        sn = spec.get("sn_order")
        if not isinstance(sn, int) or sn < 2 or sn % 2:
            raise ProblemError("sn_order must be a positive even integer")
        vt = spec.get("vacuum_treatment", "ghost_flux")
        if vt not in ("ghost_flux", "dirichlet_cell"):
            raise ProblemError(f"vacuum_treatment must be 'ghost_flux' or 'dirichlet_cell', got {vt!r}")
        if vt == "dirichlet_cell" and order == 1:
            raise ProblemError("the Dirichlet-cell vacuum treatment is an error at order 1")
        if "output" in spec:
            _strict(spec["output"], OUTPUT_KEYS, "output")
        mats = spec.get("materials")
        if isinstance(mats, str):
            with open(os.path.join(base, mats)) as f:
                mats = json.load(f)
        if mats is None:
            raise ProblemError("materials is required")
        self.materials = Materials(mats, "materials")
        regions = spec.get("regions", {})
        _strict(regions, REGION_KEYS, "regions")
        if "cell_sets" in regions and mtype == "structured":
            raise ProblemError("regions.cell_sets is for unstructured meshes only")
        bcs_spec = spec.get("boundary_conditions", {})
        if not isinstance(bcs_spec, dict):
            raise ProblemError("boundary_conditions must be an object")

        self.dim, self.sn_order, self.order, self.mesh_type = dim, sn, order, mtype
        # refused: valid problems outside this build's operator path
        if dim == 1:
            raise UnsupportedProblem("1D (the slab backend) is not part of this build")
        if disc_kind == "cg_supg":
            raise UnsupportedProblem("the CG-SUPG discretisation is not part of this build")
        if vt == "dirichlet_cell":
            raise UnsupportedProblem("the Dirichlet-cell vacuum treatment is not part of this build (ghost-flux only)")

        quad = core.Quadrature(dim, sn)
        self.quad = quad
        face_ids = FACE_IDS[dim]
        gmsh = None
        if mtype == "structured":
            n, L = mesh["n_cells"], mesh["lengths"]
            centres = self._grid_centres(n, L)
            self.n_cells_axis, self.lengths = n, L
        else:
            if has_file:
                gmsh = read_gmsh(os.path.join(base, mesh["file"]))
                coords_all = gmsh.nodes
                cells, cell_phys = gmsh.cells(dim)
                used = np.unique(cells)
                remap = np.full(len(coords_all), -1, dtype=np.int64)
                remap[used] = np.arange(len(used))
                coords = np.ascontiguousarray(coords_all[used][:, :dim])
                if dim == 2 and np.abs(coords_all[used][:, 2]).max() > 0:
                    raise ProblemError("a 2D mesh file must lie in the z = 0 plane")
                cells = remap[cells]
                self.cell_sets = cell_phys
                m = core.Mesh(coords, _cells_for(cells), label_box=False)
                fv, fl = [], []
                for verts, phys in gmsh.faces(dim):
                    for v, p in zip(remap[verts], phys):
                        if p > 0 and np.all(v >= 0):
                            fv.append(v)
                            fl.append(p)
                if fv:
                    k = max(len(v) for v in fv)
                    if any(len(v) != k for v in fv):
                        raise ProblemError("boundary faces of mixed vertex counts are not supported")
                    m.set_boundary_labels(np.asarray(fv), np.asarray(fl))
                self.file_face_sets = set(int(v) for v in fl)
            else:
                coords, cells = self._box_mesh(dim, mesh["n_cells"], mesh["lengths"], mesh.get("simplex", False))
                m = core.Mesh(coords, cells)
                self.cell_sets = None
            self.mesh = m
            centres = coords[cells].mean(axis=1) if not has_file else self._centroids(coords, cells)
        self.bcs = self._boundary_conditions(bcs_spec, face_ids, mtype, gmsh is not None)
        if mtype == "structured":
            disc = core.Disc.structured(quad, n, L, self.bcs)
        else:
            disc = core.Disc.dg(quad, self.mesh, order, self.bcs)
        self.disc = disc
        # painting: background, then cell_sets, then paint boxes by centre
        mat = np.full(disc.n_cells, self.materials.resolve(regions.get("background", 0), "regions.background"),
                      dtype=np.int64)
        for key, ref in regions.get("cell_sets", {}).items():
            try:
                val = int(key)
            except ValueError:
                raise ProblemError(f"regions.cell_sets key {key!r} is not an integer")
            mid = self.materials.resolve(ref, f"regions.cell_sets[{key}]")
            if self.cell_sets is not None:
                mat[self.cell_sets == val] = mid
        for k, p in enumerate(regions.get("paint", [])):
            _strict(p, PAINT_KEYS, f"regions.paint[{k}]")
            mid = self.materials.resolve(p.get("material"), f"regions.paint[{k}]")
            shape = "interval" if dim == 1 else "box"
            other = "box" if dim == 1 else "interval"
            if other in p or shape not in p:
                raise ProblemError(f"regions.paint[{k}] needs a '{shape}' in {dim}D")
            b = p[shape]
            if len(b) != 2 * dim:
                raise ProblemError(f"regions.paint[{k}].{shape} must have {2 * dim} numbers")
            inside = np.ones(disc.n_cells, dtype=bool)
            for d in range(dim):
                inside &= (centres[:, d] >= b[2 * d]) & (centres[:, d] <= b[2 * d + 1])
            mat[inside] = mid
        self.material = mat
        M = self.materials
        G = M.G
        sigma_t = np.ascontiguousarray(M.sigma_t[mat].T)                 # [G, cells]
        sigma_s = np.ascontiguousarray(np.transpose(M.sigma_s[mat], (1, 2, 0)))  # [from, to, cells]
        source = np.ascontiguousarray(M.source[mat].T)
        self.upscatter = bool(np.any(np.tril(np.abs(M.sigma_s).max(axis=0), -1) > 0))
        from .problems import Problem
        self.problem = Problem(self.name, quad, disc, sigma_t, sigma_s, source, n_groups=G)

    @staticmethod
    def _grid_centres(n, L):
        axes = [(np.arange(k) + 0.5) * (l / k) for k, l in zip(n, L)]
        grids = np.meshgrid(*axes[::-1], indexing="ij")  # slowest axis first
        return np.stack([g.ravel() for g in grids[::-1]], axis=1)  # x fastest

    @staticmethod
    def _centroids(coords, cells):
        return coords[cells].mean(axis=1)

    @staticmethod
    def _box_mesh(dim, n, L, simplex):
        from . import problems as P
        if dim == 2:
            coords, cells = (P.box_tris if simplex else P.box_quads)(n[0], n[1], L[0], L[1])
        else:
            coords, cells = P.box_hexes(n[0], n[1], n[2], L[0], L[1], L[2])
            if simplex:  # six tetrahedra per hex along its 0-6 diagonal
                t = [(0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)]
                cells = np.concatenate([cells[:, list(x)] for x in t])
        return coords, np.ascontiguousarray(cells, dtype=np.int32)

    def _boundary_conditions(self, bcs_spec, face_ids, mtype, from_file):
        out = {}
        for key, val in bcs_spec.items():
            if key == "_comment":
                continue
            if key in face_ids:
                fid = face_ids[key]
            else:
                try:
                    fid = int(key)
                except ValueError:
                    raise ProblemError(f"unknown face {key!r}: use {sorted(face_ids)} or a 'Face Sets' value")
                if fid < 0:
                    raise ProblemError(f"face id {key!r} must be non-negative")
                if mtype == "structured":
                    raise ProblemError(f"an integer face key ({key}) is for unstructured meshes only")
            if fid in out:
                raise ProblemError(f"face {key!r} reaches 'Face Sets' value {fid} a second time")
            if isinstance(val, str):
                val = {"type": val}
            _strict(val, BC_KEYS, f"boundary_conditions[{key}]")
            typ = val.get("type")
            if typ not in ("vacuum", "reflect"):
                raise ProblemError(f"boundary_conditions[{key}].type must be 'vacuum' or 'reflect'")
            if typ == "reflect" and ("inflow" in val or "window" in val):
                raise ProblemError(f"boundary_conditions[{key}]: a reflective face takes no inflow or window")
            entry = {"type": typ}
            if "inflow" in val:
                entry["inflow"] = float(val["inflow"])
            if "window" in val:
                w = val["window"]
                if "inflow" not in val:
                    raise ProblemError(f"boundary_conditions[{key}]: a window without an inflow restricts nothing")
                if len(w) != 2 * (self.dim - 1) or any(w[2 * k] > w[2 * k + 1] for k in range(self.dim - 1)):
                    raise ProblemError(f"boundary_conditions[{key}].window must be {self.dim - 1} [lo, hi] pairs")
                entry["window"] = [float(v) for v in w]
            out[fid] = entry
        return out


def load(path):
    """the problem file at `path`, built"""
    return ProblemFile(path)


def solve(pf, rtol=None, mixed=False, dsa=False, max_it=500, outer_rtol=1e-8, max_outer=1000, inner_gamma=0.01,
          store_moments=False, recon_steps=10, dsa_void_bridge=True, stream=None):
    """The reference driver's group sweep (tests/transportk.kokkos.cxx:925-1208)
    in the library (ubt_solver_solve_groups): per group the rhs (ghost inflow +
    source + transfers from the groups solved so far), a right-preconditioned
    FGMRES solve, the scalar flux cached for the groups below; with upscatter
    the thermal block swept again until the estimated error of its scalar flux
    is below outer_rtol. rtol None: the solves run at 1e-8, or under upscatter
    at the reference test driver's 1e-10; a given rtol is every solve's (as
    -ksp_rtol).
    store_moments: the reference's -outer_store moments (one psi, each
    thermal group warm-started from the P1 expansion of its scalar flux and
    current, refined by at most recon_steps Richardson steps). Returns the
    angular flux per group (device tensors; under the moments store the one
    vector, the last group solved) and the sweep's report. dsa_void_bridge:
    the reference's -dsa_void_bridge (False masks the voids)."""
    from . import Solver, DSA

    prob = pf.problem
    op = prob.operator()
    d = DSA(op, void_bridge=dsa_void_bridge) if dsa else None
    sol = Solver(op, d, rtol=1e-8 if rtol is None else rtol, max_it=max_it, mixed=mixed, pc_dsa=dsa)
    psi, report = sol.solve_groups(outer_rtol=outer_rtol, max_outer=max_outer, inner_gamma=inner_gamma,
                                   inner_rtol=0.0 if rtol is None else rtol, store_moments=store_moments,
                                   recon_steps=recon_steps, stream=stream)
    report["operator"] = op
    if store_moments:
        return [psi], report
    return [psi[g] for g in range(prob.n_groups)], report

