"""ctypes binding of ``include/edgefem_amd.h`` plus torch plumbing for device memory.

PyTorch only supplies device allocations and the current HIP stream; every computation
runs in the HIP library.  Loading fails loudly when the library is missing: there is no
CPU fallback anywhere in the package.
"""

from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libedgefem_amd.so")

EFM_C128 = 0
EFM_C64 = 1
EFM_METHOD_COCG = 0
EFM_METHOD_BICGSTAB = 1

# every symbol include/edgefem_amd.h declares
EXPORTS = [
    "efm_last_error", "efm_version",
    "efm_topology_create", "efm_topology_destroy", "efm_topology_get_sizes", "efm_topology_export",
    "efm_topology_export_nodes", "efm_solver_set_gradient",
    "efm_assemble_km", "efm_assemble_maxwell", "efm_port_surface_mass", "efm_combine",
    "efm_solver_create", "efm_solver_destroy", "efm_solver_set_matrix", "efm_solver_solve",
    "efm_solver_spmv", "efm_solver_profile", "efm_solver_profile_read",
    "efm_port_overlap", "efm_solver_set_spmv_variant", "efm_solver_get_spmv_variant",
    "efm_sweep_create", "efm_sweep_destroy", "efm_sweep_solve", "efm_eps_gradient",
    "efm_solver_set_lowrank", "efm_add_diagonal", "efm_periodic_reduce", "efm_solver_set_gradient_complex",
    "efm_cells_create", "efm_cells_destroy", "efm_cells_solve", "efm_cells_set_spmv_variant",
    "efm_sweep_set_variants", "efm_sweep_get_variants", "efm_sweep_set_iteration", "efm_tet_fields",
    "efm_gmsh_read", "efm_gmsh_sizes", "efm_gmsh_copy", "efm_gmsh_free",
    "efm_touchstone_write", "efm_touchstone_write_2port", "efm_touchstone_extension",
    "efm_port_overlap_noalloc", "efm_sweep_set_graph", "efm_sweep_set_correction",
    "efm_sweep_set_correction_values", "efm_sweep_set_correction_terms", "efm_direct_solve",
    "efm_array_package_json", "efm_patterns_csv", "efm_coupling_csv", "efm_scan_csv",
]


class EdgeFemError(RuntimeError):
    pass


class TopologySizes(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("n_tets", C.c_int64), ("n_tris", C.c_int64),
                ("n_edges", C.c_int64), ("nnz", C.c_int64)]


class SolveOpts(C.Structure):
    _fields_ = [("method", C.c_int), ("tolerance", C.c_double), ("max_iterations", C.c_int),
                ("check_every", C.c_int)]


class RefineOpts(C.Structure):
    _fields_ = [("inner_type", C.c_int), ("replace_factor", C.c_double)]


class SolveResultC(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("max_rel_residual", C.c_double)]


_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    # torch first: it carries its own HIP runtime, and the library must bind to that one
    # (loading the library first would bring in a second runtime, whose allocations fail)
    _torch()
    if not os.path.exists(path):
        raise EdgeFemError(
            f"HIP library not found at {path}; build it with `python -m edgefem_amd.build` "
            "(there is no CPU fallback)")
    lib = C.CDLL(path)
    vp, i64, i32, dbl = C.c_void_p, C.c_int64, C.c_int, C.c_double
    lib.efm_last_error.restype = C.c_char_p
    lib.efm_version.restype = i32
    lib.efm_topology_create.argtypes = [i64, vp, i64, vp, i64, vp, vp, C.POINTER(vp)]
    lib.efm_topology_destroy.argtypes = [vp]
    lib.efm_topology_destroy.restype = None
    lib.efm_topology_get_sizes.argtypes = [vp, C.POINTER(TopologySizes)]
    lib.efm_topology_export.argtypes = [vp] + [vp] * 7 + [vp]
    lib.efm_topology_export_nodes.argtypes = [vp, vp, vp, vp, vp]
    lib.efm_solver_set_gradient.argtypes = [vp, i64, vp, vp, vp, vp, vp]
    lib.efm_assemble_km.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
    lib.efm_assemble_maxwell.argtypes = [vp, vp, vp, vp, vp, vp, dbl, vp, vp]
    lib.efm_port_surface_mass.argtypes = [vp, vp, i64, vp, vp, vp, vp]
    lib.efm_combine.argtypes = [vp, vp, vp, dbl, i32, vp, vp, i32, vp, vp]
    lib.efm_solver_create.argtypes = [i64, i64, vp, vp, i32, i32, vp, C.POINTER(vp)]
    lib.efm_solver_destroy.argtypes = [vp]
    lib.efm_solver_destroy.restype = None
    lib.efm_solver_set_matrix.argtypes = [vp, vp, vp]
    lib.efm_solver_solve.argtypes = [vp, vp, vp, C.POINTER(SolveOpts), C.POINTER(SolveResultC),
                                     vp, vp, vp]
    lib.efm_solver_spmv.argtypes = [vp, vp, vp, vp]
    lib.efm_solver_profile.argtypes = [vp, i32]
    lib.efm_solver_profile_read.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    lib.efm_solver_set_spmv_variant.argtypes = [vp, i32]
    lib.efm_solver_get_spmv_variant.argtypes = [vp, C.POINTER(C.c_int), vp]
    lib.efm_sweep_create.argtypes = [vp, vp, vp, i32, vp, vp, i32, vp, C.POINTER(vp)]
    lib.efm_solver_set_lowrank.argtypes = [vp, i32, C.c_int64, vp, vp, vp, vp]
    lib.efm_add_diagonal.argtypes = [vp, C.c_int64, vp, vp, vp, vp]
    lib.efm_solver_set_gradient_complex.argtypes = [vp, C.c_int64, vp, vp, vp]
    lib.efm_cells_create.argtypes = [C.c_int64, vp, vp, vp, vp, C.c_int64, vp, vp, vp, i32, vp, vp, vp, vp, vp,
                                     C.POINTER(vp)]
    lib.efm_cells_destroy.argtypes = [vp]
    lib.efm_cells_destroy.restype = None
    lib.efm_cells_set_spmv_variant.argtypes = [vp, i32]
    lib.efm_cells_solve.argtypes = [vp, vp, vp, C.POINTER(SolveOpts), vp, vp, vp]
    lib.efm_periodic_reduce.argtypes = [C.c_int64, vp, vp, vp, C.c_int64, vp, vp, vp, vp, vp, vp, vp, vp,
                                        C.c_int64, C.POINTER(C.c_int64), vp]
    lib.efm_eps_gradient.argtypes = [vp, vp, C.c_int64, vp, i32, vp, i32, vp, i32, vp, vp, vp, vp, vp]
    lib.efm_sweep_destroy.argtypes = [vp]
    lib.efm_sweep_destroy.restype = None
    lib.efm_sweep_set_variants.argtypes = [vp, i32, i32]
    lib.efm_sweep_set_iteration.argtypes = [vp, i32]
    lib.efm_sweep_set_correction.argtypes = [vp, i64, vp, vp, vp, vp]
    lib.efm_sweep_set_correction_values.argtypes = [vp, i32, i32, vp, vp, vp]
    lib.efm_sweep_set_correction_terms.argtypes = [vp, i32, vp, i32, vp, vp]
    lib.efm_direct_solve.argtypes = [i64, vp, vp, vp, i32, i64, vp, vp, vp, vp, vp, i32, C.POINTER(C.c_int), vp]
    lib.efm_tet_fields.argtypes = [vp, vp, i32, vp, vp, vp, vp, vp]
    lib.efm_gmsh_read.argtypes = [C.c_char_p, i32, C.POINTER(vp)]
    lib.efm_gmsh_sizes.argtypes = [vp, vp]
    lib.efm_gmsh_copy.argtypes = [vp] * 11
    lib.efm_gmsh_free.argtypes = [vp]
    lib.efm_gmsh_free.restype = None
    lib.efm_touchstone_write.argtypes = [C.c_char_p, i64, vp, i32, vp, i32, C.c_double]
    lib.efm_touchstone_write_2port.argtypes = [C.c_char_p, i64, vp, vp]
    lib.efm_touchstone_extension.argtypes = [i32, C.c_char_p, i32]
    lib.efm_sweep_get_variants.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), vp]
    lib.efm_sweep_solve.argtypes = [vp, i32, vp, vp, vp, vp, C.POINTER(SolveOpts), C.POINTER(RefineOpts), vp, vp,
                                    vp]
    lib.efm_port_overlap.argtypes = [i64, i32, i64, vp, vp, vp, i32, vp, vp]
    lib.efm_array_package_json.argtypes = [C.c_char_p, C.c_char_p, C.c_double, i64, vp, i64, vp, i64, i32, vp, i64,
                                           i64, vp, vp, vp]
    lib.efm_patterns_csv.argtypes = [C.c_char_p, i64, vp, vp, vp]
    lib.efm_coupling_csv.argtypes = [C.c_char_p, i32, vp, vp]
    lib.efm_scan_csv.argtypes = [C.c_char_p, i64, vp, vp, vp, vp, vp]
    lib.efm_port_overlap_noalloc.argtypes = [i64, i32, i64, vp, vp, vp, i32, vp, vp]
    for name in EXPORTS:
        getattr(lib, name)  # AttributeError if the library lacks a declared symbol
    _lib = lib
    return lib


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = _lib.efm_last_error().decode(errors="replace") if _lib is not None else ""
        raise EdgeFemError(f"{what} failed (status {rc}): {msg}")


# ---------------------------------------------------------------------------
# torch plumbing
# ---------------------------------------------------------------------------
def _torch():
    import torch
    return torch


def device() -> "torch.device":
    torch = _torch()
    if not torch.cuda.is_available():
        raise EdgeFemError("no HIP device visible: the EdgeFEM MI355X path runs only on the GPU")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr() -> int:
    torch = _torch()
    return torch.cuda.current_stream().cuda_stream


def dptr(t) -> Optional[int]:
    if t is None:
        return None
    if not t.is_cuda:
        raise EdgeFemError("expected a device tensor")
    if not t.is_contiguous():
        raise EdgeFemError("expected a contiguous tensor")
    if t.is_conj() or t.is_neg():
        raise EdgeFemError("expected a materialised tensor (a lazy conj/neg view keeps the raw values in memory)")
    return t.data_ptr()


def to_dev(a, dtype=None):
    """numpy / torch -> contiguous device tensor (a lazily conjugated or negated torch view
    is materialised: the C ABI reads the raw memory)."""
    torch = _torch()
    if isinstance(a, torch.Tensor):
        t = a.resolve_conj().resolve_neg()
    else:
        t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(device()).contiguous()


# ---------------------------------------------------------------------------
# Topology
# ---------------------------------------------------------------------------
class Topology:
    """Device-resident edges, signs, CSR pattern and gather maps of one mesh."""

    def __init__(self, node_ids, tet_nodes, tri_nodes):
        torch = _torch()
        lib = load_library()
        self._lib = lib
        ids = to_dev(node_ids, torch.int64)
        tets = to_dev(np.asarray(tet_nodes).reshape(-1, 4) if not isinstance(tet_nodes, torch.Tensor) else tet_nodes, torch.int32)
        tris = to_dev(np.asarray(tri_nodes).reshape(-1, 3) if not isinstance(tri_nodes, torch.Tensor) else tri_nodes, torch.int32)
        h = C.c_void_p()
        _check(lib.efm_topology_create(ids.numel(), dptr(ids), tets.numel() // 4, dptr(tets) if tets.numel() else None,
                                       tris.numel() // 3, dptr(tris) if tris.numel() else None,
                                       stream_ptr(), C.byref(h)), "efm_topology_create")
        self._h = h
        sz = TopologySizes()
        _check(lib.efm_topology_get_sizes(h, C.byref(sz)), "efm_topology_get_sizes")
        self.n_nodes, self.n_tets, self.n_tris = sz.n_nodes, sz.n_tets, sz.n_tris
        self.n_edges, self.nnz = sz.n_edges, sz.nnz
        dev = device()
        self.edges = torch.empty((self.n_edges, 2), dtype=torch.int64, device=dev)
        self.tet_edges = torch.empty((self.n_tets, 6), dtype=torch.int32, device=dev)
        self.tet_sign = torch.empty((self.n_tets, 6), dtype=torch.int8, device=dev)
        self.tri_edges = torch.empty((self.n_tris, 3), dtype=torch.int32, device=dev)
        self.tri_sign = torch.empty((self.n_tris, 3), dtype=torch.int8, device=dev)
        self.row_ptr = torch.empty(self.n_edges + 1, dtype=torch.int32, device=dev)
        self.col_idx = torch.empty(self.nnz, dtype=torch.int32, device=dev)
        _check(lib.efm_topology_export(h, dptr(self.edges), dptr(self.tet_edges), dptr(self.tet_sign),
                                       dptr(self.tri_edges) if self.n_tris else None,
                                       dptr(self.tri_sign) if self.n_tris else None,
                                       dptr(self.row_ptr), dptr(self.col_idx), stream_ptr()),
               "efm_topology_export")
        self.edge_nodes = torch.empty((self.n_edges, 2), dtype=torch.int32, device=dev)
        self.n2e_ptr = torch.empty(self.n_nodes + 1, dtype=torch.int32, device=dev)
        self.n2e = torch.empty(2 * self.n_edges, dtype=torch.int32, device=dev)
        _check(lib.efm_topology_export_nodes(h, dptr(self.edge_nodes), dptr(self.n2e_ptr), dptr(self.n2e),
                                             stream_ptr()), "efm_topology_export_nodes")

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.efm_topology_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def assemble_km(topo: Topology, xyz, eps_tet, inv_mu_tet=None, dirichlet=None):
    torch = _torch()
    lib = load_library()
    x = to_dev(xyz, torch.float64)
    eps = to_dev(eps_tet, torch.complex128)
    imu = None if inv_mu_tet is None else to_dev(inv_mu_tet, torch.complex128)
    dm = None if dirichlet is None else to_dev(dirichlet, torch.uint8)
    K = torch.empty(topo.nnz, dtype=torch.complex128, device=x.device)
    M = torch.empty(topo.nnz, dtype=torch.complex128, device=x.device)
    _check(lib.efm_assemble_km(topo.handle, dptr(x), dptr(eps), dptr(imu), dptr(dm), dptr(K), dptr(M),
                               stream_ptr()), "efm_assemble_km")
    return K, M


def assemble_maxwell(topo: Topology, xyz, eps_tet, inv_mu_tet, dirichlet, abc_mask, omega):
    torch = _torch()
    lib = load_library()
    x = to_dev(xyz, torch.float64)
    eps = to_dev(eps_tet, torch.complex128)
    imu = None if inv_mu_tet is None else to_dev(inv_mu_tet, torch.complex128)
    dm = None if dirichlet is None else to_dev(dirichlet, torch.uint8)
    am = None if abc_mask is None else to_dev(abc_mask, torch.uint8)
    A = torch.empty(topo.nnz, dtype=torch.complex128, device=x.device)
    _check(lib.efm_assemble_maxwell(topo.handle, dptr(x), dptr(eps), dptr(imu), dptr(dm), dptr(am),
                                    float(omega), dptr(A), stream_ptr()), "efm_assemble_maxwell")
    return A


def port_surface_mass(topo: Topology, xyz, tri_sel, dirichlet=None):
    torch = _torch()
    lib = load_library()
    x = to_dev(xyz, torch.float64)
    sel = to_dev(np.asarray(tri_sel, np.int32) if not isinstance(tri_sel, torch.Tensor) else tri_sel, torch.int32)
    dm = None if dirichlet is None else to_dev(dirichlet, torch.uint8)
    Ms = torch.empty(topo.nnz, dtype=torch.float64, device=x.device)
    _check(lib.efm_port_surface_mass(topo.handle, dptr(x), sel.numel(), dptr(sel) if sel.numel() else None,
                                     dptr(dm), dptr(Ms), stream_ptr()), "efm_port_surface_mass")
    return Ms


def combine(topo: Topology, K, M, k0sq: float, Ms_list: Sequence = (), coefs: Sequence[complex] = (),
            out_dtype=None):
    torch = _torch()
    lib = load_library()
    out_dtype = out_dtype or torch.complex128
    out_type = EFM_C128 if out_dtype == torch.complex128 else EFM_C64
    A = torch.empty(topo.nnz, dtype=out_dtype, device=K.device)
    npo = len(Ms_list)
    ptrs = (C.c_void_p * max(npo, 1))(*[dptr(m) for m in Ms_list])
    co = np.zeros(2 * max(npo, 1), np.float64)
    for i, c in enumerate(coefs):
        co[2 * i] = complex(c).real
        co[2 * i + 1] = complex(c).imag
    _check(lib.efm_combine(topo.handle, dptr(K), dptr(M), float(k0sq), npo, ptrs,
                           co.ctypes.data_as(C.c_void_p), out_type, dptr(A), stream_ptr()), "efm_combine")
    return A


# ---------------------------------------------------------------------------
# Solver
# ---------------------------------------------------------------------------
class KrylovSolver:
    """Jacobi-preconditioned COCG / BiCGSTAB on a fixed CSR pattern, nrhs systems."""

    def __init__(self, row_ptr, col_idx, nrhs: int = 1, dtype=None):
        torch = _torch()
        lib = load_library()
        self._lib = lib
        self.row_ptr = to_dev(row_ptr, torch.int32)
        self.col_idx = to_dev(col_idx, torch.int32)
        self.n = self.row_ptr.numel() - 1
        self.nnz = self.col_idx.numel()
        self.nrhs = int(nrhs)
        self.dtype = dtype or torch.complex128
        vt = EFM_C128 if self.dtype == torch.complex128 else EFM_C64
        h = C.c_void_p()
        _check(lib.efm_solver_create(self.n, self.nnz, dptr(self.row_ptr), dptr(self.col_idx), self.nrhs, vt,
                                     stream_ptr(), C.byref(h)), "efm_solver_create")
        self._h = h

    def set_gradient(self, topo: "Topology", dirichlet=None):
        """Enable the auxiliary gradient preconditioner for a mesh-backed system (call
        before set_matrix)."""
        torch = _torch()
        self._aux_keep = (topo, None if dirichlet is None else to_dev(dirichlet, torch.uint8))
        dm = self._aux_keep[1]
        _check(self._lib.efm_solver_set_gradient(self._h, topo.n_nodes, dptr(topo.edge_nodes), dptr(topo.n2e_ptr),
                                                 dptr(topo.n2e), dptr(dm), stream_ptr()), "efm_solver_set_gradient")

    def set_gradient_complex(self, n_nodes: int, edge_nodes, edge_coefs):
        """Auxiliary preconditioner with a complex incidence (call before set_matrix):
        edge_nodes (n, 2) int32 node indices (-1 = none), edge_coefs (n, 2) complex128."""
        en = np.ascontiguousarray(np.asarray(edge_nodes, np.int32).reshape(-1))
        ec = np.ascontiguousarray(np.asarray(edge_coefs, np.complex128).reshape(-1))
        if en.size != 2 * self.n or ec.size != 2 * self.n:
            raise EdgeFemError("gradient arrays must have two entries per edge")
        _check(self._lib.efm_solver_set_gradient_complex(self._h, int(n_nodes), en.ctypes.data_as(C.c_void_p),
                                                         ec.ctypes.data_as(C.c_void_p), stream_ptr()),
               "efm_solver_set_gradient_complex")

    def set_lowrank(self, gidx, gval, coefs):
        """Operator term A_eff = A + sum_p coefs[p] w_p w_p^H (call before set_matrix).
        gidx: int32 device (P*len, -1 padding), gval: complex128 device (P*len)."""
        torch = _torch()
        c = np.ascontiguousarray(np.asarray(coefs, np.complex128))
        gi = to_dev(gidx, torch.int32)
        gv = to_dev(gval, torch.complex128)
        P = c.size
        L = gi.numel() // max(P, 1)
        self._lr_keep = (gi, gv)
        _check(self._lib.efm_solver_set_lowrank(self._h, P, L, dptr(gi), dptr(gv), c.ctypes.data_as(C.c_void_p),
                                                stream_ptr()), "efm_solver_set_lowrank")

    def set_matrix(self, values):
        torch = _torch()
        v = to_dev(values, self.dtype)
        if v.numel() != self.nnz:
            raise EdgeFemError(f"matrix has {v.numel()} values, pattern has {self.nnz}")
        _check(self._lib.efm_solver_set_matrix(self._h, dptr(v), stream_ptr()), "efm_solver_set_matrix")

    def solve(self, B, tol=1e-10, max_iterations=10000, method=EFM_METHOD_COCG, check_every=32, out=None):
        """B: (n, nrhs) or (n,) complex128 device tensor.  Returns (X, info)."""
        torch = _torch()
        Bd = to_dev(B, torch.complex128)
        if Bd.numel() != self.n * self.nrhs:
            raise EdgeFemError("right-hand side has the wrong size")
        X = out if out is not None else torch.empty_like(Bd)
        opts = SolveOpts(method, float(tol), int(max_iterations), int(check_every))
        res = SolveResultC()
        iters = np.zeros(self.nrhs, np.int32)
        resid = np.zeros(self.nrhs, np.float64)
        _check(self._lib.efm_solver_solve(self._h, dptr(Bd), dptr(X), C.byref(opts), C.byref(res),
                                          iters.ctypes.data_as(C.c_void_p), resid.ctypes.data_as(C.c_void_p),
                                          stream_ptr()), "efm_solver_solve")
        info = dict(iterations=res.iterations, converged=bool(res.converged),
                    max_rel_residual=res.max_rel_residual, iters=iters, resid=resid)
        return X, info

    def spmv(self, x):
        torch = _torch()
        xd = to_dev(x, torch.complex128)
        y = torch.empty_like(xd)
        _check(self._lib.efm_solver_spmv(self._h, dptr(xd), dptr(y), stream_ptr()), "efm_solver_spmv")
        return y

    def set_spmv_variant(self, variant: int):
        _check(self._lib.efm_solver_set_spmv_variant(self._h, int(variant)), "efm_solver_set_spmv_variant")

    def spmv_variant(self):
        v = C.c_int()
        ms = np.zeros(12, np.float64)
        _check(self._lib.efm_solver_get_spmv_variant(self._h, C.byref(v), ms.ctypes.data_as(C.c_void_p)),
               "efm_solver_get_spmv_variant")
        return v.value, ms

    def profile(self, enable=True):
        _check(self._lib.efm_solver_profile(self._h, 1 if enable else 0), "efm_solver_profile")

    def profile_read(self):
        ms = C.c_double()
        n = C.c_int64()
        _check(self._lib.efm_solver_profile_read(self._h, C.byref(ms), C.byref(n)), "efm_solver_profile_read")
        return ms.value, n.value

    def close(self):
        if getattr(self, "_h", None):
            self._lib.efm_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def port_overlap(gidx, gval, X, nrhs: int):
    """out[j, a] = g_j^H x_a for j over ports (rows of gidx/gval)."""
    torch = _torch()
    lib = load_library()
    gi = to_dev(gidx, torch.int32)
    gv = to_dev(gval, torch.complex128)
    P, L = gi.shape
    out = np.zeros(P * nrhs, np.complex128)
    _check(lib.efm_port_overlap(X.numel() // nrhs, P, L, dptr(gi), dptr(gv), dptr(X), int(nrhs),
                                out.ctypes.data_as(C.c_void_p), stream_ptr()), "efm_port_overlap")
    return out.reshape(P, nrhs)


class SweepSolver:
    """Batched COCG over many systems K - k0sq[s] M + sum_p coef[s,p] M_s,p sharing one
    pattern (vectors [n_edges][nsys])."""

    def __init__(self, topo: Topology, K, M, Ms_list: Sequence, dirichlet, max_systems: int):
        torch = _torch()
        lib = load_library()
        self._lib = lib
        self._keep = (topo, K, M, list(Ms_list), None if dirichlet is None else to_dev(dirichlet, torch.uint8))
        P = len(Ms_list)
        ptrs = (C.c_void_p * max(P, 1))(*[dptr(m) for m in Ms_list])
        h = C.c_void_p()
        _check(lib.efm_sweep_create(topo.handle, dptr(K), dptr(M), P, ptrs, dptr(self._keep[4]), int(max_systems),
                                    stream_ptr(), C.byref(h)), "efm_sweep_create")
        self._h = h
        self.n = topo.n_edges
        self.P = P
        self.max_systems = int(max_systems)

    def solve(self, k0sq, coef, B, tol=1e-10, max_iterations=10000, check_every=16, precision="fp64",
              replace_factor=0.01):
        """k0sq (nsys,), coef (nsys, P) complex, B (n, nsys) complex128 device tensor.
        precision "fp64": one FP64 COCG solve; "fp32": COCG on complex-FP32 vectors with
        FP64 reliable updates (the residual replaced by b - A x in FP64 each time it drops by
        replace_factor; convergence on the true residual <= tol); "fp64-refined": the same
        with FP64 vectors."""
        torch = _torch()
        k2 = np.ascontiguousarray(np.asarray(k0sq, np.float64))
        nsys = k2.size
        cf = np.ascontiguousarray(np.asarray(coef, np.complex128).reshape(nsys, self.P))
        Bd = to_dev(B, torch.complex128)
        if Bd.numel() != self.n * nsys:
            raise EdgeFemError("right-hand side block has the wrong size")
        X = torch.empty_like(Bd)
        iters = np.zeros(nsys, np.int32)
        resid = np.zeros(nsys, np.float64)
        opts = SolveOpts(EFM_METHOD_COCG, float(tol), int(max_iterations), int(check_every))
        if precision == "fp64":
            ref = None
        elif precision in ("fp32", "fp64-refined"):
            ref = C.byref(RefineOpts(EFM_C64 if precision == "fp32" else EFM_C128, float(replace_factor)))
        else:
            raise EdgeFemError(f"unknown precision {precision!r}")
        _check(self._lib.efm_sweep_solve(self._h, nsys, k2.ctypes.data_as(C.c_void_p), cf.ctypes.data_as(C.c_void_p),
                                         dptr(Bd), dptr(X), C.byref(opts), ref, iters.ctypes.data_as(C.c_void_p),
                                         resid.ctypes.data_as(C.c_void_p), stream_ptr()), "efm_sweep_solve")
        return X, dict(iters=iters, resid=resid)

    def set_variants(self, spmv: int = -1, aux: int = -1) -> None:
        """Kernel variants of the SpMV (0 one block slice per group of 64 systems, 1
        pipelined, groups in turn, 2 LDS-tiled: each tile of 8 rows loads its distinct
        columns once per group; only where every tile fits) and of the nodal
        preconditioner part (0, 1 as for the SpMV); -1: the fastest, timed on the first
        solve."""
        _check(self._lib.efm_sweep_set_variants(self._h, int(spmv), int(aux)), "efm_sweep_set_variants")

    def set_graph(self, enable: bool = True) -> None:
        """Replay each full convergence-check chunk of iterations from a HIP graph captured
        once per solve (standard forms 0 and 2)."""
        _check(self._lib.efm_sweep_set_graph(self._h, 1 if enable else 0), "efm_sweep_set_graph")

    def set_correction(self, rows, ptr, cols) -> None:
        """The sub-pattern of a per-system correction C_s (frequency-dependent media):
        correction row m is matrix row rows[m], with entries ptr[m]..ptr[m+1]-1 in columns
        cols.  Empty rows remove the correction."""
        r = np.ascontiguousarray(np.asarray(rows, np.int32))
        pp = np.ascontiguousarray(np.asarray(ptr, np.int32))
        cc = np.ascontiguousarray(np.asarray(cols, np.int32))
        if r.size and pp.size != r.size + 1:
            raise EdgeFemError("correction ptr must have len(rows) + 1 entries")
        self._corr_vals = None
        _check(self._lib.efm_sweep_set_correction(self._h, int(r.size), r.ctypes.data_as(C.c_void_p),
                                                  pp.ctypes.data_as(C.c_void_p), cc.ctypes.data_as(C.c_void_p),
                                                  stream_ptr()), "efm_sweep_set_correction")

    def set_correction_terms(self, terms, alpha) -> None:
        """Shared correction terms: terms a complex128 device tensor (entries, 8) (columns
        past len(alpha[0]) zero), alpha (nsys, nterms) complex weights per system."""
        torch = _torch()
        t = to_dev(terms, torch.complex128)
        al = np.ascontiguousarray(np.asarray(alpha, np.complex128))
        if t.dim() != 2 or t.shape[1] != 8 or al.ndim != 2:
            raise EdgeFemError("correction terms must be (entries, 8) and weights (nsys, nterms)")
        self._corr_vals = t
        _check(self._lib.efm_sweep_set_correction_terms(self._h, int(al.shape[1]), dptr(t), int(al.shape[0]),
                                                        al.ctypes.data_as(C.c_void_p), stream_ptr()),
               "efm_sweep_set_correction_terms")

    def set_correction_values(self, vals, col_of_system) -> None:
        """Values of the corrections: a complex128 device tensor (entries, ncols), kept
        alive here until replaced; system s of the later solves uses column
        col_of_system[s] (their count is the solves' number of systems)."""
        torch = _torch()
        v = to_dev(vals, torch.complex128)
        if v.dim() != 2:
            raise EdgeFemError("correction values must be (entries, ncols)")
        cm = np.ascontiguousarray(np.asarray(col_of_system, np.int32))
        self._corr_vals = v
        _check(self._lib.efm_sweep_set_correction_values(self._h, int(cm.size), int(v.shape[1]), dptr(v),
                                                         cm.ctypes.data_as(C.c_void_p), stream_ptr()),
               "efm_sweep_set_correction_values")

    def set_iteration(self, form: int = 2) -> None:
        """Iteration form of the batched COCG: 0 standard, with p = u + beta p in a pass of
        its own and two reductions per iteration; 1 Chronopoulos-Gear, one reduction per
        iteration (s = A p carried by a recurrence, which costs iterations on indefinite
        systems); 2 (the library's default) standard with u = D^-1 r + G y rebuilt in the
        p pass instead of stored."""
        _check(self._lib.efm_sweep_set_iteration(self._h, int(form)), "efm_sweep_set_iteration")

    def variants(self):
        """(SpMV variant, nodal variant, mean ms per launch over the tuning iterations
        [SpMV 0, 1, 2, nodal 0, 1]; 0 for a variant that was not timed)."""
        v, a = C.c_int(), C.c_int()
        ms = np.zeros(5, np.float64)
        _check(self._lib.efm_sweep_get_variants(self._h, C.byref(v), C.byref(a), ms.ctypes.data_as(C.c_void_p)),
               "efm_sweep_get_variants")
        return v.value, a.value, ms

    def close(self):
        if getattr(self, "_h", None):
            self._lib.efm_sweep_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gmsh_read(path: str, n_threads: int = 0) -> dict:
    """gmsh v2.2 ASCII $Nodes/$Elements (efm_gmsh_read, host code, n_threads 0: up to 16).
    Returns host arrays: node_ids (N,), xyz (N,3), tet_ids (T,4) / tri_ids (F,3) / line_ids
    (L,2) as gmsh node IDs, tet_phys, tri_phys, line_phys, tet_elem, tri_elem."""
    lib = load_library()
    h = C.c_void_p()
    _check(lib.efm_gmsh_read(str(path).encode(), int(n_threads), C.byref(h)), "efm_gmsh_read")
    try:
        sz = np.zeros(4, np.int64)
        _check(lib.efm_gmsh_sizes(h, sz.ctypes.data_as(C.c_void_p)), "efm_gmsh_sizes")
        N, T, F, L = (int(v) for v in sz)
        out = dict(node_ids=np.empty(N, np.int64), xyz=np.empty((N, 3)), tet_ids=np.empty((T, 4), np.int64),
                   tet_phys=np.empty(T, np.int32), tet_elem=np.empty(T, np.int64), tri_ids=np.empty((F, 3), np.int64),
                   tri_phys=np.empty(F, np.int32), tri_elem=np.empty(F, np.int64),
                   line_ids=np.empty((L, 2), np.int64), line_phys=np.empty(L, np.int32))
        order = ("node_ids", "xyz", "tet_ids", "tet_phys", "tet_elem", "tri_ids", "tri_phys", "tri_elem",
                 "line_ids", "line_phys")
        _check(lib.efm_gmsh_copy(h, *[out[k].ctypes.data_as(C.c_void_p) for k in order]), "efm_gmsh_copy")
    finally:
        lib.efm_gmsh_free(h)
    return out


def touchstone_write(path: str, freq, S, fmt: int, z0: float) -> None:
    """efm_touchstone_write: freq (F,) float64, S (F, P, P) complex128, host arrays."""
    lib = load_library()
    f = np.ascontiguousarray(freq, np.float64).reshape(-1)
    s = np.ascontiguousarray(S, np.complex128)
    P = s.shape[1] if s.ndim == 3 else 0
    _check(lib.efm_touchstone_write(str(path).encode(), f.size, f.ctypes.data_as(C.c_void_p), P,
                                    s.ctypes.data_as(C.c_void_p), int(fmt), float(z0)), "efm_touchstone_write")


def _vp(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def array_package_json(path: str, model_name: str, z0: float, positions, freqs, S, n_pattern_freqs: int,
                       scan_theta, scan_phi, scan_loss_db) -> None:
    """efm_array_package_json: positions (N, 3), freqs (F,), S (F_S, P, P) complex128, scan
    arrays (K,) in radians / dB; host arrays."""
    lib = load_library()
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1, 3)
    fr = np.ascontiguousarray(freqs, np.float64).reshape(-1)
    s = np.ascontiguousarray(S, np.complex128)
    nS, P = (s.shape[0], s.shape[1]) if s.ndim == 3 else (0, 0)
    th = np.ascontiguousarray(scan_theta, np.float64).reshape(-1)
    ph = np.ascontiguousarray(scan_phi, np.float64).reshape(-1)
    sl = np.ascontiguousarray(scan_loss_db, np.float64).reshape(-1)
    _check(lib.efm_array_package_json(str(path).encode(), str(model_name).encode(), float(z0), pos.shape[0], _vp(pos),
                                      fr.size, _vp(fr), nS, P, _vp(s), int(n_pattern_freqs), th.size, _vp(th),
                                      _vp(ph), _vp(sl)), "efm_array_package_json")


def patterns_csv(path: str, port_index, npts, rows) -> None:
    """efm_patterns_csv: port_index (K,) int32, npts (K,) int64, rows (sum npts, 5)."""
    lib = load_library()
    pi = np.ascontiguousarray(port_index, np.int32).reshape(-1)
    npt = np.ascontiguousarray(npts, np.int64).reshape(-1)
    r = np.ascontiguousarray(rows, np.float64).reshape(-1, 5)
    _check(lib.efm_patterns_csv(str(path).encode(), pi.size, _vp(pi), _vp(npt), _vp(r)), "efm_patterns_csv")


def coupling_csv(path: str, S, positions=None) -> None:
    """efm_coupling_csv: S (P, P) complex128, positions (P, 3) or None."""
    lib = load_library()
    s = np.ascontiguousarray(S, np.complex128)
    pos = None if positions is None else np.ascontiguousarray(positions, np.float64).reshape(-1, 3)
    _check(lib.efm_coupling_csv(str(path).encode(), s.shape[0], _vp(s), _vp(pos)), "efm_coupling_csv")


def scan_csv(path: str, theta, phi, nelem, gamma, Z) -> None:
    """efm_scan_csv: theta, phi (K,) radians, nelem (K,) int64, gamma and Z concatenated complex128."""
    lib = load_library()
    th = np.ascontiguousarray(theta, np.float64).reshape(-1)
    ph = np.ascontiguousarray(phi, np.float64).reshape(-1)
    ne = np.ascontiguousarray(nelem, np.int64).reshape(-1)
    g = np.ascontiguousarray(gamma, np.complex128).reshape(-1)
    z = np.ascontiguousarray(Z, np.complex128).reshape(-1)
    _check(lib.efm_scan_csv(str(path).encode(), th.size, _vp(th), _vp(ph), _vp(ne), _vp(g), _vp(z)), "efm_scan_csv")


def touchstone_write_2port(path: str, freq, S4) -> None:
    """efm_touchstone_write_2port: S4 (F, 4) complex128 as s11, s21, s12, s22."""
    lib = load_library()
    f = np.ascontiguousarray(freq, np.float64).reshape(-1)
    s = np.ascontiguousarray(S4, np.complex128).reshape(-1, 4)
    _check(lib.efm_touchstone_write_2port(str(path).encode(), f.size, f.ctypes.data_as(C.c_void_p),
                                          s.ctypes.data_as(C.c_void_p)), "efm_touchstone_write_2port")


def touchstone_extension(num_ports: int) -> str:
    lib = load_library()
    buf = C.create_string_buffer(16)
    _check(lib.efm_touchstone_extension(int(num_ports), buf, 16), "efm_touchstone_extension")
    return buf.value.decode()


def tet_fields(topo: Topology, xyz, X, lam=None, with_curl: bool = True):
    """E at a point of every tet (barycentric weights lam (T, 4), None: the centroid) and
    curl E (efm_tet_fields).  X: device complex128 (n,) or (n, R), R in 1, 2, 3, 4, 8.
    Returns (E, curl) as complex128 device tensors (T, R, 3); curl is None without
    with_curl."""
    torch = _torch()
    lib = load_library()
    Xd = to_dev(X, torch.complex128)
    R = Xd.shape[1] if Xd.dim() == 2 else 1
    if Xd.shape[0] != topo.n_edges:
        raise EdgeFemError("solution length differs from the number of edges")
    T = topo.n_tets
    ld = None
    if lam is not None:
        ld = to_dev(lam, torch.float64)
        if tuple(ld.shape) != (T, 4):
            raise EdgeFemError("barycentric weights must be (n_tets, 4)")
    E = torch.empty((T, R, 3), dtype=torch.complex128, device=Xd.device)
    C = torch.empty((T, R, 3), dtype=torch.complex128, device=Xd.device) if with_curl else None
    _check(lib.efm_tet_fields(topo.handle, dptr(xyz), int(R), dptr(Xd), dptr(ld), dptr(E), dptr(C), stream_ptr()),
           "efm_tet_fields")
    return E, C


def eps_gradient(topo: Topology, xyz, design, Lam, X, terms_l, terms_r, weights):
    """grad[d] = sum_t Re(w_t * lam_{l_t}^T M_e x_{r_t}) over design tets (efm_eps_gradient).
    design: int array of tet indices (host or device) or None for all tets; Lam, X: device
    complex128 (n, nl) / (n, nx) system-minor; returns a float64 device tensor."""
    torch = _torch()
    lib = load_library()
    if design is None:
        dd = None
        D = topo.n_tets
    else:
        dh = np.asarray(design.cpu().numpy() if hasattr(design, "cpu") else design, np.int64)
        if dh.size and (dh.min() < 0 or dh.max() >= topo.n_tets):
            raise EdgeFemError("design tet index out of range")
        dd = to_dev(dh.astype(np.int32), torch.int32)
        D = dh.size
    Ld = to_dev(Lam, torch.complex128)
    Xd = to_dev(X, torch.complex128)
    nl = Ld.shape[1] if Ld.dim() == 2 else 1
    nx = Xd.shape[1] if Xd.dim() == 2 else 1
    tl = np.ascontiguousarray(np.asarray(terms_l, np.int32))
    tr = np.ascontiguousarray(np.asarray(terms_r, np.int32))
    w = np.ascontiguousarray(np.asarray(weights, np.complex128))
    if not (tl.size == tr.size == w.size):
        raise EdgeFemError("term arrays differ in length")
    out = torch.empty(D, dtype=torch.float64, device=Xd.device)
    _check(lib.efm_eps_gradient(topo.handle, dptr(xyz), int(D), dptr(dd), int(nl), dptr(Ld), int(nx), dptr(Xd),
                                int(tl.size), tl.ctypes.data_as(C.c_void_p), tr.ctypes.data_as(C.c_void_p),
                                w.ctypes.data_as(C.c_void_p), dptr(out), stream_ptr()), "efm_eps_gradient")
    return out


def add_diagonal(topo: Topology, rows, vals, A):
    """A[r, r] += vals[i] for the distinct rows[i] (in place on the device values A)."""
    torch = _torch()
    lib = load_library()
    r = np.ascontiguousarray(np.asarray(rows, np.int32))
    if r.size == 0:
        return A
    if np.unique(r).size != r.size or r.min() < 0 or r.max() >= topo.n_edges:
        raise EdgeFemError("add_diagonal: rows must be distinct edge indices")
    rd = to_dev(r, torch.int32)
    vd = to_dev(np.asarray(vals, np.complex128).reshape(-1), torch.complex128)
    _check(lib.efm_add_diagonal(topo.handle, int(r.size), dptr(rd), dptr(vd), dptr(A), stream_ptr()),
           "efm_add_diagonal")
    return A


def periodic_reduce(row_ptr, col_idx, values, master, slave, coef_row, coef_col, dirichlet=None):
    """Sparse periodic elimination (efm_periodic_reduce): returns the reduced CSR
    (row_ptr, col_idx, values) as device tensors."""
    torch = _torch()
    lib = load_library()
    rp = to_dev(row_ptr, torch.int32)
    ci = to_dev(col_idx, torch.int32)
    v = to_dev(values, torch.complex128)
    n = rp.numel() - 1
    m = np.ascontiguousarray(np.asarray(master, np.int32))
    s_ = np.ascontiguousarray(np.asarray(slave, np.int32))
    cr = np.ascontiguousarray(np.asarray(coef_row, np.complex128))
    cc = np.ascontiguousarray(np.asarray(coef_col, np.complex128))
    dh = None if dirichlet is None else np.ascontiguousarray(np.asarray(dirichlet, np.uint8))
    rp_out = torch.empty(n + 1, dtype=torch.int32, device=rp.device)
    nnz = C.c_int64(0)
    args = (m.ctypes.data_as(C.c_void_p), s_.ctypes.data_as(C.c_void_p), cr.ctypes.data_as(C.c_void_p),
            cc.ctypes.data_as(C.c_void_p), None if dh is None else dh.ctypes.data_as(C.c_void_p))
    _check(lib.efm_periodic_reduce(n, dptr(rp), dptr(ci), dptr(v), int(m.size), *args, dptr(rp_out), None, None, 0,
                                   C.byref(nnz), stream_ptr()), "efm_periodic_reduce")
    col_out = torch.empty(max(nnz.value, 1), dtype=torch.int32, device=rp.device)
    val_out = torch.empty(max(nnz.value, 1), dtype=torch.complex128, device=rp.device)
    _check(lib.efm_periodic_reduce(n, dptr(rp), dptr(ci), dptr(v), int(m.size), *args, dptr(rp_out), dptr(col_out),
                                   dptr(val_out), int(nnz.value), C.byref(nnz), stream_ptr()), "efm_periodic_reduce")
    return rp_out, col_out[:nnz.value], val_out[:nnz.value]


class CellBatch:
    """Block-diagonal batch of independent BiCGSTAB systems (efm_cells_*) in global
    (concatenated) numbering: row_ptr, col, val device tensors; the rest host arrays."""

    def __init__(self, row_off, row_ptr, col, val, node_off, edge_nodes, edge_coefs, nrhs, port_off, port_idx,
                 port_w, port_coef):
        torch = _torch()
        lib = load_library()
        self._lib = lib
        self._dev = (to_dev(row_ptr, torch.int32), to_dev(col, torch.int32), to_dev(val, torch.complex128))
        arrs = dict(row_off=np.ascontiguousarray(row_off, np.int64),
                    node_off=np.ascontiguousarray(node_off, np.int64),
                    edge_nodes=np.ascontiguousarray(edge_nodes, np.int32).reshape(-1),
                    edge_coefs=np.ascontiguousarray(edge_coefs, np.complex128).reshape(-1),
                    port_off=np.ascontiguousarray(port_off, np.int64), port_idx=np.ascontiguousarray(port_idx, np.int32),
                    port_w=np.ascontiguousarray(port_w, np.complex128),
                    port_coef=np.ascontiguousarray(port_coef, np.complex128))
        self.ncells = arrs["row_off"].size - 1
        self.n = int(arrs["row_off"][-1])
        self.nrhs = int(nrhs)
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)
        h = C.c_void_p()
        rp, cl, vl = self._dev
        _check(lib.efm_cells_create(self.ncells, ptr(arrs["row_off"]), dptr(rp), dptr(cl), dptr(vl), int(cl.numel()),
                                    ptr(arrs["node_off"]), ptr(arrs["edge_nodes"]),
                                    ptr(arrs["edge_coefs"]), self.nrhs, ptr(arrs["port_off"]), ptr(arrs["port_idx"]),
                                    ptr(arrs["port_w"]), ptr(arrs["port_coef"]), stream_ptr(), C.byref(h)),
               "efm_cells_create")
        self._h = h

    def solve(self, B, tol=1e-10, max_iterations=10000, check_every=32):
        """B: (n, nrhs) complex128 device tensor.  Returns (X, iters, resid)."""
        torch = _torch()
        Bd = to_dev(B, torch.complex128)
        if Bd.numel() != self.n * self.nrhs:
            raise EdgeFemError("right-hand side has the wrong size")
        X = torch.empty_like(Bd)
        iters = np.zeros(self.ncells * self.nrhs, np.int32)
        resid = np.zeros(self.ncells * self.nrhs, np.float64)
        opts = SolveOpts(EFM_METHOD_BICGSTAB, float(tol), int(max_iterations), int(check_every))
        _check(self._lib.efm_cells_solve(self._h, dptr(Bd), dptr(X), C.byref(opts), iters.ctypes.data_as(C.c_void_p),
                                         resid.ctypes.data_as(C.c_void_p), stream_ptr()), "efm_cells_solve")
        return X, iters.reshape(self.ncells, self.nrhs), resid.reshape(self.ncells, self.nrhs)

    def set_spmv_variant(self, variant):
        """1: entry sub-tiles (default), 0: rows over 8 lanes (the only one when a row
        holds more than 1024 entries)."""
        _check(self._lib.efm_cells_set_spmv_variant(self._h, int(variant)), "efm_cells_set_spmv_variant")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.efm_cells_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def direct_solve(row_ptr, col_idx, values, B, nrhs: int, lowrank=None):
    """Dense-LU direct solve on the GPU (efm_direct_solve).  lowrank: (gidx, gval, coefs) as
    set_lowrank takes them, or None.  Returns (X, info): X a device complex128 tensor like B,
    info > 0 for a singular matrix."""
    torch = _torch()
    lib = load_library()
    rp = to_dev(row_ptr, torch.int32)
    ci = to_dev(col_idx, torch.int32)
    v = to_dev(values, torch.complex128)
    Bd = to_dev(B, torch.complex128)
    n = rp.numel() - 1
    if Bd.numel() != n * nrhs:
        raise EdgeFemError("right-hand side block has the wrong size")
    X = torch.empty_like(Bd)
    info = C.c_int()
    if lowrank:
        gi = to_dev(lowrank[0], torch.int32)
        gv = to_dev(lowrank[1], torch.complex128)
        cf = np.ascontiguousarray(np.asarray(lowrank[2], np.complex128))
        P = cf.size
        L = gi.numel() // max(P, 1)
        _check(lib.efm_direct_solve(n, dptr(rp), dptr(ci), dptr(v), P, L, dptr(gi), dptr(gv),
                                    cf.ctypes.data_as(C.c_void_p), dptr(Bd), dptr(X), int(nrhs), C.byref(info),
                                    stream_ptr()), "efm_direct_solve")
    else:
        _check(lib.efm_direct_solve(n, dptr(rp), dptr(ci), dptr(v), 0, 0, None, None, None, dptr(Bd), dptr(X),
                                    int(nrhs), C.byref(info), stream_ptr()), "efm_direct_solve")
    return X, int(info.value)
