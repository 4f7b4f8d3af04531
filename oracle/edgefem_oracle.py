"""numpy/scipy restatement of the EdgeFEM hot path (test infrastructure only).

Functions and the reference code each one follows:

==========================  =========================================================
load_gmsh_v2                src/mesh_gmsh.cpp:15-101
build_edges                 src/mesh_gmsh.cpp:104-146, include/edgefem/mesh.hpp:56-60
gradients_and_volume        src/edge_basis.cpp:14-26
whitney_curl_curl_matrix    src/edge_basis.cpp:47-63
whitney_mass_matrix         src/edge_basis.cpp:65-86 (lambda_int 28-30)
triangle_whitney_mass       src/edge_basis.cpp:88-130
build_edge_pec              src/bc.cpp:47-80
assemble_maxwell            src/assemble_maxwell.cpp:46-350 (with the PML stretch)
pml_stretch                 src/assemble_maxwell.cpp:58-172 (per-tet scalar stretch,
                            diagnostics)
debye_eps, lorentz_eps,     include/edgefem/materials/dispersive.hpp:37-287
drude_eps, drude_lorentz_eps
assemble_maxwell_km         src/sweep.cpp:90-172
port_surface_mass           src/ports/wave_port.cpp:548-585
solve_port_mode_2d          src/ports/wave_port.cpp:410-514
build_wave_port_2d          src/ports/wave_port.cpp:516-546
solve_te10_mode             src/ports/port_eigensolve.cpp:45-63
sparams_eigenmode           src/assemble_maxwell.cpp:637-787
calculate_sparams           src/assemble_maxwell.cpp:352-394 (port loading 209-242)
frequency_sweep             src/sweep.cpp:174-329 (K/M reuse branch)
extract_surface_mesh        src/ports/wave_port.cpp:54-111
solve_port_eigens           src/ports/port_eigensolve.cpp:97-275
build_wave_port             src/ports/wave_port.cpp:113-205
build_periodic_pairs        src/periodic.cpp:48-164
floquet_phase_from_angle    src/periodic.cpp:204-225 (and set_floquet_phase)
assemble_maxwell_periodic   src/assemble_maxwell.cpp:496-574 (sparse restatement;
                            periodic_eliminate_dense is the literal dense one)
calculate_sparams_periodic  src/assemble_maxwell.cpp:576-635
port_abc_coefficient        src/assemble_maxwell.cpp:274-320
compute_barycentric         src/edge_basis.cpp:132-151
evaluate_edge_field         src/edge_basis.cpp:162-192
tet_fields                  the per-tet loop of src/io/vtk_export.cpp:14-40 and
                            src/post/huygens_surface.cpp:95-125 (E at a point, curl E)
build_lumped_port           src/ports/lumped_port.cpp:21-160
normalize_port_weights      src/assemble_maxwell.cpp:396-494
eps_gradient_eigenmode      adjoint of sparams_eigenmode w.r.t. per-tet eps_r (no
                            reference counterpart; pinned by central differences)
touchstone_text             src/io/touchstone.cpp:68-135 (write_touchstone_nport;
                            iostream precision 12 = printf %.12g)
touchstone_text_2port       src/io/touchstone.cpp:50-62 (legacy write_touchstone)
array_package_json_text     src/io/array_export.cpp:33-123
patterns_csv_text           src/io/array_export.cpp:125-145
coupling_csv_text           src/io/array_export.cpp:147-165 (extract_all_couplings,
                            src/array/coupling_extract.cpp:21-57)
scan_csv_text               src/io/array_export.cpp:167-189 (active_vswr,
                            src/array/active_impedance.cpp:117-123)
==========================  =========================================================

Third-party routines replaced: Eigen ``setFromTriplets`` (duplicates summed) ->
``scipy.sparse.coo_matrix(...).tocsr()``; Eigen ``GeneralizedSelfAdjointEigenSolver`` ->
``scipy.linalg.eigh(K, M)`` (same LAPACK problem class; eigenvector sign is fixed by
the physically anchored rule in ``orient_port_mode``, because the reference's sign is
whatever Eigen returns and cannot be reproduced); Eigen ``SparseLU`` ->
``scipy.sparse.linalg.splu`` (complex128 SuperLU).

Parity pins (independent of the formulas restated here), all exercised in tests/:

* topology: brute-force invariants of tests/test_edge_indexing.cpp:11-31 on the shipped
  cube_cavity.msh fixture; exact edge counts of the structured meshes (closed form);
  Euler characteristic V - E + F - T = 1 of a ball;
* element matrices: rational closed forms on the unit tet; an independent 4-point
  quadrature (exact for degree 2) of the mass integrand; K_e . G = 0 for the discrete
  gradient; the 3 triangles of tests/test_triangle_mass_matrix.cpp:88-102 against a
  7-point degree-5 quadrature at 1e-12;
* assembly: a dense brute-force element loop on small meshes; A == K - k0^2 M; A is
  Hermitian for real materials, non-Hermitian when lossy (tests/test_maxwell.cpp);
* port mode: kc^2 against the analytic (pi/a)^2 as h -> 0, scipy eigh residual; the nodal
  solve_port_eigens TE10 (kc, cos(pi x/a), Z0 = omega mu0/beta, sin(pi x/a) edge weights) and
  TM11/TM21 (kc^2 = (m pi/a)^2 + (n pi/b)^2 at the P1 convergence rate, sin sin field zero on
  the wall, Z0 = beta/(omega eps0), the power scaling by an independent gradient integral);
* S-parameters: |S21| ~ 1, |S11| ~ 0, arg S21 ~ -beta L, reciprocity, and the
  reference-measured values on examples/rect_waveguide.msh (docs/validation.md:20-27).
"""

from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional, Sequence

import numpy as np
import scipy.linalg
import scipy.sparse as sp
import scipy.sparse.linalg as spla

C0 = 299792458.0
MU0 = 4.0 * math.pi * 1e-7
EPS0 = 1.0 / (MU0 * C0 * C0)
ETA0_ASM = MU0 * C0            # assemble_maxwell.cpp:43
ETA0_PORT = 376.730313668      # port_eigensolve.cpp:15

TET_PAIRS = np.array([(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)])   # edge_basis.cpp:9-12
TRI_PAIRS = np.array([(0, 1), (1, 2), (2, 0)])                           # mesh_gmsh.cpp:110-112


# ---------------------------------------------------------------------------
# Mesh
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class OMesh:
    node_ids: np.ndarray      # (N,) int64 gmsh IDs in $Nodes order
    xyz: np.ndarray           # (N,3)
    tet_ids: np.ndarray       # (T,4) int64 gmsh IDs
    tet_phys: np.ndarray
    tri_ids: np.ndarray       # (F,3)
    tri_phys: np.ndarray
    edges: np.ndarray = None          # (n,2) int64 IDs (min,max)
    tet_edges: np.ndarray = None      # (T,6)
    tet_sign: np.ndarray = None       # (T,6) +-1
    tri_edges: np.ndarray = None      # (F,3)
    tri_sign: np.ndarray = None
    node_index: Dict[int, int] = None

    @property
    def n_edges(self):
        return self.edges.shape[0]

    def tet_xyz(self):
        idx = np.vectorize(self.node_index.__getitem__)(self.tet_ids) if self.tet_ids.size else self.tet_ids
        return self.xyz[idx]

    def tri_xyz(self):
        idx = np.vectorize(self.node_index.__getitem__)(self.tri_ids) if self.tri_ids.size else self.tri_ids
        return self.xyz[idx]


def load_gmsh_v2(path: str) -> OMesh:
    """gmsh v2.2 ASCII (mesh_gmsh.cpp:15-101), as the reference's stream reads it: whole
    lines are compared with "$Nodes" / "$Elements"; inside a section every number is the
    next whitespace-separated token (blank lines included); $Nodes gives ``id x y z``;
    $Elements records ``id type ntags tags... nodes...`` keep types 2 (Tri3) and 4 (Tet4)
    with the first tag as ``phys`` (0 without tags); type 1 (Line2) reads its two nodes and
    drops them; any other type skips the rest of its line; after each section two lines
    are consumed (the end of the last record and the closing marker)."""
    with open(path, newline="") as f:
        text = f.read()
    pos, end = 0, len(text)

    def token() -> str:
        nonlocal pos
        while pos < end and text[pos].isspace():
            pos += 1
        s = pos
        while pos < end and not text[pos].isspace():
            pos += 1
        return text[s:pos]

    def getline() -> str:
        nonlocal pos
        e = text.find("\n", pos)
        e = end if e < 0 else e
        line, pos = text[pos:e], e + 1
        return line

    node_ids, xyz, tets, tetp, tris, trip = [], [], [], [], [], []
    while pos < end:
        line = getline()
        if line == "$Nodes":
            for _ in range(int(token())):
                node_ids.append(int(token()))
                xyz.append([float(token()), float(token()), float(token())])
            getline(), getline()
        elif line == "$Elements":
            for _ in range(int(token())):
                token()  # element ID
                typ, ntag = int(token()), int(token())
                tags = [int(token()) for _ in range(ntag)]
                phys = tags[0] if tags else 0
                if typ == 2:
                    tris.append([int(token()) for _ in range(3)]); trip.append(phys)
                elif typ == 1:
                    token(), token()
                elif typ == 4:
                    tets.append([int(token()) for _ in range(4)]); tetp.append(phys)
                else:
                    getline()
            getline(), getline()
    mesh = OMesh(node_ids=np.asarray(node_ids, np.int64), xyz=np.asarray(xyz, float).reshape(-1, 3),
                 tet_ids=np.asarray(tets, np.int64).reshape(-1, 4), tet_phys=np.asarray(tetp, np.int32),
                 tri_ids=np.asarray(tris, np.int64).reshape(-1, 3), tri_phys=np.asarray(trip, np.int32))
    mesh.node_index = {int(v): k for k, v in enumerate(mesh.node_ids)}
    build_edges(mesh)
    return mesh


def mesh_from_arrays(node_ids, xyz, tet_ids, tet_phys, tri_ids, tri_phys) -> OMesh:
    mesh = OMesh(node_ids=np.asarray(node_ids, np.int64), xyz=np.asarray(xyz, float),
                 tet_ids=np.asarray(tet_ids, np.int64), tet_phys=np.asarray(tet_phys, np.int32),
                 tri_ids=np.asarray(tri_ids, np.int64), tri_phys=np.asarray(tri_phys, np.int32))
    mesh.node_index = {int(v): k for k, v in enumerate(mesh.node_ids)}
    build_edges(mesh)
    return mesh


def build_edges(mesh: OMesh) -> None:
    """First-seen global edge numbering (mesh_gmsh.cpp:104-146): scan tets in file
    order with local pairs (0,1),(0,2),(0,3),(1,2),(1,3),(2,3), then tris with pairs
    (0,1),(1,2),(2,0).  Edge stored as (min ID, max ID); orientation +1 iff
    conn[a] < conn[b] (gmsh IDs); key = (a<<32) ^ b of the sorted pair."""
    index: Dict[int, int] = {}
    edges: List = []
    T = mesh.tet_ids.shape[0]
    F = mesh.tri_ids.shape[0]
    tet_edges = np.empty((T, 6), np.int32)
    tet_sign = np.empty((T, 6), np.int8)
    tri_edges = np.empty((F, 3), np.int32)
    tri_sign = np.empty((F, 3), np.int8)
    for conn, pairs, out_e, out_s in ((mesh.tet_ids, TET_PAIRS, tet_edges, tet_sign),
                                      (mesh.tri_ids, TRI_PAIRS, tri_edges, tri_sign)):
        cl = conn.tolist()
        for t, c in enumerate(cl):
            for e, (pa, pb) in enumerate(pairs):
                a = c[pa]; b = c[pb]
                lo, hi = (a, b) if a < b else (b, a)
                key = ((lo << 32) ^ hi) & 0xFFFFFFFFFFFFFFFF
                idx = index.get(key)
                if idx is None:
                    idx = len(edges)
                    index[key] = idx
                    edges.append((lo, hi))
                out_e[t, e] = idx
                out_s[t, e] = 1 if a < b else -1
    mesh.edges = np.asarray(edges, np.int64).reshape(-1, 2)
    mesh.tet_edges, mesh.tet_sign, mesh.tri_edges, mesh.tri_sign = tet_edges, tet_sign, tri_edges, tri_sign


# ---------------------------------------------------------------------------
# Element matrices (batched over tets)
# ---------------------------------------------------------------------------
def gradients_and_volume(X: np.ndarray):
    """X (T,4,3).  B = [v0-v3, v1-v3, v2-v3] (columns); grad lambda_i = column i of
    B^{-T} for i < 3, grad lambda_3 = -sum; V = |det B| / 6 (edge_basis.cpp:14-26)."""
    B = np.stack([X[:, 0] - X[:, 3], X[:, 1] - X[:, 3], X[:, 2] - X[:, 3]], axis=2)  # columns
    Binv = np.linalg.inv(B)
    g = np.empty((X.shape[0], 4, 3))
    g[:, 0:3, :] = Binv            # row i of B^{-1} == column i of B^{-T}
    g[:, 3, :] = -(g[:, 0] + g[:, 1] + g[:, 2])
    V = np.abs(np.linalg.det(B)) / 6.0
    return g, V


def whitney_curl_curl_matrix(X: np.ndarray) -> np.ndarray:
    """K_e(i,j) = V c_i . c_j with c_i = 2 grad(l_a) x grad(l_b) (edge_basis.cpp:47-63)."""
    g, V = gradients_and_volume(X)
    c = 2.0 * np.cross(g[:, TET_PAIRS[:, 0]], g[:, TET_PAIRS[:, 1]])   # (T,6,3)
    return V[:, None, None] * np.einsum("tik,tjk->tij", c, c)


def whitney_mass_matrix(X: np.ndarray) -> np.ndarray:
    """M_e(i,j) = gb.gd I(a,c) - gb.gc I(a,d) - ga.gd I(b,c) + ga.gc I(b,d) with
    I(i,j) = V/10 (i==j) or V/20 (edge_basis.cpp:65-86)."""
    g, V = gradients_and_volume(X)
    G = np.einsum("tik,tjk->tij", g, g)     # (T,4,4) grad dot products
    I = np.where(np.eye(4, dtype=bool)[None], V[:, None, None] / 10.0, V[:, None, None] / 20.0)
    a = TET_PAIRS[:, 0][:, None]; b = TET_PAIRS[:, 1][:, None]
    c = TET_PAIRS[:, 0][None, :]; d = TET_PAIRS[:, 1][None, :]
    M = (G[:, b, d] * I[:, a, c] - G[:, b, c] * I[:, a, d]
         - G[:, a, d] * I[:, b, c] + G[:, a, c] * I[:, b, d])
    return M


def compute_barycentric(v: np.ndarray, p: np.ndarray) -> np.ndarray:
    """Barycentric coordinates of p in the tet v (4,3): solve [v0-v3, v1-v3, v2-v3] l = p - v3,
    lambda_3 = 1 - sum (edge_basis.cpp:132-151)."""
    v = np.asarray(v, np.float64)
    T = np.stack([v[0] - v[3], v[1] - v[3], v[2] - v[3]], axis=1)
    l012 = np.linalg.solve(T, np.asarray(p, np.float64) - v[3])
    return np.array([l012[0], l012[1], l012[2], 1.0 - l012.sum()])


def evaluate_edge_field(v: np.ndarray, edge_orient, edge_dofs, p: np.ndarray) -> np.ndarray:
    """E(p) = sum_e s_e x_e (l_a grad l_b - l_b grad l_a) over the local edges (a, b) in
    TET_PAIRS order (edge_basis.cpp:162-192)."""
    v = np.asarray(v, np.float64)
    lam = compute_barycentric(v, p)
    g, _ = gradients_and_volume(v[None])
    g = g[0]
    E = np.zeros(3, np.complex128)
    for e, (a, b) in enumerate(TET_PAIRS):
        W = lam[a] * g[b] - lam[b] * g[a]
        E += complex(edge_dofs[e]) * float(edge_orient[e]) * W
    return E


def tet_fields(mesh: OMesh, x: np.ndarray, lam: Optional[np.ndarray] = None):
    """E at a point of every tet (lam (T,4) barycentric weights; None: the centroid, as the
    VTK export and the Huygens surface evaluate it) and the tet-constant curl E =
    sum_e s_e x_e 2 grad l_a x grad l_b (whitney_edge_curls, edge_basis.cpp:34-45).
    x (n,) or (n, R).  Returns E, curl: (T, R, 3) complex (R = 1 for a vector x)."""
    X = mesh.xyz[np.vectorize(mesh.node_index.__getitem__)(mesh.tet_ids)]      # (T,4,3)
    g, _ = gradients_and_volume(X)
    xs = np.asarray(x, np.complex128)
    if xs.ndim == 1:
        xs = xs[:, None]
    c = xs[mesh.tet_edges] * mesh.tet_sign[:, :, None]                          # (T,6,R)
    if lam is None:
        lam = np.full((X.shape[0], 4), 0.25)
    a, b = TET_PAIRS[:, 0], TET_PAIRS[:, 1]
    W = lam[:, a, None] * g[:, b] - lam[:, b, None] * g[:, a]                    # (T,6,3)
    curls = 2.0 * np.cross(g[:, a], g[:, b])                                    # (T,6,3)
    E = np.einsum("ter,tek->trk", c, W)
    curl = np.einsum("ter,tek->trk", c, curls)
    return E, curl


def triangle_whitney_mass(v: np.ndarray) -> np.ndarray:
    """3x3 surface Whitney mass on triangles v (F,3,3), local edges (0,1),(1,2),(2,0);
    in-plane gradients g_k = n x (opposite edge) / 2A; I = A/6 or A/12
    (edge_basis.cpp:88-130).  Degenerate triangles (2A < 1e-30) give zeros."""
    v = np.asarray(v, float).reshape(-1, 3, 3)
    normal = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    area2 = np.linalg.norm(normal, axis=1)
    ok = area2 >= 1e-30
    safe = np.where(ok, area2, 1.0)
    nh = normal / safe[:, None]
    g = np.stack([np.cross(nh, v[:, 2] - v[:, 1]), np.cross(nh, v[:, 0] - v[:, 2]),
                  np.cross(nh, v[:, 1] - v[:, 0])], axis=1) / safe[:, None, None]
    area = safe / 2.0
    G = np.einsum("tik,tjk->tij", g, g)
    I = np.where(np.eye(3, dtype=bool)[None], area[:, None, None] / 6.0, area[:, None, None] / 12.0)
    a = TRI_PAIRS[:, 0][:, None]; b = TRI_PAIRS[:, 1][:, None]
    c = TRI_PAIRS[:, 0][None, :]; d = TRI_PAIRS[:, 1][None, :]
    M = (G[:, b, d] * I[:, a, c] - G[:, b, c] * I[:, a, d]
         - G[:, a, d] * I[:, b, c] + G[:, a, c] * I[:, b, d])
    M[~ok] = 0.0
    return M


def triangle_surface_curls(v: np.ndarray):
    """Per-triangle (curls[3], area) with curl_e = 2 (g_a x g_b) . n (wave_port.cpp:463-476)."""
    v = np.asarray(v, float).reshape(-1, 3, 3)
    normal = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    area2 = np.linalg.norm(normal, axis=1)
    ok = area2 >= 1e-30
    safe = np.where(ok, area2, 1.0)
    nh = normal / safe[:, None]
    g = np.stack([np.cross(nh, v[:, 2] - v[:, 1]), np.cross(nh, v[:, 0] - v[:, 2]),
                  np.cross(nh, v[:, 1] - v[:, 0])], axis=1) / safe[:, None, None]
    curls = 2.0 * np.einsum("tek,tk->te", np.cross(g[:, TRI_PAIRS[:, 0]], g[:, TRI_PAIRS[:, 1]]), nh)
    return curls, safe / 2.0, ok, g


# ---------------------------------------------------------------------------
# Boundary conditions and materials
# ---------------------------------------------------------------------------
def build_edge_pec(mesh: OMesh, pec_tag: int) -> np.ndarray:
    """Dirichlet edges = all edges of tris with the PEC tag (bc.cpp:47-80); returns a
    boolean mask over global edges."""
    mask = np.zeros(mesh.n_edges, bool)
    sel = mesh.tri_phys == pec_tag
    mask[mesh.tri_edges[sel].ravel()] = True
    return mask


@dataclasses.dataclass
class Params:
    """Subset of MaxwellParams (maxwell.hpp:46-143) the path uses.  Material precedence:
    per-tet array (an extension beyond the reference) > region map > global value."""
    omega: float = 0.0
    eps_r: complex = 1.0
    mu_r: complex = 1.0
    eps_r_regions: Dict[int, complex] = dataclasses.field(default_factory=dict)
    mu_r_regions: Dict[int, complex] = dataclasses.field(default_factory=dict)
    eps_tet: Optional[np.ndarray] = None
    mu_tet: Optional[np.ndarray] = None
    use_abc: bool = False
    abc_surface_tags: set = dataclasses.field(default_factory=set)
    port_abc_scale: float = 1.0
    use_port_abc: bool = False
    port_abc_type: str = "beta"     # none | beta | beta_norm | impedance_match | modal_admittance
    # dispersive models: tag -> callable omega -> complex (take precedence over the regions)
    eps_models: Dict[int, object] = dataclasses.field(default_factory=dict)
    mu_models: Dict[int, object] = dataclasses.field(default_factory=dict)
    pml_sigma: float = 0.0
    pml_regions: set = dataclasses.field(default_factory=set)
    # tag -> (sigma_max[3], thickness[3], grading_order)
    pml_tensor: Dict[int, tuple] = dataclasses.field(default_factory=dict)
    enforce_pml_heuristics: bool = True


def debye_eps(omega, eps_s, eps_inf, tau):
    """dispersive.hpp:50-54: eps_inf + (eps_s - eps_inf) / (1 + j omega tau)."""
    return eps_inf + (eps_s - eps_inf) / complex(1.0, omega * tau)


def lorentz_eps(omega, eps_inf, poles):
    """dispersive.hpp:110-121: eps_inf + sum d w0^2 / (w0^2 - w^2 + j g w); poles (d, w0, g)."""
    e = complex(eps_inf)
    for d, w0, g in poles:
        e += d * w0 * w0 / complex(w0 * w0 - omega * omega, g * omega)
    return e


def drude_eps(omega, omega_p, gamma):
    """dispersive.hpp:166-178: 1 - wp^2 / (w^2 + j g w), -1e30 at w = 0."""
    if omega == 0.0:
        return complex(-1e30, 0.0)
    return 1.0 - omega_p * omega_p / complex(omega * omega, gamma * omega)


def drude_lorentz_eps(omega, eps_inf, omega_p, gamma_d, poles):
    """dispersive.hpp:240-262: Drude term on eps_inf (the whole value -1e30 at w = 0), then
    the Lorentz poles added."""
    if omega != 0.0:
        e = eps_inf - omega_p * omega_p / complex(omega * omega, gamma_d * omega)
    else:
        e = complex(-1e30, 0.0)
    return lorentz_eps(omega, 0.0, poles) + e


def tet_materials(mesh: OMesh, p: Params):
    """Per-tet eps, mu at p.omega: per-tet array > dispersive model > region > global
    (maxwell.hpp:92-143)."""
    T = mesh.tet_ids.shape[0]

    def one(tag, regions, models, default):
        m = models.get(int(tag))
        return complex(m(p.omega)) if m is not None else complex(regions.get(int(tag), default))
    if p.eps_tet is not None:
        eps = np.asarray(p.eps_tet, np.complex128)
    else:
        eps = np.array([one(t, p.eps_r_regions, p.eps_models, p.eps_r) for t in mesh.tet_phys], np.complex128) if T else np.zeros(0, complex)
    if p.mu_tet is not None:
        mu = np.asarray(p.mu_tet, np.complex128)
    else:
        mu = np.array([one(t, p.mu_r_regions, p.mu_models, p.mu_r) for t in mesh.tet_phys], np.complex128) if T else np.zeros(0, complex)
    return eps, mu


def pml_stretch(mesh: OMesh, p: Params):
    """Per-tet scalar PML stretch and per-region reflection estimates, tet by tet as
    assemble_maxwell.cpp:58-172 computes them.  Returns (s[T], {tag: reflection[3]})."""
    T = mesh.tet_ids.shape[0]
    X = mesh.tet_xyz()
    refl = {}
    for tag, (smax, thick, order) in p.pml_tensor.items():
        r = []
        for a in range(3):
            if smax[a] <= 0.0 or thick[a] <= 0.0 or abs(p.omega) <= 0.0:
                r.append(1.0)
            else:
                r.append(math.exp(-2.0 * (smax[a] * thick[a] / (order + 1.0)) / abs(p.omega)))
        refl[int(tag)] = r
    bounds = {}
    for t in range(T):
        tag = int(mesh.tet_phys[t])
        if tag not in p.pml_tensor:
            continue
        lo, hi = bounds.get(tag, (None, None))
        for i in range(4):
            x = X[t, i]
            lo = x.copy() if lo is None else np.minimum(lo, x)
            hi = x.copy() if hi is None else np.maximum(hi, x)
        bounds[tag] = (lo, hi)
    s = np.ones(T, np.complex128)
    if p.omega == 0.0:
        return s, refl
    for t in range(T):
        tag = int(mesh.tet_phys[t])
        if tag in p.pml_tensor:
            smax, thick, order = p.pml_tensor[tag]
            c = (X[t, 0] + X[t, 1] + X[t, 2] + X[t, 3]) / 4.0
            lo, hi = bounds[tag]
            st = [1.0 + 0j, 1.0 + 0j, 1.0 + 0j]
            for a in range(3):
                if smax[a] <= 0.0 or thick[a] <= 0.0:
                    continue
                md = min(c[a] - lo[a], hi[a] - c[a])
                xi = 1.0 - md / thick[a] if md < thick[a] else 0.0
                if p.enforce_pml_heuristics and xi > 0.0:
                    xi = max(xi, 1e-3)
                sig = smax[a] * xi ** order
                if p.enforce_pml_heuristics and smax[a] > 0.0 and sig < smax[a] * 1e-3:
                    sig = smax[a] * 1e-3
                st[a] = complex(1.0, sig / p.omega)
            s[t] = (st[0] + st[1] + st[2]) / 3.0
        elif tag in p.pml_regions:
            s[t] = complex(1.0, p.pml_sigma / p.omega)
    return s, refl


# ---------------------------------------------------------------------------
# Assembly
# ---------------------------------------------------------------------------
def _triplets(mesh: OMesh):
    T = mesh.tet_ids.shape[0]
    rows = np.repeat(mesh.tet_edges, 6, axis=1).reshape(T, 6, 6)
    cols = np.tile(mesh.tet_edges, (1, 6)).reshape(T, 6, 6)
    sign = (mesh.tet_sign[:, :, None].astype(float) * mesh.tet_sign[:, None, :])
    return rows, cols, sign


def _dirichlet_zero(A: sp.csr_matrix, mask: np.ndarray, diag_value) -> sp.csr_matrix:
    """Zero rows and columns of Dirichlet edges in place (explicit zeros kept), then set
    the diagonal (assemble_maxwell.cpp:322-347)."""
    A = A.tocsr(copy=True)
    A.sort_indices()
    rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
    kill = mask[rows] | mask[A.indices]
    A.data[kill] = 0
    if diag_value is not None:
        d = np.nonzero(mask)[0]
        pos = _find(A, d, d)
        A.data[pos] = diag_value
    return A


def _find(A: sp.csr_matrix, r: np.ndarray, c: np.ndarray) -> np.ndarray:
    """Positions of (r, c) in a sorted CSR matrix; raises if absent."""
    out = np.empty(len(r), np.int64)
    for k, (ri, ci) in enumerate(zip(r, c)):
        lo, hi = A.indptr[ri], A.indptr[ri + 1]
        j = lo + np.searchsorted(A.indices[lo:hi], ci)
        if j >= hi or A.indices[j] != ci:
            raise KeyError((ri, ci))
        out[k] = j
    return out


def volume_pattern(mesh: OMesh) -> sp.csr_matrix:
    rows, cols, _ = _triplets(mesh)
    n = mesh.n_edges
    P = sp.coo_matrix((np.ones(rows.size), (rows.ravel(), cols.ravel())), shape=(n, n)).tocsr()
    P.sort_indices()
    return P


def assemble_maxwell_km(mesh: OMesh, p: Params, dirichlet: np.ndarray):
    """K = sum (K_e / mu_r) s_i s_j, M = sum (eps_r M_e) s_i s_j, duplicates summed in
    triplet order; Dirichlet rows/cols zeroed, K diagonal 1, M diagonal 0
    (sweep.cpp:90-172).  Returns sorted CSR matrices with the full volume pattern."""
    X = mesh.tet_xyz()
    Ke = whitney_curl_curl_matrix(X)
    Me = whitney_mass_matrix(X)
    # static materials: the reference's K/M assembly reads the region maps only (sweep.cpp:110-118)
    eps, mu = tet_materials(mesh, dataclasses.replace(p, eps_models={}, mu_models={}))
    rows, cols, sign = _triplets(mesh)
    n = mesh.n_edges
    kv = (Ke.astype(np.complex128) / mu[:, None, None]) * sign
    mv = (eps[:, None, None] * Me) * sign
    K = sp.coo_matrix((kv.ravel(), (rows.ravel(), cols.ravel())), shape=(n, n)).tocsr()
    M = sp.coo_matrix((mv.ravel(), (rows.ravel(), cols.ravel())), shape=(n, n)).tocsr()
    K.sort_indices(); M.sort_indices()
    K = _dirichlet_zero(K, dirichlet, 1.0)
    M = _dirichlet_zero(M, dirichlet, None)
    return K, M


def assemble_maxwell(mesh: OMesh, p: Params, dirichlet: np.ndarray, ports=(), active=-1):
    """A = sum (K_e/mu - k0^2 eps M_e) s_i s_j (+ rank-1 port loading w w^H / Z0, + j k0
    ABC on boundary edges), Dirichlet rows/cols zeroed with unit diagonal, b = 2w/sqrt(Z0)
    on the active port (assemble_maxwell.cpp:46-350, PML stretch s = 1)."""
    X = mesh.tet_xyz()
    Ke = whitney_curl_curl_matrix(X)
    Me = whitney_mass_matrix(X)
    eps, mu = tet_materials(mesh, p)
    s, _ = pml_stretch(mesh, p)
    k0 = p.omega / C0
    rows, cols, sign = _triplets(mesh)
    n = mesh.n_edges
    Ks = Ke / s[:, None, None]       # Kloc = K_e / s, Mloc = M_e s (assemble_maxwell.cpp:168-171)
    Ms = Me * s[:, None, None]
    val = (Ks / mu[:, None, None] - (k0 * k0) * eps[:, None, None] * Ms) * sign
    R = [rows.ravel()]; C = [cols.ravel()]; V = [val.ravel()]
    b = np.zeros(n, np.complex128)
    for i, port in enumerate(ports):
        if port.Z0 == 0 or len(port.weights) != len(port.edges):
            continue
        e = np.asarray(port.edges)
        w = np.asarray(port.weights, np.complex128)
        keep = ~dirichlet[e]
        e = e[keep]; w = w[keep]
        R.append(np.repeat(e, len(e))); C.append(np.tile(e, len(e)))
        V.append((w[:, None] * np.conj(w)[None, :] / port.Z0).ravel())
        if i == active:
            np.add.at(b, e, 2.0 / np.sqrt(complex(port.Z0)) * w)
    A = sp.coo_matrix((np.concatenate(V), (np.concatenate(R), np.concatenate(C))), shape=(n, n)).tocsr()
    A.sort_indices()
    if p.use_abc:
        sel = np.ones(mesh.tri_ids.shape[0], bool) if not p.abc_surface_tags else np.isin(mesh.tri_phys, list(p.abc_surface_tags))
        be = np.unique(mesh.tri_edges[sel].ravel())
        be = be[~dirichlet[be]]
        A = A + sp.csr_matrix((np.full(be.size, 1j * k0), (be, be)), shape=(n, n))
        A.sort_indices()
    if p.use_port_abc and p.port_abc_type != "none":
        for port in ports:
            coeff = port_abc_coefficient(p, port)
            if coeff is None:
                continue
            e = np.asarray(port.edges)
            e = e[~dirichlet[e]]
            A = A + sp.csr_matrix((np.full(e.size, coeff), (e, e)), shape=(n, n))
            A.sort_indices()
    A = _dirichlet_zero(A, dirichlet, 1.0)
    b[dirichlet] = 0
    return A, b


def port_abc_coefficient(p: Params, port) -> Optional[complex]:
    """Diagonal port ABC term of assemble_maxwell.cpp:274-320: only when k0^2 > kc^2;
    j beta, j beta/k0, j beta sqrt(Re Z0/eta0) or j omega eps0/Re Z0, with eta0 = mu0 c0."""
    if port.Z0 == 0:
        return None
    k0 = p.omega / C0
    bsq = k0 * k0 - port.kc * port.kc
    if bsq <= 0:
        return None
    beta = math.sqrt(bsq)
    z0r = complex(port.Z0).real
    t = p.port_abc_type
    if t == "beta":
        return 1j * beta
    if t == "beta_norm":
        return 1j * beta / k0
    if t == "impedance_match":
        return 1j * beta * math.sqrt(z0r / ETA0_ASM)
    if t == "modal_admittance":
        return 1j * p.omega * EPS0 / z0r
    raise ValueError(f"unknown port ABC type {t!r}")


def port_surface_mass(mesh: OMesh, surface_tag: int, dirichlet: np.ndarray) -> sp.csr_matrix:
    """M_s = sum over port-tag tris of s_i s_j M_tri(i,j), Dirichlet rows/cols skipped
    (wave_port.cpp:548-585)."""
    sel = np.nonzero(mesh.tri_phys == surface_tag)[0]
    n = mesh.n_edges
    if sel.size == 0:
        return sp.csr_matrix((n, n))
    idx = np.vectorize(mesh.node_index.__getitem__)(mesh.tri_ids[sel])
    Mt = triangle_whitney_mass(mesh.xyz[idx])
    e = mesh.tri_edges[sel]; s = mesh.tri_sign[sel].astype(float)
    val = Mt * s[:, :, None] * s[:, None, :]
    r = np.repeat(e, 3, axis=1).reshape(-1, 3, 3)
    c = np.tile(e, (1, 3)).reshape(-1, 3, 3)
    keep = ~(dirichlet[r] | dirichlet[c])
    Ms = sp.coo_matrix((val[keep], (r[keep], c[keep])), shape=(n, n)).tocsr()
    Ms.sort_indices()
    return Ms


# ---------------------------------------------------------------------------
# Ports
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class OPort:
    surface_tag: int
    kc: float
    Z0: complex
    beta: complex
    edges: np.ndarray
    weights: np.ndarray
    kc_sq_discrete: float = 0.0


def solve_te10_mode(a: float, freq: float):
    """kc = pi/a, beta = sqrt(max(0, k^2 - kc^2)), Z0 = eta0 k / beta with the literal
    eta0 = 376.730313668 (port_eigensolve.cpp:45-63)."""
    k = 2 * math.pi * freq / C0
    kc = math.pi / a
    beta = math.sqrt(max(0.0, k * k - kc * kc))
    Z0 = ETA0_PORT * k / beta if beta > 0 else 0.0
    return kc, beta, Z0


def orient_port_mode(mesh: OMesh, surface_tag: int, vfull: np.ndarray) -> np.ndarray:
    """Fix the eigenvector sign: p = sum_e w_e int_S N_e . y dS > 0, with the closed form
    int_tri N_e dS = (A/3)(grad l_b - grad l_a) (lumped_port.cpp:87-90); falls back to x
    when |p| is negligible.  The reference leaves the sign to Eigen; this rule is the
    build's documented convention."""
    sel = np.nonzero(mesh.tri_phys == surface_tag)[0]
    idx = np.vectorize(mesh.node_index.__getitem__)(mesh.tri_ids[sel])
    _, area, ok, g = triangle_surface_curls(mesh.xyz[idx])
    integ = (area / 3.0)[:, None, None] * (g[:, TRI_PAIRS[:, 1]] - g[:, TRI_PAIRS[:, 0]])  # (F,3,3)
    w = vfull[mesh.tri_edges[sel]] * mesh.tri_sign[sel]
    pvec = np.einsum("fe,fek->k", w, integ)
    scale = np.abs(vfull).max() * np.sqrt(area.sum()) + 1e-300
    comp = pvec[1] if abs(pvec[1]) > 1e-9 * scale else pvec[0]
    return -vfull if comp < 0 else vfull


def solve_port_mode_2d(mesh: OMesh, surface_tag: int, dirichlet: np.ndarray, target_kc_sq: float):
    """Dense generalised eigenproblem K_s e = kc^2 M_s e on the free port edges, with
    K_s(i,j) = s_i s_j c_i c_j A and M_s from triangle_whitney_mass; picks the eigenvalue
    closest to the target among those >= 1e-6 max(target, 1) (wave_port.cpp:410-514).
    Free edges are indexed in first-seen order over the port tris."""
    sel = np.nonzero(mesh.tri_phys == surface_tag)[0]
    local = {}
    order = []
    for t in sel:
        for le in range(3):
            ge = int(mesh.tri_edges[t, le])
            if dirichlet[ge] or ge in local:
                continue
            local[ge] = len(order)
            order.append(ge)
    n = len(order)
    m = mesh.n_edges
    if n == 0:
        return np.zeros(m), 0.0
    idx = np.vectorize(mesh.node_index.__getitem__)(mesh.tri_ids[sel])
    v = mesh.xyz[idx]
    curls, area, ok, _ = triangle_surface_curls(v)
    Mt = triangle_whitney_mass(v)
    K = np.zeros((n, n)); M = np.zeros((n, n))
    for k, t in enumerate(sel):
        if not ok[k]:
            continue
        for i in range(3):
            li = local.get(int(mesh.tri_edges[t, i]))
            if li is None:
                continue
            for j in range(3):
                lj = local.get(int(mesh.tri_edges[t, j]))
                if lj is None:
                    continue
                s = float(mesh.tri_sign[t, i]) * float(mesh.tri_sign[t, j])
                K[li, lj] += s * curls[k, i] * curls[k, j] * area[k]
                M[li, lj] += s * Mt[k, i, j]
    evals, evecs = scipy.linalg.eigh(K, M)
    kc_min = 1e-6 * max(target_kc_sq, 1.0)
    best = -1
    best_d = np.inf
    for i, ev in enumerate(evals):
        if ev < kc_min:
            continue
        d = abs(ev - target_kc_sq)
        if d < best_d:
            best_d = d; best = i
    if best < 0:
        return np.zeros(m), 0.0
    vfull = np.zeros(m)
    vfull[np.asarray(order)] = evecs[:, best]
    return orient_port_mode(mesh, surface_tag, vfull), float(evals[best])


def build_wave_port_2d(mesh: OMesh, surface_tag: int, kc: float, Z0: complex, beta: complex,
                       dirichlet: np.ndarray, target_kc_sq: float) -> OPort:
    """Port edges = sorted set of all port-tri edges (PEC included, weight 0); weights =
    the 2D mode on them; mode.kc <- sqrt(discrete kc^2) (wave_port.cpp:516-546)."""
    v, kc2 = solve_port_mode_2d(mesh, surface_tag, dirichlet, target_kc_sq)
    if kc2 > 0:
        kc = math.sqrt(kc2)
    edges = np.unique(mesh.tri_edges[mesh.tri_phys == surface_tag].ravel())
    return OPort(surface_tag=surface_tag, kc=kc, Z0=Z0, beta=beta, edges=edges,
                 weights=v[edges].astype(np.complex128), kc_sq_discrete=kc2)


def port_eps_mu(mesh: OMesh, p: Params, surface_tag: int) -> complex:
    """eps*mu of the first tet (file order) owning a face on the port surface, faces
    scanned {0,1,2},{0,1,3},{0,2,3},{1,2,3} (assemble_maxwell.cpp:652-686)."""
    faces = {tuple(sorted(t)) for t in mesh.tri_ids[mesh.tri_phys == surface_tag].tolist()}
    eps, mu = tet_materials(mesh, p)
    for t, c in enumerate(mesh.tet_ids.tolist()):
        for f in ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)):
            if tuple(sorted((c[f[0]], c[f[1]], c[f[2]]))) in faces:
                return complex(eps[t] * mu[t])
    return 1.0 + 0j


def eigenmode_setup(mesh: OMesh, p: Params, dirichlet: np.ndarray, ports: Sequence[OPort]):
    """betas, M_s per port and M_s-normalised port vectors (assemble_maxwell.cpp:644-728).
    Returns None for betas when a port is below cutoff."""
    k0 = p.omega / C0
    betas = []
    for port in ports:
        em = port_eps_mu(mesh, p, port.surface_tag)
        b2 = em * k0 * k0 - port.kc * port.kc
        if b2.real <= 0:
            return None, None, None
        betas.append(np.sqrt(complex(b2)))
    Ms = [port_surface_mass(mesh, port.surface_tag, dirichlet) for port in ports]
    vecs = []
    n = mesh.n_edges
    for port, M in zip(ports, Ms):
        e = np.zeros(n, np.complex128)
        keep = ~dirichlet[port.edges]
        e[port.edges[keep]] = port.weights[keep]
        nsq = np.real(np.vdot(e, M @ e))
        if nsq > 1e-30:
            e = e / np.sqrt(nsq)
        vecs.append(e)
    return np.array(betas), Ms, vecs


def sparams_eigenmode(mesh: OMesh, p: Params, dirichlet: np.ndarray, ports: Sequence[OPort],
                      solver="splu", return_systems=False):
    """A = assemble_maxwell(no ports) + sum_i j beta_i scale M_s,i; b_a = 2 j beta_a scale
    M_s,a e_a; S(j,a) = e_j^H M_s,j x - delta_ja (assemble_maxwell.cpp:637-787).  The two
    excitations share A, so one LU serves both.  Below cutoff the result is NaN (the
    reference returns an uninitialised matrix there)."""
    P = len(ports)
    betas, Ms, vecs = eigenmode_setup(mesh, p, dirichlet, ports)
    if betas is None:
        return np.full((P, P), np.nan + 1j * np.nan)
    A, _ = assemble_maxwell(mesh, p, dirichlet)
    for i in range(P):
        A = A + (1j * p.port_abc_scale * betas[i]) * Ms[i]
    A = A.tocsc()
    B = np.stack([2.0 * (1j * p.port_abc_scale * betas[a]) * (Ms[a] @ vecs[a]) for a in range(P)], axis=1)
    if solver == "splu":
        lu = spla.splu(A, permc_spec="MMD_AT_PLUS_A")
        X = lu.solve(B)
    else:
        X = solver(A, B)
    S = np.empty((P, P), np.complex128)
    for a in range(P):
        for j in range(P):
            S[j, a] = np.vdot(vecs[j], Ms[j] @ X[:, a]) - (1.0 if j == a else 0.0)
    if return_systems:
        return S, dict(A=A, B=B, X=X, betas=betas, Ms=Ms, vecs=vecs)
    return S


def aux_preconditioner(mesh: OMesh, dirichlet: np.ndarray, A: sp.csr_matrix):
    """P = D^-1 + G diag(G^T A G)^-1 G^T with G the signed edge-node incidence (tail -1,
    head +1 by gmsh ID, Dirichlet rows removed): point Jacobi plus one nodal Jacobi step
    on the gradient space, the null space of the curl-curl part.  Returned as a function
    of an (n, k) block."""
    n = A.shape[0]
    order = np.argsort(mesh.node_ids)
    pos = lambda ids: order[np.searchsorted(mesh.node_ids, ids, sorter=order)]
    ends = np.stack([pos(mesh.edges[:, 0]), pos(mesh.edges[:, 1])], 1)
    G = sp.csr_matrix((np.tile([-1.0, 1.0], n), (np.repeat(np.arange(n), 2), ends.ravel())),
                      shape=(n, mesh.xyz.shape[0]))
    G = (sp.diags((~dirichlet).astype(float)) @ G).tocsr()
    GT = G.T.tocsr()
    dn = (GT @ A @ G).diagonal()
    dninv = np.where(dn != 0, 1.0 / np.where(dn != 0, dn, 1.0), 0.0)
    dinv = 1.0 / A.diagonal()
    return lambda r: dinv[:, None] * r + G @ (dninv[:, None] * (GT @ r))


def cocg_solve(A: sp.csr_matrix, B: np.ndarray, P, tol: float, maxit: int, refine: int = 0,
               log=None):
    """FP64 stand-in for the LU solve where a sparse factor of A does not fit in memory:
    preconditioned conjugate orthogonal CG (the bilinear form u^T v, valid because A is
    complex symmetric) on every column of B, to ||b - A x|| <= tol ||b|| with the true
    residual checked at the end, then `refine` steps of x += COCG(A, b - A x).  Returns X and
    the final true relative residual of every column."""
    def cocg(Rhs, tol_):
        X = np.zeros_like(Rhs)
        R = Rhs.copy()
        Z = P(R)
        Pd = Z.copy()
        rho = np.einsum("ij,ij->j", R, Z)
        nb = np.linalg.norm(Rhs, axis=0)
        for it in range(maxit):
            Q = A @ Pd
            al = rho / np.einsum("ij,ij->j", Pd, Q)
            X += al * Pd
            R -= al * Q
            rn = np.linalg.norm(R, axis=0) / nb
            if log is not None and it % 500 == 0:
                log(f"  cocg it {it}: {rn}")
            if np.all(rn <= tol_):
                break
            Z = P(R)
            rho_n = np.einsum("ij,ij->j", R, Z)
            Pd = Z + (rho_n / rho) * Pd
            rho = rho_n
        return X, it + 1

    X, its = cocg(B, tol)
    iters = [its]
    for _ in range(refine):
        D, its = cocg(B - A @ X, 1e-6)
        X = X + D
        iters.append(its)
    res = np.linalg.norm(B - A @ X, axis=0) / np.linalg.norm(B, axis=0)
    return X, res, iters


def calculate_sparams(mesh: OMesh, p: Params, dirichlet: np.ndarray, ports: Sequence[OPort]):
    """Rank-1 port loading path with FP64 LU in place of the Krylov solve
    (assemble_maxwell.cpp:352-394): V_j = sum conj(w_k) x_k over free edges,
    S_ii = (V_i - sqrt Z0_i)/sqrt Z0_i, S_ji = V_j / sqrt Z0_i."""
    P = len(ports)
    S = np.empty((P, P), np.complex128)
    for i in range(P):
        A, b = assemble_maxwell(mesh, p, dirichlet, ports, i)
        x = spla.splu(A.tocsc(), permc_spec="MMD_AT_PLUS_A").solve(b)
        vinc = np.sqrt(complex(ports[i].Z0))
        for j in range(P):
            keep = ~dirichlet[ports[j].edges]
            V = np.sum(np.conj(ports[j].weights[keep]) * x[ports[j].edges[keep]])
            S[j, i] = (V - vinc) / vinc if i == j else V / vinc
    return S


def frequency_sweep(mesh: OMesh, p: Params, dirichlet: np.ndarray, ports: Sequence[OPort], freqs):
    """K/M branch of frequency_sweep (sweep.cpp:174-329): K and M assembled once; per
    frequency A = K - k0^2 M + Y_port (+ j k0 on the ABC edges), Y_port = sum w w^H / Z0;
    b = 2 w / sqrt(Z0) per active port; S as in calculate_sparams; FP64 LU per frequency
    (port ABC diagonals are not part of this branch, as in the reference)."""
    K, M = assemble_maxwell_km(mesh, p, dirichlet)
    n = mesh.n_edges
    R, C, V = [], [], []
    for port in ports:
        if port.Z0 == 0 or len(port.weights) != len(port.edges):
            continue
        e = np.asarray(port.edges)
        w = np.asarray(port.weights, np.complex128)
        keep = ~dirichlet[e]
        e, w = e[keep], w[keep]
        R.append(np.repeat(e, e.size)); C.append(np.tile(e, e.size))
        V.append((w[:, None] * np.conj(w)[None, :] / port.Z0).ravel())
    Y = sp.coo_matrix((np.concatenate(V) if V else np.zeros(0), (np.concatenate(R) if R else np.zeros(0, int),
                       np.concatenate(C) if C else np.zeros(0, int))), shape=(n, n)).tocsr()
    abc = np.zeros(0, int)
    if p.use_abc:
        sel = np.ones(mesh.tri_ids.shape[0], bool) if not p.abc_surface_tags else np.isin(mesh.tri_phys, list(p.abc_surface_tags))
        abc = np.unique(mesh.tri_edges[sel].ravel())
        abc = abc[~dirichlet[abc]]
    out = []
    P = len(ports)
    for f in np.asarray(freqs, float):
        k0 = 2 * math.pi * f / C0
        A = (K - (k0 * k0) * M + Y).tocsc()
        if abc.size:
            A = (A + sp.csc_matrix((np.full(abc.size, 1j * k0), (abc, abc)), shape=(n, n))).tocsc()
        lu = spla.splu(A, permc_spec="MMD_AT_PLUS_A")
        S = np.empty((P, P), np.complex128)
        for i, port in enumerate(ports):
            b = np.zeros(n, np.complex128)
            e = np.asarray(port.edges)
            keep = ~dirichlet[e]
            b[e[keep]] += 2.0 / np.sqrt(complex(port.Z0)) * np.asarray(port.weights)[keep]
            x = lu.solve(b)
            vinc = np.sqrt(complex(port.Z0))
            for j in range(P):
                kj = ~dirichlet[ports[j].edges]
                Vj = np.sum(np.conj(np.asarray(ports[j].weights)[kj]) * x[np.asarray(ports[j].edges)[kj]])
                S[j, i] = (Vj - vinc) / vinc if i == j else Vj / vinc
        out.append(S)
    return np.array(out)


def extract_surface_mesh(mesh: OMesh, tag: int):
    """(node ids in first-use order, xy, tri node ids, boundary edges used once, volume tri
    indices) of the tris of ``tag`` (wave_port.cpp:54-111)."""
    sel = np.nonzero(mesh.tri_phys == tag)[0]
    tri = mesh.tri_ids[sel]
    ids = []
    seen = set()
    for t in tri:
        for v in t:
            if int(v) not in seen:
                seen.add(int(v))
                ids.append(int(v))
    ids = np.array(ids, np.int64)
    xy = mesh.xyz[[mesh.node_index[int(v)] for v in ids]][:, :2]
    cnt = {}
    for t in tri:
        for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0])):
            k = (min(int(a), int(b)), max(int(a), int(b)))
            cnt[k] = cnt.get(k, 0) + 1
    lines = sorted(k for k, c in cnt.items() if c == 1)
    return ids, xy, tri, lines, sel


def solve_port_eigens(surf, num_modes: int, omega: float, eps_r=1.0, mu_r=1.0, pol: str = "TE"):
    """Nodal P1 port modes (port_eigensolve.cpp:97-275), element loop restated with
    explicit 2x2 algebra; returns a list of dicts (kc, fc, beta, Z0, field).  Eigenvector
    sign: made real with its largest entry positive, entries within 1e-8 of the largest
    magnitude counting as tied and the first in node order deciding (the reference's sign
    is Eigen's).  A
    degenerate pair is resolved to the member with E along y (the least y-variation).  The
    zero-mode cut is kc^2 < max(1e-12, 1e-10 kc^2_max): with the reference's absolute
    1e-12 the Neumann constant mode, whose computed kc^2 is roundoff of order 1e-9 on
    millimetre meshes, would come back as a spurious first TE mode."""
    ids, xy, tri, lines, _ = surf
    N = ids.size
    pos = {int(v): k for k, v in enumerate(ids)}
    pec = np.zeros(N, bool)
    if pol == "TM":
        for a, b in lines:
            pec[pos[a]] = pec[pos[b]] = True
    dof = -np.ones(N, int)
    dof[~pec] = np.arange((~pec).sum())
    nd = int((~pec).sum())
    A = np.zeros((nd, nd)); Bm = np.zeros((nd, nd)); Ky = np.zeros((nd, nd))
    tris = []
    for t in tri:
        ii = [pos[int(v)] for v in t]
        p = xy[ii]
        d = (p[1, 0] - p[0, 0]) * (p[2, 1] - p[0, 1]) - (p[2, 0] - p[0, 0]) * (p[1, 1] - p[0, 1])
        area = abs(d) / 2.0
        # grad l_i = perpendicular of the opposite edge / (2 * signed area)
        g = np.array([[p[1, 1] - p[2, 1], p[2, 0] - p[1, 0]],
                      [p[2, 1] - p[0, 1], p[0, 0] - p[2, 0]],
                      [p[0, 1] - p[1, 1], p[1, 0] - p[0, 0]]]) / d
        tris.append((ii, area, g))
        for a in range(3):
            for b in range(3):
                if dof[ii[a]] >= 0 and dof[ii[b]] >= 0:
                    A[dof[ii[a]], dof[ii[b]]] += area * g[a] @ g[b]
                    Bm[dof[ii[a]], dof[ii[b]]] += area / 12.0 * (2.0 if a == b else 1.0)
                    Ky[dof[ii[a]], dof[ii[b]]] += area * g[a][1] * g[b][1]
    w, V = scipy.linalg.eigh(A, Bm)
    # a degenerate cluster (square port: TE10/TE01) is rotated so that its first vector
    # minimises int |df/dy|^2 over the cluster's span (E along y); the reference returns
    # whatever member of the plane Eigen gives, which nothing can reproduce
    i = 0
    while i < w.size:
        j = i + 1
        while j < w.size and abs(w[j] - w[i]) <= 1e-4 * max(abs(w[i]), 1e-300):
            j += 1
        if j - i > 1:
            Vc = V[:, i:j]
            V[:, i:j] = Vc @ np.linalg.eigh(Vc.T @ Ky @ Vc)[1]
        i = j
    eps = EPS0 * complex(eps_r); mu = MU0 * complex(mu_r)
    k = omega * np.sqrt(mu * eps)
    out = []
    for i in range(w.size):
        kc2 = w[i]
        if not np.isfinite(kc2) or kc2 < max(1e-12, 1e-10 * abs(w[-1])):
            continue
        kc = math.sqrt(kc2)
        beta = np.sqrt(complex(k * k - kc2))
        Z0 = (omega * mu / beta if pol == "TE" else beta / (omega * eps)) if beta != 0 else 0j
        v = V[:, i].astype(complex)
        av = np.abs(v)
        j = int(np.nonzero(av >= (1.0 - 1e-8) * av.max())[0][0])   # ties within 1e-8: node order
        v = v if v[j].real >= 0 else -v
        f = np.zeros(N, complex); f[~pec] = v
        pw = 0.0
        for ii, area, g in tris:
            gr = g.T @ f[ii]
            if pol == "TE":
                fe = 1j * omega * mu / kc2; fh = beta / kc2
                Et = (-fe * gr[1], fe * gr[0]); Ht = (fh * gr[0], fh * gr[1])
            else:
                fe = -beta / kc2; fh = 1.0 / (1j * omega * mu)
                Et = (fe * gr[0], fe * gr[1]); Ht = (-fh * gr[1], fh * gr[0])
            pw += 0.5 * area * (Et[0] * np.conj(Ht[1]) - Et[1] * np.conj(Ht[0]))
        pr = pw.real if pw.real > 0 else abs(pw)
        if pr <= 0:
            continue
        out.append(dict(kc=kc, fc=kc * C0 / (2 * math.pi), beta=beta, Z0=Z0, field=f / math.sqrt(pr),
                        omega=omega, mu=mu, eps=eps, pol=pol))
        if len(out) >= num_modes:
            break
    out.sort(key=lambda md: md["fc"])
    return out[:num_modes]


def build_wave_port(mesh: OMesh, surf, mode) -> OPort:
    """Edge weights from a nodal mode (wave_port.cpp:113-205): tangential E per tri dotted
    with the edge vector, averaged over the tris of each edge, times the sign of the
    summed z normal; edges sorted."""
    ids, xy, tri, _, sel = surf
    pos = {int(v): k for k, v in enumerate(ids)}
    acc, cnt = {}, {}
    nz = 0.0
    kc2 = mode["kc"] ** 2
    for t3, f in zip(tri, sel):
        ii = [pos[int(v)] for v in t3]
        p = xy[ii]
        d = (p[1, 0] - p[0, 0]) * (p[2, 1] - p[0, 1]) - (p[2, 0] - p[0, 0]) * (p[1, 1] - p[0, 1])
        g = np.array([[p[1, 1] - p[2, 1], p[2, 0] - p[1, 0]],
                      [p[2, 1] - p[0, 1], p[0, 0] - p[2, 0]],
                      [p[0, 1] - p[1, 1], p[1, 0] - p[0, 0]]]) / d
        gr = g.T @ mode["field"][ii]
        if mode["pol"] == "TE":
            fe = 1j * mode["omega"] * mode["mu"] / kc2
            Et = np.array([fe * gr[1], -fe * gr[0], 0.0])
        else:
            fe = -mode["beta"] / kc2
            Et = np.array([fe * gr[0], fe * gr[1], 0.0])
        v = mesh.xyz[[mesh.node_index[int(x)] for x in mesh.tri_ids[f]]]
        nz += np.cross(v[1] - v[0], v[2] - v[0])[2]
        for le in range(3):
            e = int(mesh.tri_edges[f, le])
            pa = mesh.xyz[mesh.node_index[int(mesh.edges[e, 0])]]
            pb = mesh.xyz[mesh.node_index[int(mesh.edges[e, 1])]]
            acc[e] = acc.get(e, 0) + Et @ (pb - pa)
            cnt[e] = cnt.get(e, 0) + 1
    sign = 1.0 if nz >= 0 else -1.0
    edges = np.array(sorted(acc), np.int64)
    w = np.array([sign * acc[e] / cnt[e] for e in edges], np.complex128)
    return OPort(surface_tag=int(mesh.tri_phys[sel[0]]), kc=mode["kc"], Z0=mode["Z0"], beta=mode["beta"],
                 edges=edges, weights=w)


@dataclasses.dataclass
class OPeriodic:
    master: np.ndarray           # edge indices
    slave: np.ndarray
    orient: np.ndarray           # master_orient * slave_orient
    period: np.ndarray
    phase: complex = 1.0 + 0j


def _edge_centroids(mesh: OMesh, edges: np.ndarray) -> np.ndarray:
    i0 = np.array([mesh.node_index[int(v)] for v in mesh.edges[edges, 0]], np.int64)
    i1 = np.array([mesh.node_index[int(v)] for v in mesh.edges[edges, 1]], np.int64)
    return 0.5 * (mesh.xyz[i0] + mesh.xyz[i1])


def build_periodic_pairs(mesh: OMesh, master_tag: int, slave_tag: int, period, tol: float = 1e-9) -> OPeriodic:
    """Pairs of edges on two faces whose centroids differ by the period (periodic.cpp:48-164):
    lookup by rounded centroid, brute force when that misses, error on an unmatched master
    or a slave matched twice.  Orientation of an edge on a face = its local sign in the
    LAST tri of that face (mesh order) containing it, +1 if none (periodic.cpp:135-146).
    Pairs are listed in ascending master edge (the reference lists them in hash-set
    order; the elimination does not depend on the order)."""
    period = np.asarray(period, float)
    me = np.unique(mesh.tri_edges[mesh.tri_phys == master_tag].ravel())
    se = np.unique(mesh.tri_edges[mesh.tri_phys == slave_tag].ravel())
    if me.size == 0 or se.size == 0:
        raise ValueError("no edges on the master or slave surface")
    cm = _edge_centroids(mesh, me) + period
    cs = _edge_centroids(mesh, se)
    key = lambda c: tuple(np.round(c / tol).astype(np.int64))
    table = {key(c): k for k, c in enumerate(cs)}
    match = np.empty(me.size, np.int64)
    used = set()
    for i, c in enumerate(cm):
        k = table.get(key(c), -1)
        if k < 0 or np.linalg.norm(cs[k] - c) >= tol:
            d = np.linalg.norm(cs - c, axis=1)
            hit = np.nonzero(d < tol)[0]
            k = int(hit[0]) if hit.size else -1
        if k < 0:
            raise ValueError(f"no slave edge for master edge {me[i]}")
        if k in used:
            raise ValueError(f"slave edge {se[k]} matched twice")
        used.add(k)
        match[i] = k
    slave = se[match]

    def last_orient(tag, edges):
        o = np.ones(edges.size, np.int64)
        pos = {int(e): i for i, e in enumerate(edges)}
        for f in np.nonzero(mesh.tri_phys == tag)[0]:
            for le in range(3):
                i = pos.get(int(mesh.tri_edges[f, le]), -1)
                if i >= 0:
                    o[i] = mesh.tri_sign[f, le]
        return o
    orient = last_orient(master_tag, me) * last_orient(slave_tag, slave)
    return OPeriodic(master=me, slave=slave, orient=orient, period=period)


def floquet_phase_from_angle(period, theta: float, phi: float, k0: float) -> complex:
    """exp(j k_inc . L), k_inc = k0 (sin t cos p, sin t sin p, -cos t) (periodic.cpp:213-225)."""
    k = k0 * np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), -math.cos(theta)])
    return complex(np.exp(1j * float(k @ np.asarray(period, float))))


def set_floquet_phase(pbc: OPeriodic, k_t) -> None:
    """phase = exp(j (kx Lx + ky Ly)) (periodic.cpp:204-211)."""
    pbc.phase = complex(np.exp(1j * (k_t[0] * pbc.period[0] + k_t[1] * pbc.period[1])))


def _periodic_coefs(pbc: OPeriodic, mode: str):
    """(row, column, recovery) coefficient per pair.  "reference": rows += phi o R_s, cols
    += conj(phi) o C_s, x_s = phi o x_m (assemble_maxwell.cpp:536-546, 606-613).
    "consistent": rows += conj(phi) o, cols += phi o, x_s = phi o x_m, so the column
    operation matches the recovery ansatz."""
    o = pbc.orient.astype(float)
    ph = complex(pbc.phase)
    if mode == "reference":
        return ph * o, np.conj(ph) * o, ph * o
    if mode == "consistent":
        return np.conj(ph) * o, ph * o, ph * o
    raise ValueError(f"unknown periodic mode {mode!r}")


def periodic_eliminate(A: sp.csr_matrix, b: np.ndarray, pbc: OPeriodic, dirichlet: np.ndarray,
                       mode: str = "reference"):
    """Sparse restatement of the elimination (assemble_maxwell.cpp:496-574):
    A' = Z (E A F) Z + I_S, b' = Z E b, E = I + sum a_k e_m e_s^T, F = I + sum c_k e_s e_m^T
    over pairs with neither edge PEC, Z zeroing the non-PEC slaves S.  Exact zeros stay
    in the pattern (the reference's sparseView drops them)."""
    n = A.shape[0]
    a, c, _ = _periodic_coefs(pbc, mode)
    m, s = pbc.master, pbc.slave
    act = ~(dirichlet[m] | dirichlet[s])
    E = sp.identity(n, complex, format="csr") + sp.csr_matrix((a[act], (m[act], s[act])), shape=(n, n))
    F = sp.identity(n, complex, format="csr") + sp.csr_matrix((c[act], (s[act], m[act])), shape=(n, n))
    S = s[~dirichlet[s]]
    keep = np.ones(n)
    keep[S] = 0.0
    Z = sp.diags(keep)
    Ap = (Z @ (E @ A @ F) @ Z).tocsr() + sp.csr_matrix((np.ones(S.size), (S, S)), shape=(n, n))
    Ap.sort_indices()
    bp = keep * (E @ b)
    return Ap.tocsr(), bp


def periodic_eliminate_dense(A: sp.csr_matrix, b: np.ndarray, pbc: OPeriodic, dirichlet: np.ndarray,
                             mode: str = "reference"):
    """Literal dense restatement: sequential row/column updates per pair, then zero slave
    rows/columns with unit diagonal (assemble_maxwell.cpp:526-571).  Small cells only."""
    D = A.toarray().astype(np.complex128)
    bb = np.array(b, np.complex128)
    a, c, _ = _periodic_coefs(pbc, mode)
    for k in range(pbc.master.size):
        mm, ss = int(pbc.master[k]), int(pbc.slave[k])
        if dirichlet[mm] or dirichlet[ss]:
            continue
        D[mm, :] += a[k] * D[ss, :]
        D[:, mm] += c[k] * D[:, ss]
        bb[mm] += a[k] * bb[ss]
    for ss in pbc.slave:
        if dirichlet[ss]:
            continue
        D[ss, :] = 0
        D[:, ss] = 0
        D[ss, ss] = 1
        bb[ss] = 0
    return D, bb


def calculate_sparams_periodic(mesh: OMesh, p: Params, dirichlet: np.ndarray, pbc: OPeriodic,
                               ports: Sequence[OPort], mode: str = "reference"):
    """assemble_maxwell_periodic + solve + slave recovery x_s = r_k x_m + the rank-1 S
    extraction (assemble_maxwell.cpp:576-635), FP64 LU."""
    P = len(ports)
    S = np.empty((P, P), np.complex128)
    _, _, rec = _periodic_coefs(pbc, mode)
    for i in range(P):
        A, b = assemble_maxwell(mesh, p, dirichlet, ports, i)
        Ap, bp = periodic_eliminate(A, b, pbc, dirichlet, mode)
        x = spla.splu(Ap.tocsc(), permc_spec="MMD_AT_PLUS_A").solve(bp)
        x[pbc.slave] = rec * x[pbc.master]
        vinc = np.sqrt(complex(ports[i].Z0))
        for j in range(P):
            kj = ~dirichlet[ports[j].edges]
            Vj = np.sum(np.conj(np.asarray(ports[j].weights)[kj]) * x[np.asarray(ports[j].edges)[kj]])
            S[j, i] = (Vj - vinc) / vinc if i == j else Vj / vinc
    return S


def build_lumped_port(mesh: OMesh, surface_tag: int, z0: float = 50.0, e_dir=(0.0, 0.0, 1.0),
                      mode: str = "surface_integral") -> OPort:
    """Lumped port on the tris of ``surface_tag`` (lumped_port.cpp:21-160).  Weights: the
    surface integral w_e = sum_tri s (A/3)(grad l_j - grad l_i).e (in-plane gradients,
    local edges (0,1),(1,2),(2,0)), or the projection of the edge vector (min ID -> max
    ID) on e; then scaled to ||w|| = sqrt(z0).  The reference keeps the edges in hash-set
    order; here they are sorted (S does not depend on the order)."""
    sel = np.nonzero(mesh.tri_phys == surface_tag)[0]
    if sel.size == 0:
        raise ValueError(f"no triangles with surface tag {surface_tag}")
    edges = np.unique(mesh.tri_edges[sel].ravel())
    pos = {int(e): k for k, e in enumerate(edges)}
    ed = np.asarray(e_dir, float)
    ed = ed / np.linalg.norm(ed)
    w = np.zeros(edges.size, np.complex128)
    if mode == "surface_integral":
        for f in sel:
            v = mesh.xyz[[mesh.node_index[int(i)] for i in mesh.tri_ids[f]]]
            nrm = np.cross(v[1] - v[0], v[2] - v[0])
            a2 = np.linalg.norm(nrm)
            if a2 < 1e-30:
                continue
            nh = nrm / a2
            g = [np.cross(nh, v[2] - v[1]) / a2, np.cross(nh, v[0] - v[2]) / a2, np.cross(nh, v[1] - v[0]) / a2]
            for le, (li, lj) in enumerate(TRI_PAIRS):
                ge = int(mesh.tri_edges[f, le])
                w[pos[ge]] += mesh.tri_sign[f, le] * (a2 / 2.0 / 3.0) * np.dot(g[lj] - g[li], ed)
    elif mode == "projection":
        for k, e in enumerate(edges):
            p0 = mesh.xyz[mesh.node_index[int(mesh.edges[e, 0])]]
            p1 = mesh.xyz[mesh.node_index[int(mesh.edges[e, 1])]]
            w[k] = np.dot(p1 - p0, ed)
    else:
        raise ValueError(f"unknown lumped port weight mode {mode!r}")
    nw = np.linalg.norm(w)
    if nw > 1e-15:
        w = w * (math.sqrt(z0) / nw)
    return OPort(surface_tag=surface_tag, kc=0.0, Z0=complex(z0), beta=0.0, edges=edges, weights=w)


def normalize_port_weights(mesh: OMesh, p: Params, dirichlet: np.ndarray, ports: Sequence[OPort]):
    """Scale each port's weights so that |w^H A^{-1} w| = Re Z0 with A assembled without
    ports (assemble_maxwell.cpp:396-494); FP64 LU.  Returns new ports."""
    A, _ = assemble_maxwell(mesh, p, dirichlet)
    lu = spla.splu(A.tocsc(), permc_spec="MMD_AT_PLUS_A")
    out = []
    for port in ports:
        q = dataclasses.replace(port, weights=np.array(port.weights, np.complex128))
        if port.Z0 == 0 or len(port.weights) != len(port.edges):
            out.append(q)
            continue
        e = np.asarray(port.edges)
        keep = ~dirichlet[e]
        rhs = np.zeros(A.shape[0], np.complex128)
        rhs[e[keep]] = q.weights[keep]
        y = lu.solve(rhs)
        c = abs(np.sum(np.conj(q.weights[keep]) * y[e[keep]]))
        target = complex(port.Z0).real
        if np.linalg.norm(q.weights) >= 1e-15 and c > 1e-15 and target > 1e-15:
            q.weights = q.weights * math.sqrt(target / c)
        out.append(q)
    return out


def eps_gradient_eigenmode(mesh: OMesh, p: Params, dirichlet: np.ndarray, ports: Sequence[OPort],
                           G: np.ndarray, design: Optional[np.ndarray] = None):
    """Adjoint gradient of a real loss L(S) of the eigenmode S matrix at one frequency with
    respect to the real part of the per-tet permittivity, given G = dL/d conj(S) (PyTorch's
    convention: L = |S21|^2 gives G[1,0] = 2 S21).  With A x_a = b_a, S_ja = g_j^H x_a -
    delta_ja and dA/d eps_e = -k0^2 M_e (the element mass scattered with edge signs):
        dS_ja/d eps_e = k0^2 lam_j^T M_e x_a,   lam_j = A^{-T} conj(g_j),
        dL/d eps_e = sum_ja Re(conj(G_ja) dS_ja/d eps_e).
    beta_i is held fixed, so tets with a face on a port must not be design tets.
    Returns (S, grad[design])."""
    P = len(ports)
    betas, Ms, vecs = eigenmode_setup(mesh, p, dirichlet, ports)
    if betas is None:
        raise ValueError("frequency below a port cutoff")
    A, _ = assemble_maxwell(mesh, p, dirichlet)
    for i in range(P):
        A = A + (1j * p.port_abc_scale * betas[i]) * Ms[i]
    A = A.tocsc()
    lu = spla.splu(A, permc_spec="MMD_AT_PLUS_A")
    B = np.stack([2.0 * (1j * p.port_abc_scale * betas[a]) * (Ms[a] @ vecs[a]) for a in range(P)], axis=1)
    X = lu.solve(B)
    g = np.stack([Ms[j] @ vecs[j] for j in range(P)], axis=1)           # g_j^H x = vdot(e_j, M_s x)
    S = np.conj(g).T @ X - np.eye(P)
    Lam = lu.solve(np.conj(g), trans="T")
    k0 = p.omega / C0
    T = mesh.tet_ids.shape[0]
    design = np.arange(T) if design is None else np.asarray(design)
    Me = whitney_mass_matrix(mesh.tet_xyz()[design])                   # (D,6,6)
    sg = mesh.tet_sign[design].astype(float)
    te = mesh.tet_edges[design]
    xe = X[te] * sg[:, :, None]                                         # (D,6,P)
    le = Lam[te] * sg[:, :, None]
    dS = k0 * k0 * np.einsum("dkj,dkl,dla->dja", le, Me, xe)            # (D,P,P)
    grad = np.einsum("ja,dja->d", np.conj(G), dS).real
    return S, grad


def touchstone_text(freq, S, fmt: int = 0, z0: float = 50.0) -> str:
    """Text of write_touchstone_nport (touchstone.cpp:68-135).  fmt 0 RI, 1 MA, 2 DB.
    Every number goes through iostream at precision 12 with no float-field flag, i.e.
    printf "%.12g".  MA: |s| and arg(s)*180/pi; DB: 20 log10(max(|s|, 1e-20)) and the same
    angle.  More than 2 ports: a newline before every 4th value after the first, counted
    over the whole row-major matrix; each value is written with its leading space."""
    g = lambda v: "%.12g" % v
    S = np.asarray(S, np.complex128)
    P = S.shape[1]
    out = ["! Touchstone file generated by EdgeFEM\n", "! Number of ports: %d\n" % P,
           "# Hz S %s R %s\n" % (("RI", "MA", "DB")[fmt], g(z0))]
    for fi, f in enumerate(np.asarray(freq, np.float64)):
        line = g(f)
        for q, v in enumerate(S[fi].reshape(-1)):
            if P > 2 and q > 0 and q % 4 == 0:
                line += "\n"
            if fmt == 0:
                a, b = v.real, v.imag
            else:
                mag = abs(v)
                a = mag if fmt == 1 else 20.0 * math.log10(max(mag, 1e-20))
                b = math.atan2(v.imag, v.real) * 180.0 / math.pi
            line += " " + g(a) + " " + g(b)
        out.append(line + "\n")
    return "".join(out)


def touchstone_text_2port(freq, S4) -> str:
    """Text of the legacy write_touchstone (touchstone.cpp:50-62): option line
    "# Hz S RI R 50", then freq and s11, s21, s12, s22 as real/imaginary pairs at %.12g."""
    out = ["# Hz S RI R 50\n"]
    for f, row in zip(np.asarray(freq, np.float64), np.asarray(S4, np.complex128).reshape(-1, 4)):
        out.append(" ".join(["%.12g" % f] + ["%.12g %.12g" % (v.real, v.imag) for v in row]) + "\n")
    return "".join(out)


# ---------------------------------------------------------------------------
# Array-package text (iostream at precision 12 = %.12g; degrees = rad * 180 / pi)
# ---------------------------------------------------------------------------
def _g(v) -> str:
    return "%.12g" % float(v)


def _deg(r) -> float:
    return float(r) * 180.0 / math.pi


def array_package_json_text(model_name, z0, positions, freqs, S_list, n_pattern_freqs, scan) -> str:
    """export_array_package_json (array_export.cpp:33-123); scan = [(theta, phi, loss_db)]."""
    L = ["{", '  "format": "EdgeFEM_ArrayPackage",', '  "version": "1.0",', f'  "model_name": "{model_name}",',
         f'  "reference_impedance": {_g(z0)},', f'  "num_elements": {len(positions)},', '  "element_positions": [']
    for i, p in enumerate(positions):
        L.append(f"    [{_g(p[0])}, {_g(p[1])}, {_g(p[2])}]" + ("," if i < len(positions) - 1 else ""))
    L.append("  ],")
    L.append(f'  "num_frequencies": {len(freqs)},')
    L.append('  "frequencies": [' + ", ".join(_g(f) for f in freqs) + "],")
    L.append('  "S_matrices": [')
    for fi, S in enumerate(S_list):
        S = np.asarray(S)
        L.append("    {")
        L.append(f'      "frequency": {_g(freqs[fi])},')
        L.append('      "data": [')
        for i in range(S.shape[0]):
            row = ", ".join(f"[{_g(S[i, j].real)}, {_g(S[i, j].imag)}]" for j in range(S.shape[1]))
            L.append(f"        [{row}]" + ("," if i < S.shape[0] - 1 else ""))
        L.append("      ]")
        L.append("    }" + ("," if fi < len(S_list) - 1 else ""))
    L.append("  ],")
    L.append(f'  "has_embedded_patterns": {"false" if n_pattern_freqs == 0 else "true"},')
    L.append(f'  "num_pattern_frequencies": {n_pattern_freqs},')
    L.append(f'  "has_scan_data": {"false" if not scan else "true"},')
    if scan:
        L.append(f'  "num_scan_points": {len(scan)},')
        L.append('  "scan_angles": [')
        for i, (t, p, l) in enumerate(scan):
            L.append(f'    {{"theta": {_g(_deg(t))}, "phi": {_g(_deg(p))}, "scan_loss_db": {_g(l)}}}'
                     + ("," if i < len(scan) - 1 else ""))
        L.append("  ]")
    else:
        L.append('  "num_scan_points": 0')
    L.append("}")
    return "\n".join(L) + "\n"


def patterns_csv_text(patterns) -> str:
    """export_patterns_csv (array_export.cpp:125-145); patterns = [(port, [(theta, m0, p0, m90, p90)])]."""
    out = ["port_index,theta_deg,phi0_mag,phi0_phase_deg,phi90_mag,phi90_phase_deg"]
    for port, rows in patterns:
        for t, m0, p0, m90, p90 in rows:
            out.append(f"{port},{_g(_deg(t))},{_g(m0)},{_g(_deg(p0))},{_g(m90)},{_g(_deg(p90))}")
    return "\n".join(out) + "\n"


def coupling_csv_text(S, positions=None) -> str:
    """export_coupling_csv with extract_all_couplings: |S| = hypot, dB = 20 log10 max(|S|,
    1e-20), phase = atan2 in degrees, distance sqrt((dx^2 + dy^2) + dz^2) or 0."""
    S = np.asarray(S)
    out = ["port_i,port_j,S_real,S_imag,magnitude_db,phase_deg,distance_m"]
    for i in range(S.shape[0]):
        for j in range(S.shape[0]):
            if i == j:
                continue
            re, im = float(S[i, j].real), float(S[i, j].imag)
            d = 0.0
            if positions is not None:
                dx, dy, dz = (float(positions[i][k]) - float(positions[j][k]) for k in range(3))
                d = math.sqrt((dx * dx + dy * dy) + dz * dz)
            db = 20.0 * math.log10(max(math.hypot(re, im), 1e-20))
            out.append(f"{i},{j},{_g(re)},{_g(im)},{_g(db)},{_g(_deg(math.atan2(im, re)))},{_g(d)}")
    return "\n".join(out) + "\n"


def scan_csv_text(scan) -> str:
    """export_scan_csv (array_export.cpp:167-189); scan = [(theta, phi, gammas, Zs)]."""
    out = ["theta_deg,phi_deg,element_idx,Gamma_real,Gamma_imag,Z_real,Z_imag,VSWR"]
    for t, p, gs, zs in scan:
        for i, (gm, z) in enumerate(zip(gs, zs)):
            rho = math.hypot(gm.real, gm.imag)
            vswr = 1e12 if rho >= 1.0 - 1e-12 else (1.0 + rho) / (1.0 - rho)
            out.append(f"{_g(_deg(t))},{_g(_deg(p))},{i},{_g(gm.real)},{_g(gm.imag)},{_g(z.real)},{_g(z.imag)},{_g(vswr)}")
    return "\n".join(out) + "\n"
