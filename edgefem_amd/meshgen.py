"""Structured Kuhn-split tetrahedral meshes of a rectangular box, written as gmsh v2.2 ASCII.

The reference meshes every design through the external ``gmsh`` binary
(python/edgefem/designs/waveguide.py:214-227), which this machine does not have, so
the synthetic inputs of the waveguide workloads come from this generator instead.

Geometry and numbering:

* box ``[0,a] x [0,b] x [0,L]`` cut into ``nx*ny*nz`` hexes; every hex is split into the
  six Kuhn tetrahedra that share its (0,0,0)-(1,1,1) diagonal, vertex order chosen so
  that every tet has positive signed volume (mesh_quality.cpp:10-18 convention);
* gmsh node ID = ``1 + i + (nx+1)*(j + (ny+1)*k)``;
* only interior nodes (0<i<nx, 0<j<ny, 0<k<nz) are jittered, by ``U(-f, f) * h_axis``
  per axis from a counter-based splitmix64 stream, so the walls and the port planes
  stay planar;
* boundary quads are split along the same face diagonal the Kuhn tets use, so the
  boundary triangles are faces of the tets;
* physical tags follow waveguide.py:49-52: PEC walls 1, port z=0 -> 2, port z=L -> 3,
  volume 100.

Element lines are ``id type 2 phys elementary nodes...``: triangles first, then tets.
Tets are written hex by hex (k slowest, then j, then i), six per hex.
"""

from __future__ import annotations

import dataclasses
import io
from typing import Optional

import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, counter: np.ndarray) -> np.ndarray:
    """Counter-based splitmix64: one 64-bit output per counter value."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + (counter.astype(np.uint64) + np.uint64(1)) * _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, counter: np.ndarray) -> np.ndarray:
    """Uniform doubles in [0, 1) from the top 53 bits of splitmix64."""
    return (splitmix64(seed, counter) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


# Kuhn simplices of the unit cube: for an axis permutation (p0, p1, p2) the tet is
# 0 -> e_p0 -> e_p0 + e_p1 -> (1,1,1).  Odd permutations get their last two vertices
# swapped so the signed volume is positive.
_PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def _kuhn_local():
    out = []
    for p in _PERMS:
        e = [np.zeros(3, int) for _ in range(4)]
        e[1][p[0]] = 1
        e[2][p[0]] = 1
        e[2][p[1]] = 1
        e[3][:] = 1
        det = np.linalg.det(np.array([e[1] - e[0], e[2] - e[0], e[3] - e[0]], float))
        if det < 0:
            e[2], e[3] = e[3], e[2]
        out.append([tuple(v) for v in e])
    return out


KUHN_TETS = _kuhn_local()


@dataclasses.dataclass
class BoxMesh:
    """Arrays of a generated mesh (0-based node indices; IDs are index + 1)."""

    xyz: np.ndarray          # (N, 3) float64
    node_ids: np.ndarray     # (N,) int64
    tets: np.ndarray         # (T, 4) int32 node indices
    tet_phys: np.ndarray     # (T,) int32
    tris: np.ndarray         # (F, 3) int32 node indices
    tri_phys: np.ndarray     # (F,) int32
    dims: tuple              # (a, b, L)
    grid: tuple              # (nx, ny, nz)
    zplanes: Optional[np.ndarray] = None

    @property
    def tet_ids(self) -> np.ndarray:
        return self.node_ids[self.tets]

    @property
    def tri_ids(self) -> np.ndarray:
        return self.node_ids[self.tris]

    def centroids(self) -> np.ndarray:
        return self.xyz[self.tets].mean(axis=1)

    def write_msh(self, path: str) -> None:
        write_msh(path, self)


def box_mesh(a: float, b: float, L: float, nx: int, ny: int, nz: int, *, seed: int = 0,
             jitter: float = 0.15, zplanes: Optional[np.ndarray] = None,
             tags=(1, 2, 3, 100), side_tags=None, zflat=(), periodic_faces: bool = False) -> BoxMesh:
    """Generate the Kuhn-split box mesh.

    ``zplanes`` optionally gives the nz+1 z coordinates of the grid planes (for layered
    cells); otherwise they are uniform.  ``tags`` = (pec, port_z0, port_zL, volume).
    ``side_tags`` optionally overrides the wall tags per side as a dict with keys
    ``x0, x1, y0, y1`` (used for periodic unit cells).  Nodes on the z planes listed in
    ``zflat`` are not jittered along z (a flat interior interface).  With
    ``periodic_faces`` the x = a and y = b triangles are translated copies of the x = 0 and
    y = 0 ones (same local node order), as a periodic mesher writes a slave surface;
    otherwise every boundary triangle is wound with its normal pointing outward.
    """
    pec, p1, p2, vol = tags
    side = {"x0": pec, "x1": pec, "y0": pec, "y1": pec}
    if side_tags:
        side.update(side_tags)
    xs = np.linspace(0.0, a, nx + 1)
    ys = np.linspace(0.0, b, ny + 1)
    zs = np.linspace(0.0, L, nz + 1) if zplanes is None else np.asarray(zplanes, float)
    assert zs.shape == (nz + 1,)
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    i = i.ravel(); j = j.ravel(); k = k.ravel()
    xyz = np.stack([xs[i], ys[j], zs[k]], axis=1)
    N = xyz.shape[0]
    node_idx = np.arange(N, dtype=np.int64)
    interior = (i > 0) & (i < nx) & (j > 0) & (j < ny) & (k > 0) & (k < nz)
    if jitter > 0 and interior.any():
        hx = a / nx
        hy = b / ny
        # local z spacing: the smaller neighbouring gap keeps planes ordered
        dz_lo = zs[k] - zs[np.maximum(k - 1, 0)]
        dz_hi = zs[np.minimum(k + 1, nz)] - zs[k]
        hz = np.minimum(dz_lo, dz_hi)
        ctr = node_idx * 3
        flat = np.isin(k, np.asarray(list(zflat), dtype=np.int64))
        for ax, h in ((0, hx), (1, hy), (2, hz)):
            u = uniform01(seed, ctr + ax) * 2.0 - 1.0
            d = u * jitter * (h if np.ndim(h) == 0 else h)
            mov = interior & ~flat if ax == 2 else interior
            xyz[:, ax] = np.where(mov, xyz[:, ax] + d, xyz[:, ax])

    def nid(ii, jj, kk):
        return ii + (nx + 1) * (jj + (ny + 1) * kk)

    # tets, hex-major (k slowest), six per hex
    hk, hj, hi = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    hi = hi.ravel(); hj = hj.ravel(); hk = hk.ravel()
    H = hi.size
    tets = np.empty((H, 6, 4), dtype=np.int64)
    for t, verts in enumerate(KUHN_TETS):
        for v, (dx, dy, dz) in enumerate(verts):
            tets[:, t, v] = nid(hi + dx, hj + dy, hk + dz)
    tets = tets.reshape(-1, 4).astype(np.int32)
    tet_phys = np.full(tets.shape[0], vol, dtype=np.int32)

    # boundary triangles: quad (u0,v0)-(u1,v0)-(u1,v1)-(u0,v1) split on the (u0,v0)-(u1,v1) diagonal
    tris = []
    phys = []

    def add_face(c00, c10, c11, c01, tag, flip):
        # two triangles per quad; 'flip' reverses winding so the normal points outward
        t1 = np.stack([c00, c10, c11], axis=-1)
        t2 = np.stack([c00, c11, c01], axis=-1)
        if flip:
            t1 = t1[:, [0, 2, 1]]
            t2 = t2[:, [0, 2, 1]]
        tt = np.stack([t1, t2], axis=1).reshape(-1, 3)
        tris.append(tt)
        phys.append(np.full(tt.shape[0], tag, dtype=np.int32))

    # z = 0 and z = L (ports): u = x, v = y
    fj, fi = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    fi = fi.ravel(); fj = fj.ravel()
    for kk, tag, flip in ((0, p1, True), (nz, p2, False)):
        add_face(nid(fi, fj, kk), nid(fi + 1, fj, kk), nid(fi + 1, fj + 1, kk), nid(fi, fj + 1, kk), tag, flip)
    # y = 0 and y = b: u = x, v = z
    fk, fi = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    fi = fi.ravel(); fk = fk.ravel()
    for jj, tag, flip in ((0, side["y0"], False), (ny, side["y1"], not periodic_faces)):
        add_face(nid(fi, jj, fk), nid(fi + 1, jj, fk), nid(fi + 1, jj, fk + 1), nid(fi, jj, fk + 1), tag, flip)
    # x = 0 and x = a: u = y, v = z
    fk, fj = np.meshgrid(np.arange(nz), np.arange(ny), indexing="ij")
    fj = fj.ravel(); fk = fk.ravel()
    for ii, tag, flip in ((0, side["x0"], True), (nx, side["x1"], periodic_faces)):
        add_face(nid(ii, fj, fk), nid(ii, fj + 1, fk), nid(ii, fj + 1, fk + 1), nid(ii, fj, fk + 1), tag, flip)
    tris = np.concatenate(tris).astype(np.int32)
    tri_phys = np.concatenate(phys)
    return BoxMesh(xyz=xyz, node_ids=node_idx + 1, tets=tets, tet_phys=tet_phys, tris=tris,
                   tri_phys=tri_phys, dims=(a, b, L), grid=(nx, ny, nz), zplanes=zs)


def write_msh(path: str, m: BoxMesh) -> None:
    """Write gmsh v2.2 ASCII (triangles, then tets), the layout load_gmsh_v2 reads."""
    F = m.tris.shape[0]
    T = m.tets.shape[0]
    with open(path, "w") as f:
        f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n")
        f.write("$Nodes\n%d\n" % m.xyz.shape[0])
        buf = io.StringIO()
        np.savetxt(buf, np.column_stack([m.node_ids.astype(np.float64), m.xyz]),
                   fmt=["%d", "%.17g", "%.17g", "%.17g"])
        f.write(buf.getvalue())
        f.write("$EndNodes\n$Elements\n%d\n" % (F + T))
        ids = m.node_ids
        tri_rows = np.column_stack([np.arange(1, F + 1), np.full(F, 2), np.full(F, 2), m.tri_phys,
                                    m.tri_phys, ids[m.tris]])
        buf = io.StringIO()
        np.savetxt(buf, tri_rows, fmt="%d")
        f.write(buf.getvalue())
        tet_rows = np.column_stack([np.arange(F + 1, F + T + 1), np.full(T, 4), np.full(T, 2), m.tet_phys,
                                    m.tet_phys, ids[m.tets]])
        buf = io.StringIO()
        np.savetxt(buf, tet_rows, fmt="%d")
        f.write(buf.getvalue())
        f.write("$EndElements\n")


# ---------------------------------------------------------------------------
# The named workloads (sizes, seeds and materials)
# ---------------------------------------------------------------------------
WR90_A = 22.86e-3
WR90_B = 10.16e-3


def wr90_guide(nx=16, ny=8, nz=32, length=50e-3, seed=1):
    """Air-filled WR-90 guide, PEC walls, ports at z=0 and z=L."""
    return box_mesh(WR90_A, WR90_B, length, nx, ny, nz, seed=seed)


def slab_eps(mesh: BoxMesh, z0=30e-3, z1=70e-3, lo=1.0, hi=4.0, tan_d=0.002, seed=2) -> np.ndarray:
    """Per-tet complex permittivity: tets whose centroid z lies in [z0, z1] get
    eps' ~ U[lo, hi] (seeded per tet) times (1 - j tan_d); the rest are air."""
    T = mesh.tets.shape[0]
    zc = mesh.centroids()[:, 2]
    u = uniform01(seed, np.arange(T, dtype=np.int64))
    epsr = lo + (hi - lo) * u
    inside = (zc >= z0) & (zc <= z1)
    eps = np.ones(T, dtype=np.complex128)
    eps[inside] = epsr[inside] * (1.0 - 1j * tan_d)
    return eps


# Tags of the generated unit cells: periodic faces x0/x1/y0/y1, ports z=0 (2) and z=top
# (3), PEC patch 1, substrate 101, air 100.
CELL_TAGS = dict(x0=10, x1=11, y0=12, y1=13, port_bottom=2, port_top=3, patch=1, substrate=101, air=100)


def unit_cell(period: float = 5e-3, n: int = 10, nz: int = 12, h_sub: float = 0.5e-3, h_air: float = 10e-3,
              patch=None, seed: int = 0, jitter: float = 0.15) -> BoxMesh:
    """Periodic unit cell: n x n x nz hexes over period x period x (h_sub + h_air); the
    substrate takes n_sub = max(2, round(nz h_sub / (h_sub + h_air))) hex layers; the
    optional PEC patch (w, l), centred, is snapped to grid lines on the substrate top
    z = h_sub (a flat plane) and written as tagged triangles (tag 1)."""
    t = CELL_TAGS
    n_sub = max(2, int(round(nz * h_sub / (h_sub + h_air))))
    zp = np.concatenate([np.linspace(0.0, h_sub, n_sub + 1), np.linspace(h_sub, h_sub + h_air, nz - n_sub + 1)[1:]])
    g = box_mesh(period, period, h_sub + h_air, n, n, nz, seed=seed, jitter=jitter, zplanes=zp,
                 tags=(t["patch"], t["port_bottom"], t["port_top"], t["air"]),
                 side_tags={"x0": t["x0"], "x1": t["x1"], "y0": t["y0"], "y1": t["y1"]}, zflat=(n_sub,),
                 periodic_faces=True)
    hexk = (np.arange(g.tets.shape[0]) // 6) // (n * n)
    g.tet_phys = np.where(hexk < n_sub, t["substrate"], t["air"]).astype(np.int32)
    if patch is not None:
        w, l = patch
        h = period / n
        i0 = int(round((period - w) / 2 / h))
        j0 = int(round((period - l) / 2 / h))
        i1 = i0 + max(1, int(round(w / h)))
        j1 = j0 + max(1, int(round(l / h)))
        fj, fi = np.meshgrid(np.arange(j0, min(j1, n)), np.arange(i0, min(i1, n)), indexing="ij")
        fi = fi.ravel(); fj = fj.ravel()
        nid = lambda ii, jj, kk: ii + (n + 1) * (jj + (n + 1) * kk)
        c00, c10 = nid(fi, fj, n_sub), nid(fi + 1, fj, n_sub)
        c11, c01 = nid(fi + 1, fj + 1, n_sub), nid(fi, fj + 1, n_sub)
        pt = np.stack([np.stack([c00, c10, c11], -1), np.stack([c00, c11, c01], -1)], 1).reshape(-1, 3)
        g.tris = np.concatenate([g.tris, pt.astype(np.int32)])
        g.tri_phys = np.concatenate([g.tri_phys, np.full(pt.shape[0], t["patch"], np.int32)])
    return g
