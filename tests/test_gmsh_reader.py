"""The library's gmsh v2.2 reader (efm_gmsh_read, host code) against the oracle's
load_gmsh_v2 (mesh_gmsh.cpp:15-101) on the golden meshes and a generated one, against the
generator's own arrays (exact: coordinates are written with 17 significant digits), and on
hand-written files covering the record forms the reference reads: Line2 boundary lines,
skipped element types, zero and extra tags, blank and CRLF lines, IDs that are not 1..N."""
import glob
import os

import numpy as np
import pytest

from conftest import ROOT

GOLDEN = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "*.msh")))


def _read(path, threads=0):
    from edgefem_amd import _capi
    return _capi.gmsh_read(path, threads)


def _same_as_oracle(d, om):
    assert np.array_equal(d["node_ids"], om.node_ids)
    assert np.array_equal(d["xyz"], om.xyz)
    assert np.array_equal(d["tet_ids"], om.tet_ids) and np.array_equal(d["tet_phys"], om.tet_phys)
    assert np.array_equal(d["tri_ids"], om.tri_ids) and np.array_equal(d["tri_phys"], om.tri_phys)


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_golden_meshes_match_oracle(path):
    from oracle import edgefem_oracle as O
    d = _read(path)
    _same_as_oracle(d, O.load_gmsh_v2(path))
    # the oracle drops Line2 records; count them from the file's type column
    with open(path) as f:
        text = f.read()
    body = text.split("$Elements\n", 1)[1].split("$EndElements", 1)[0].split("\n")[1:]
    types = [int(r.split()[1]) for r in body if r.strip()]
    assert d["line_ids"].shape == (types.count(1), 2)
    assert d["tet_ids"].shape[0] == types.count(4) and d["tri_ids"].shape[0] == types.count(2)


def test_generated_mesh_exact_and_thread_count_independent(tmp_path):
    from edgefem_amd import meshgen
    from oracle import edgefem_oracle as O
    m = meshgen.box_mesh(22.86e-3, 10.16e-3, 50e-3, 7, 5, 9, seed=3)
    path = str(tmp_path / "box.msh")
    meshgen.write_msh(path, m)
    F, T = m.tris.shape[0], m.tets.shape[0]
    ref = _read(path, 1)
    assert np.array_equal(ref["node_ids"], m.node_ids)
    assert np.array_equal(ref["xyz"], m.xyz)  # %.17g round-trips every double
    assert np.array_equal(ref["tet_ids"], m.node_ids[m.tets])
    assert np.array_equal(ref["tri_ids"], m.node_ids[m.tris])
    assert np.array_equal(ref["tet_phys"], m.tet_phys) and np.array_equal(ref["tri_phys"], m.tri_phys)
    # the writer numbers triangles 1..F, then tets F+1..F+T
    assert np.array_equal(ref["tri_elem"], np.arange(1, F + 1))
    assert np.array_equal(ref["tet_elem"], np.arange(F + 1, F + T + 1))
    _same_as_oracle(ref, O.load_gmsh_v2(path))
    for k in (2, 3, 7, 16, 64):
        d = _read(path, k)
        for key, v in ref.items():
            assert np.array_equal(d[key], v), (k, key)


HAND = """$MeshFormat
2.2 0 8
$EndMeshFormat
$PhysicalNames
1
3 100 "air"
$EndPhysicalNames
$Nodes
5
10 0 0 0
20 1.5 0 0

30 0 2.5e-1 0
40 0 0 +3
50 1 1 1
$EndNodes
$Elements
7
1 15 2 9 9 10
2 1 2 5 5 10 20
3 2 2 7 7 10 20 30
4 4 2 100 100 10 20 30 40
5 4 0 20 30 40 50
6 2 3 8 1 2 20 30 40
7 1 1 6 30 40
$EndElements
"""


def _check_hand(d):
    assert d["node_ids"].tolist() == [10, 20, 30, 40, 50]
    assert d["xyz"].tolist() == [[0, 0, 0], [1.5, 0, 0], [0, 0.25, 0], [0, 0, 3], [1, 1, 1]]
    assert d["tet_ids"].tolist() == [[10, 20, 30, 40], [20, 30, 40, 50]]
    assert d["tet_phys"].tolist() == [100, 0] and d["tet_elem"].tolist() == [4, 5]
    assert d["tri_ids"].tolist() == [[10, 20, 30], [20, 30, 40]]
    assert d["tri_phys"].tolist() == [7, 8] and d["tri_elem"].tolist() == [3, 6]
    assert d["line_ids"].tolist() == [[10, 20], [30, 40]] and d["line_phys"].tolist() == [5, 6]


@pytest.mark.parametrize("eol", ["\n", "\r\n"], ids=["lf", "crlf"])
def test_hand_written_records(tmp_path, eol):
    path = str(tmp_path / "hand.msh")
    with open(path, "w", newline="") as f:
        f.write(HAND.replace("\n", eol))
    for k in (1, 4):
        _check_hand(_read(path, k))
    if eol == "\n":
        from oracle import edgefem_oracle as O
        _same_as_oracle(_read(path), O.load_gmsh_v2(path))


def test_errors(tmp_path):
    from edgefem_amd import _capi
    missing = str(tmp_path / "nope.msh")
    with pytest.raises(_capi.EdgeFemError, match="Failed to open mesh file: " + missing):
        _read(missing)
    bad = str(tmp_path / "short.msh")
    with open(bad, "w") as f:
        f.write(HAND.replace("$Nodes\n5\n", "$Nodes\n6\n"))
    with pytest.raises(_capi.EdgeFemError, match=r"\$Nodes count"):
        _read(bad)
    bad2 = str(tmp_path / "garbled.msh")
    with open(bad2, "w") as f:
        f.write(HAND.replace("4 4 2 100 100 10 20 30 40", "4 4 2 100 100 10 20 x 40"))
    with pytest.raises(_capi.EdgeFemError, match="malformed"):
        _read(bad2)


def test_dielectric_slab_fixture_regions():
    """tests/test_dielectric_slab_fresnel.cpp:93-170 on its own fixture
    (examples/dielectric_slab.msh): more than 100 nodes, non-empty port surfaces 1 and 2,
    and every tet in the air (100, 102) or dielectric (101) regions, read the same by the
    library's reader and the oracle."""
    from oracle import edgefem_oracle as O
    path = os.path.join(ROOT, "tests", "golden", "dielectric_slab.msh")
    d = _read(path)
    assert d["node_ids"].size > 100
    assert np.all(np.isin(d["tet_phys"], [100, 101, 102]))
    assert np.count_nonzero(d["tet_phys"] == 101) > 0 and np.count_nonzero(d["tet_phys"] != 101) > 0
    m = O.load_gmsh_v2(path)
    for tag in (1, 2):
        ids, xy, tri, lines, sel = O.extract_surface_mesh(m, tag)
        assert tri.shape[0] > 0 and sel.size == np.count_nonzero(d["tri_phys"] == tag)
