"""Pins the refinement oracle (oracle/refine.py) by what does not depend on it.

* ElementAssociation: the reference's own loops (the O(n_node n_elem nnod)
  node scan and the nested face/edge scans, ElementAssociation.h:62-245)
  run literally on small meshes, including a refined mesh whose hanging
  elements hold parent nodes, equal the oracle's counting form; on a
  structured grid the face neighbours are the grid neighbours.
* MeshRefine2D: refining every element of a uniform grid gives the grid at
  half spacing (node positions, element squares, boundary codes, Dirichlet
  flags, no hanging nodes); a linear field is interpolated exactly; one
  element refined leaves hanging rows whose interpolated corners are the
  children's geometric corners; refining the neighbour afterwards removes
  them (do_hanging_nodes_2d/3d) and the two passes end at the uniform grid.
  Quirk Q14 (x, y-only node match) on one hex: 13 nodes instead of 27.
* AdaptiveRefineList: a hand-derived dependency closure with level buckets,
  and the recursion of find_dependancies on elements[k] (k the position in
  the face list) that never ends when element k is refined and has a
  coarser neighbour.
"""
import numpy as np
import pytest

from cfdv_amd import mesh as M
from oracle import refine as OR


def _rm(mesh, U=None, level=None):
    if U is None:
        U = np.zeros((mesh.n_node, mesh.neqn))
    return OR.RMesh.from_arrays(mesh.ndim, mesh.nnod, mesh.neqn, mesh.conn, mesh.bc, mesh.x, U, U.copy(),
                                mesh.dirichlet, mesh.adap, mesh.hnm, level)


def _box(nx, ny, nz, h=0.5, codes=dict(zmin=1, ymin=2, xmax=-1, ymax=2, xmin=-1, zmax=6)):
    """axis-aligned block of cubes of edge h"""
    x, conn = M.box_3d(nx, ny, nz, lx=nx * h, ly=ny * h, lz=nz * h)
    bc = M._face_bc_3d(nx, ny, nz, **codes)
    d = M.dirichlet_flags(3, 5, x.shape[0], conn, bc, False)
    return M.Mesh(3, 8, 5, conn, x, bc, d)


def _brute_association(m: OR.RMesh):
    """ElementAssociation::reset as written (ElementAssociation.h:55-117)"""
    nodes = []
    for i in range(m.n_node):
        lst = []
        for j in range(m.n_elem):
            for k in range(m.nnod):
                if m.conn[j][k] == i:
                    lst.append(j)
        nodes.append(lst)
    F, B, E = [], [], []
    for i in range(m.n_elem):
        f = [[] for _ in range(m.nface)]
        edge = [[] for _ in range(12)]
        if m.nnod == 4:
            for j in range(m.nface):
                n0, n1 = (m.conn[i][t] for t in OR.FACE_NODES_2D[j])
                for a in nodes[n0]:
                    for b in nodes[n1]:
                        if a == b and b != i:
                            f[j].append(a)
        else:
            for j in range(m.nface):
                n0, n1, n2, n3 = (m.conn[i][t] for t in OR.FACE_NODES_3D[j])
                for a in nodes[n0]:
                    for b in nodes[n1]:
                        for c in nodes[n2]:
                            for d in nodes[n3]:
                                if a == b and b == c and c == d and b != i:
                                    f[j].append(a)
            for j, (p, q) in enumerate(OR.EDGES_3D):
                n0, n1 = m.conn[i][p], m.conn[i][q]
                for a in nodes[n0]:
                    for b in nodes[n1]:
                        if a == b and b != i:
                            match = any(a == x for fl in f for x in fl)
                            if not match:
                                edge[j].append(a)
        bnd = [False] * m.nface
        for face, _ in m.faces[i]:
            bnd[face] = True
        F.append(f)
        B.append(bnd)
        E.append(edge)
    return F, B, E


def _refined_once(mesh, elems, mode="3d"):
    rm = _rm(mesh)
    flags = np.zeros(mesh.n_elem, dtype=bool)
    flags[list(elems)] = True
    el, rl = OR.adaptive_refine_list(rm, flags)
    OR.MeshRefine(rm, el, rl, mode=mode)
    return rm


@pytest.mark.parametrize("case", ["lid3", "box", "lid3_refined", "box_refined"])
def test_association_equals_reference_loops(case):
    if case.startswith("lid3"):
        m, _, U = M.lid_cavity(3)
        rm = _refined_once(m, [4], "parity") if case.endswith("refined") else _rm(m, U)
    else:
        m = _box(3, 2, 2)
        rm = _refined_once(m, [1, 8], "3d") if case.endswith("refined") else _rm(m)
    ea = OR.element_association(rm)
    F, B, E = _brute_association(rm)
    assert ea.f == F
    assert ea.boundary == B
    assert ea.edge == E


def test_association_structured_grid():
    nx, ny, nz = 3, 2, 2
    rm = _rm(_box(nx, ny, nz))
    ea = OR.element_association(rm)
    # face index -> (di, dj, dk) of the neighbour (faces 0 z-min, 1 y-min,
    # 2 x-max, 3 y-max, 4 x-min, 5 z-max)
    step = {0: (0, 0, -1), 1: (0, -1, 0), 2: (1, 0, 0), 3: (0, 1, 0), 4: (-1, 0, 0), 5: (0, 0, 1)}
    for e in range(nx * ny * nz):
        i, j, k = e % nx, (e // nx) % ny, e // (nx * ny)
        for f, (di, dj, dk) in step.items():
            a, b, c = i + di, j + dj, k + dk
            want = [a + nx * (b + ny * c)] if (0 <= a < nx and 0 <= b < ny and 0 <= c < nz) else []
            assert ea.f[e][f] == want
        # edge neighbours: the elements across an edge that are no face neighbour
        want_edges = set()
        for di in (-1, 0, 1):
            for dj in (-1, 0, 1):
                for dk in (-1, 0, 1):
                    if sorted(map(abs, (di, dj, dk))) != [0, 1, 1]:
                        continue
                    a, b, c = i + di, j + dj, k + dk
                    if 0 <= a < nx and 0 <= b < ny and 0 <= c < nz:
                        want_edges.add(a + nx * (b + ny * c))
        got = set(x for lst in ea.edge[e] for x in lst)
        assert got == want_edges


def _key(p):
    return tuple(np.round(np.asarray(p) * 1e9).astype(np.int64))


def _match_uniform(rm: OR.RMesh, fine: M.Mesh):
    """rm must equal the mesh ``fine`` up to numbering: node positions,
    elements (by their node positions in local order), bc per element,
    Dirichlet flags per node; no hanging nodes."""
    X = np.array(rm.x)
    assert X.shape[0] == fine.n_node
    pos = {_key(p): i for i, p in enumerate(fine.x)}
    assert len(pos) == fine.n_node
    node_map = np.array([pos[_key(p)] for p in X])      # exact positions of the fine grid
    assert len(set(node_map.tolist())) == fine.n_node
    conn, bc, _, _, _, d, adap, _, _ = rm.arrays()
    assert not adap.any()
    fine_elems = {tuple(r): e for e, r in enumerate(fine.conn.tolist())}
    for e in range(rm.n_elem):
        fe = fine_elems[tuple(node_map[conn[e]].tolist())]
        assert np.array_equal(bc[e], fine.bc[fe])
    assert np.array_equal(d, fine.dirichlet[node_map])


def test_uniform_refinement_2d():
    m, _, U = M.lid_cavity(3)
    rm = _refined_once(m, range(m.n_elem), "parity")
    f, _, _ = M.lid_cavity(6)
    _match_uniform(rm, f)
    assert rm.level == [1] * 36


def test_uniform_refinement_3d():
    m = _box(2, 2, 2, h=0.5)
    rm = _refined_once(m, range(m.n_elem), "3d")
    _match_uniform(rm, _box(4, 4, 4, h=0.25))


def test_linear_field_interpolated_exactly():
    m = _box(2, 2, 1)
    rng = np.random.default_rng(1)
    A = rng.standard_normal(5)
    B = rng.standard_normal((5, 3))
    U = A[None, :] + m.x @ B.T
    rm = _rm(m, U)
    flags = np.ones(m.n_elem, dtype=bool)
    el, rl = OR.adaptive_refine_list(rm, flags)
    OR.MeshRefine(rm, el, rl, mode="3d")
    X = np.array(rm.x)
    want = A[None, :] + X @ B.T
    assert np.allclose(np.array(rm.U), want, rtol=0, atol=1e-14)
    assert np.allclose(np.array(rm.U0), want, rtol=0, atol=1e-14)


def _hanging_geometry(rm: OR.RMesh, h, elems):
    """the elements (hanging rows applied) are axis-aligned cubes of edge h
    with the reference's corner order"""
    sgn = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
    X = np.array(rm.x)
    for e in elems:
        geo = rm.hnm[e] @ X[rm.conn[e]]
        assert np.allclose(geo - geo[0], sgn * h, atol=1e-14), e


def test_hanging_nodes_2d_two_passes():
    m, _, U = M.lid_cavity(2)
    rm = _rm(m, U)
    OR.MeshRefine(rm, *OR.adaptive_refine_list(rm, np.array([1, 0, 0, 0], bool)), mode="parity")
    # element 0 split; its children along faces 1 and 2 hang on the coarse faces
    assert rm.n_elem == 7 and rm.n_node == 9 + 3   # two boundary mid-edge nodes and the centre
    assert any(rm.adap)
    assert all(np.allclose(r.sum(axis=1), 1.0) for r in rm.hnm)
    h = 0.25
    for e in [0, 4, 5, 6]:
        geo = rm.hnm[e] @ np.array(rm.x)[rm.conn[e]]
        assert np.allclose(geo - geo[0], np.array([[0, 0], [1, 0], [1, 1], [0, 1]]) * h, atol=1e-14)
    # second pass: the other three; the result is the uniform 4x4 grid
    flags = np.zeros(rm.n_elem, bool)
    flags[[1, 2, 3]] = True
    el, rl = OR.adaptive_refine_list(rm, flags)
    OR.MeshRefine(rm, el, rl, mode="parity")
    f, _, _ = M.lid_cavity(4)
    _match_uniform(rm, f)


def test_hanging_nodes_3d_two_passes():
    m = _box(2, 2, 2, h=0.5)
    rm = _rm(m)
    flags = np.zeros(8, bool)
    flags[0] = True
    OR.MeshRefine(rm, *OR.adaptive_refine_list(rm, flags), mode="3d")
    assert rm.n_elem == 15
    # children on the three interior faces of element 0 (x-max 2, y-max 3,
    # z-max 5; addFaces_3d receivers 0,1,4,5 / 1,2,5,6 / 3,4,5,6 -> +8) hang
    assert [e for e in range(15) if rm.adap[e]] == [8, 9, 10, 11, 12, 13, 14]
    _hanging_geometry(rm, 0.25, [0] + list(range(8, 15)))
    # the other seven in a second pass: every hanging row is replaced
    # (face centres at 0.25, edges at 0.5, do_hanging_nodes_3d) and the
    # result is the uniform grid
    flags = np.zeros(rm.n_elem, bool)
    flags[1:8] = True
    el, rl = OR.adaptive_refine_list(rm, flags)
    OR.MeshRefine(rm, el, rl, mode="3d")
    _match_uniform(rm, _box(4, 4, 4, h=0.25))


def test_q14_xy_only_node_match():
    m = _box(1, 1, 1, h=1.0)
    rp = _refined_once(m, [0], "parity")
    r3 = _refined_once(m, [0], "3d")
    assert r3.n_node == 27 and rp.n_node == 13
    # the 5 new nodes of parity mode are the centres of faces 0..4 seen in x, y
    newx = {tuple(p[:2]) for p in np.array(rp.x)[8:]}
    assert newx == {(0.5, 0.5), (0.5, 0.0), (1.0, 0.5), (0.5, 1.0), (0.0, 0.5)}


def _renumber(mesh, order):
    """the same mesh with element order[i] as element i"""
    return M.Mesh(mesh.ndim, mesh.nnod, mesh.neqn, mesh.conn[order], mesh.x, mesh.bc[order], mesh.dirichlet)


def test_refine_list_closure_and_buckets():
    m, _, U = M.lid_cavity(3)
    rm = _rm(m, U)
    OR.MeshRefine(rm, [4], [0, 1], mode="parity")
    assert rm.level[4] == 1 and rm.level[9:12] == [1, 1, 1]
    # flag child 10 (upper-right quarter of element 4): its coarse face
    # neighbours are elements 5 (right) and 7 (top); levels bucket them first
    flags = np.zeros(rm.n_elem, bool)
    flags[10] = True
    el, rl = OR.adaptive_refine_list(rm, flags)
    assert el == [5, 7, 10] and rl == [0, 2, 3]
    # no level above 0: the flagged elements in order, one bucket
    el, rl = OR.adaptive_refine_list(_rm(m, U), np.array([0, 1, 0, 0, 1, 0, 0, 0, 1], bool))
    assert el == [1, 4, 8] and rl == [0, 3]


def test_refine_list_recursion_on_face_list_position():
    m, _, U = M.lid_cavity(3)
    order = [4, 1, 2, 3, 0, 5, 6, 7, 8]          # the centre element first
    m = _renumber(m, order)
    rm = _rm(m, U)
    OR.MeshRefine(rm, [0], [0, 1], mode="parity")
    flags = np.zeros(rm.n_elem, bool)
    flags[9] = True                                # lower-right quarter of element 0
    with pytest.raises(OR.RefineError):
        OR.adaptive_refine_list(rm, flags)
