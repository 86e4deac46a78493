"""CPU restatement of the reference's adaptive refinement -- TEST INFRASTRUCTURE.

Restates, loop for loop, with element and node objects replaced by indices:

* ``ElementAssociation``  (src/adap/ElementAssociation.h:42-249)
* ``AdaptiveRefineList``  (src/adap/RefineListAdaptive.h:8-126)
* ``MeshRefine2D``        (src/adap/MeshRefine2D.h:39-872), 2D quad4 and 3D hex8

Imported only by tests/ (the checker).  The product path is the C++
implementation in cfdv_amd/host/refine.cpp behind the C ABI.

Model: element ``e`` has ``conn[e]`` (node numbers; the reference's
``Element::node`` pointers -- after refinement a hanging local node keeps
the parent's node, exactly as the reference keeps the pointer), ``faces[e]``
(list of ``(face index, bc)`` in the reference's ``Element::face`` order;
Face objects are never mutated after loading, so sharing a pointer and
copying the pair are the same), ``hnm[e]`` (nnod x nnod, H(N, M)),
``adap[e]`` and ``level[e]``.  Node ``n`` has ``x``, ``U``, ``U0`` and the
Dirichlet mask.  Element and node numbers are their indices
(LoadBinary.h:198,230; MeshRefine2D.h:357,469).
"""
from __future__ import annotations

import dataclasses
from typing import List

import numpy as np

# get_nodes_from_face_3d (get_face.h:106-165) and get_nodes_from_face (:80-89)
FACE_NODES_3D = ((3, 2, 1, 0), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7), (6, 7, 4, 5))
FACE_NODES_2D = ((0, 1), (1, 2), (2, 3), (3, 0))
# get_opposing_face_3d (get_face.h:201-226), get_opposing_face (:92-97)
OPPOSING_3D = (5, 3, 4, 1, 2, 0)
# edge_association_3d (ElementAssociation.h:140-156)
EDGES_3D = ((0, 1), (1, 2), (2, 3), (3, 0), (0, 4), (1, 5), (2, 6), (3, 7), (4, 5), (5, 6), (6, 7), (7, 4))
# get_edge_from_nodes (get_face.h:232-292): (ni, nj), ni < nj -> candidate node index
EDGE_CANDIDATE = {(0, 1): 7, (1, 2): 8, (2, 3): 9, (0, 3): 10, (0, 4): 11, (1, 5): 12,
                  (2, 6): 13, (3, 7): 14, (4, 5): 15, (5, 6): 16, (6, 7): 17, (4, 7): 18}

# candidate node -> isoparametric position: create_nodes_3d (MeshRefine2D.h:289-331)
ZETA_3D = {0: (0., 0., -1.), 1: (0., -1., 0.), 2: (1., 0., 0.), 3: (0., 1., 0.), 4: (-1., 0., 0.),
           5: (0., 0., 1.), 6: (0., 0., 0.),
           7: (0., -1., -1.), 8: (1., 0., -1.), 9: (0., 1., -1.), 10: (-1., 0., -1.),
           11: (-1., -1., 0.), 12: (1., -1., 0.), 13: (1., 1., 0.), 14: (-1., 1., 0.),
           15: (0., -1., 1.), 16: (1., 0., 1.), 17: (0., 1., 1.), 18: (-1., 0., 1.)}
# edge candidates and the two face candidates that must both be set (:317-330)
EDGE_FACES_3D = ((7, 0, 1), (8, 0, 2), (9, 0, 3), (10, 0, 4), (11, 4, 1), (12, 1, 2), (13, 2, 3),
                 (14, 3, 4), (15, 5, 1), (16, 5, 2), (17, 5, 3), (18, 5, 4))
ZETA_2D = {0: (0., -1.), 1: (1., 0.), 2: (0., 1.), 3: (-1., 0.), 4: (0., 0.)}

# Hnm rows per candidate (createElements_3d, MeshRefine2D.h:504-523; _2d :467-475)
_H3 = np.zeros((19, 8))
for _c, _cols in {0: (0, 1, 2, 3), 1: (0, 1, 4, 5), 2: (1, 2, 5, 6), 3: (2, 3, 6, 7), 4: (0, 3, 4, 7),
                  5: (4, 5, 6, 7)}.items():
    _H3[_c, list(_cols)] = 0.25
for _c, _cols in {7: (0, 1), 8: (1, 2), 9: (2, 3), 10: (0, 3), 11: (0, 4), 12: (1, 5), 13: (2, 6),
                  14: (3, 7), 15: (4, 5), 16: (5, 6), 17: (6, 7), 18: (4, 7)}.items():
    _H3[_c, list(_cols)] = 0.5
_H2 = np.zeros((5, 4))
for _c, _cols in {0: (0, 1), 1: (1, 2), 2: (2, 3), 3: (0, 3)}.items():
    _H2[_c, list(_cols)] = 0.5
HROWS_3D = _H3
HROWS_2D = _H2

# set_Element_Node calls: child -> [(local node, candidate)], in call order
# (MeshRefine2D.h:525-600; the parent is child 7, 'e'); 2D (:477-495)
CHILD_NODES_3D = (
    ((0, 7), (2, 8), (3, 0), (4, 1), (5, 12), (6, 2), (7, 6)),
    ((0, 0), (1, 8), (3, 9), (5, 2), (6, 13), (7, 3), (4, 6)),
    ((0, 10), (1, 0), (2, 9), (4, 4), (6, 3), (7, 14), (5, 6)),
    ((0, 11), (1, 1), (3, 4), (5, 15), (6, 5), (7, 18), (2, 6)),
    ((0, 1), (1, 12), (2, 2), (4, 15), (6, 16), (7, 5), (3, 6)),
    ((1, 2), (2, 13), (3, 3), (4, 5), (5, 16), (7, 17), (0, 6)),
    ((0, 4), (2, 3), (3, 14), (4, 18), (5, 5), (6, 17), (1, 6)),
    ((1, 7), (2, 0), (3, 10), (4, 11), (5, 1), (7, 4), (6, 6)),
)
CHILD_NODES_2D = (
    ((0, 0), (2, 1), (3, 4)),
    ((1, 1), (3, 2), (0, 4)),
    ((0, 3), (2, 2), (1, 4)),
    ((1, 0), (3, 3), (2, 4)),
)
# addFaces_3d (:706-761): face index -> receivers ('e' = -1, children 0..6)
FACE_RECEIVERS_3D = {0: (-1, 0, 1, 2), 4: (-1, 2, 3, 6), 1: (-1, 0, 3, 4), 2: (0, 1, 4, 5),
                     3: (1, 2, 5, 6), 5: (3, 4, 5, 6)}


class RefineError(RuntimeError):
    """A state in which the reference prints and blocks on stdin, or recurses
    without end; the restatement stops there."""


@dataclasses.dataclass
class RMesh:
    ndim: int
    nnod: int
    neqn: int
    conn: List[List[int]]
    faces: List[List[tuple]]
    hnm: List[np.ndarray]
    adap: List[bool]
    level: List[int]
    x: List[np.ndarray]
    U: List[np.ndarray]
    U0: List[np.ndarray]
    dirichlet: List[np.ndarray]

    @property
    def nface(self):
        return 6 if self.ndim == 3 else 4

    @property
    def n_elem(self):
        return len(self.conn)

    @property
    def n_node(self):
        return len(self.x)

    @classmethod
    def from_arrays(cls, ndim, nnod, neqn, conn, bc, x, U, U0, dirichlet, adap=None, hnm=None, level=None):
        """Faces in loader order: one per face index with bc != 0, ascending
        (loadElements, LoadBinary.h:228-234)."""
        ne = conn.shape[0]
        nface = bc.shape[1]
        faces = [[(f, int(bc[e, f])) for f in range(nface) if bc[e, f] != 0] for e in range(ne)]
        H = [np.eye(nnod) if hnm is None else np.array(hnm[e], dtype=np.float64) for e in range(ne)]
        return cls(ndim, nnod, neqn, [list(map(int, r)) for r in conn], faces, H,
                   [bool(adap[e]) if adap is not None else False for e in range(ne)],
                   [int(level[e]) if level is not None else 0 for e in range(ne)],
                   [np.array(r, dtype=np.float64) for r in x], [np.array(r, dtype=np.float64) for r in U],
                   [np.array(r, dtype=np.float64) for r in U0],
                   [np.array(r, dtype=bool) for r in dirichlet])

    def arrays(self):
        """conn, bc (face index -> code, SaveBinary.h:66-69), x, U, U0,
        dirichlet, adap, hnm [e][N][M], level"""
        ne = self.n_elem
        bc = np.zeros((ne, self.nface), dtype=np.int32)
        for e in range(ne):
            for f, b in self.faces[e]:
                bc[e, f] = b
        return (np.array(self.conn, dtype=np.int32), bc, np.array(self.x), np.array(self.U),
                np.array(self.U0), np.array(self.dirichlet, dtype=np.uint8),
                np.array(self.adap, dtype=np.uint8), np.array(self.hnm), np.array(self.level, dtype=np.int32))


# ---------------------------------------------------------------------------
# ElementAssociation (ElementAssociation.h:55-249)
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class Association:
    f: list          # [e][face] -> list of element numbers (with repeats)
    boundary: list   # [e][face] -> bool
    edge: list       # [e][12] -> list of element numbers (3D)


def element_association(m: RMesh) -> Association:
    # 0. node correlation (:62-77): node i lists element j once per local node
    # of j equal to i, j ascending
    node_e = [[] for _ in range(m.n_node)]
    for j in range(m.n_elem):
        for k in range(m.nnod):
            node_e[m.conn[j][k]].append(j)
    f = []
    bnd = []
    edg = []
    for i in range(m.n_elem):
        fi = [[] for _ in range(m.nface)]
        ei = [[] for _ in range(12)]
        if m.nnod == 4:
            # face_association_2d (:222-245)
            for j in range(m.nface):
                L0 = node_e[m.conn[i][FACE_NODES_2D[j][0]]]
                L1 = node_e[m.conn[i][FACE_NODES_2D[j][1]]]
                for a in L0:
                    for b in L1:
                        if a == b and b != i:
                            fi[j].append(a)
        else:
            # face_association_3d (:100-136)
            for j in range(m.nface):
                L = [node_e[m.conn[i][n]] for n in FACE_NODES_3D[j]]
                for a in L[0]:
                    c = L[1].count(a) * L[2].count(a) * L[3].count(a)
                    if a != i:
                        fi[j].extend([a] * c)
            # edge_association_3d (:139-213)
            face_members = set()
            for j in range(m.nface):
                face_members.update(fi[j])
            for j, (n0, n1) in enumerate(EDGES_3D):
                L0 = node_e[m.conn[i][n0]]
                L1 = node_e[m.conn[i][n1]]
                for a in L0:
                    for b in L1:
                        if a == b and b != i and a not in face_members:
                            ei[j].append(a)
        # boundary flags from the element's faces (:87-98)
        bi = [False] * m.nface
        for face, _bc in m.faces[i]:
            bi[face] = True
        f.append(fi)
        bnd.append(bi)
        edg.append(ei)
    return Association(f, bnd, edg)


# ---------------------------------------------------------------------------
# AdaptiveRefineList (RefineListAdaptive.h:8-126)
# ---------------------------------------------------------------------------
def adaptive_refine_list(m: RMesh, flags, max_depth: int = 256):
    """flags: s1 > 0.8 per element (the FDVParam pass, :28-38).
    Returns (elelist, refine_level breakpoints)."""
    ne = m.n_elem
    n = 0
    for i in range(ne):
        if m.level[i] > n:
            n = m.level[i]
    rl = [0] * (n + 2)
    ea = element_association(m)
    refine_list = [i for i in range(ne) if flags[i]]

    def find_dependancies(refine_level, elno, depth):
        if depth > max_depth:
            raise RefineError("find_dependancies recurses without end (RefineListAdaptive.h:84-88 "
                              "recurses on elements[k], k the index within the face list)")
        for j in range(m.nface):
            if not ea.boundary[elno][j]:
                for k, nb in enumerate(ea.f[elno][j]):
                    if m.level[nb] < refine_level:
                        refine_list.append(nb)
                        if m.level[k] > 0:
                            find_dependancies(m.level[k], k, depth + 1)
        for j in range(len(ea.edge[elno])):
            for nb in ea.edge[elno][j]:
                if m.level[nb] < refine_level:
                    refine_list.append(nb)

    rsize = len(refine_list)
    for i in range(rsize):
        el = refine_list[i]
        if m.level[el] != 0:
            find_dependancies(m.level[el], el, 0)
    # remove_redundant_elements (:114-126): keeps first occurrences
    seen = set()
    dedup = []
    for v in refine_list:
        if v not in seen:
            seen.add(v)
            dedup.append(v)
    elelist = []
    for i in range(1, n + 2):
        rl[i] = rl[i - 1]
        for v in dedup:
            if m.level[v] == i - 1:
                elelist.append(v)
                rl[i] += 1
    return elelist, rl


# ---------------------------------------------------------------------------
# MeshRefine2D (MeshRefine2D.h:39-872)
# ---------------------------------------------------------------------------
def _phi(ndim, zeta):
    z = zeta
    if ndim == 3:
        return [0.125 * (1 - z[0]) * (1 - z[1]) * (1 - z[2]), 0.125 * (1 + z[0]) * (1 - z[1]) * (1 - z[2]),
                0.125 * (1 + z[0]) * (1 + z[1]) * (1 - z[2]), 0.125 * (1 - z[0]) * (1 + z[1]) * (1 - z[2]),
                0.125 * (1 - z[0]) * (1 - z[1]) * (1 + z[2]), 0.125 * (1 + z[0]) * (1 - z[1]) * (1 + z[2]),
                0.125 * (1 + z[0]) * (1 + z[1]) * (1 + z[2]), 0.125 * (1 - z[0]) * (1 + z[1]) * (1 + z[2])]
    return [0.25 * (1 - z[0]) * (1 - z[1]), 0.25 * (1 + z[0]) * (1 - z[1]),
            0.25 * (1 + z[0]) * (1 + z[1]), 0.25 * (1 - z[0]) * (1 + z[1])]


class MeshRefine:
    """mode "parity": node dedupe on x and y only, |d| < tol (Q14, :337-361);
    mode "3d": on every coordinate."""

    def __init__(self, m: RMesh, elelist, refine_level, mode="parity", tol=10E-4):
        self.m = m
        self.mode = mode
        self.tol = tol
        self.nrnod = 5 if m.nnod == 4 else 19
        self.ea = element_association(m)
        n = len(refine_level) - 2
        for j in range(n + 1):
            for i in range(refine_level[j], refine_level[j + 1]):
                self.refine(elelist[i], elelist, n)
            if j != n:
                self.ea = element_association(m)

    # findNeighbors (:827-862)
    def find_neighbors(self, i, refinelist):
        m = self.m
        face = [False] * self.nrnod
        jlist = 0
        for j, v in enumerate(refinelist):
            if v == i:
                jlist = j
        for j in range(m.nface):
            for k in self.ea.f[i][j]:
                for l, v in enumerate(refinelist):
                    if v == k:
                        if m.level[k] == m.level[i] and l < jlist:
                            pass
                        else:
                            face[j] = True
        return face

    # addPriorRefine (:786-821)
    def add_prior_refine(self, nodelist, i):
        m = self.m
        bordering = 2 if m.nnod == 4 else 4
        prior = []
        for j in range(m.nface):
            if len(self.ea.f[i][j]) == bordering:
                nodelist[j] = True
                for nb in self.ea.f[i][j]:
                    if m.nnod == 4:
                        local = FACE_NODES_2D[(j + 2) % 4]
                    else:
                        local = FACE_NODES_3D[OPPOSING_3D[j]]
                    prior.append((nb, j, list(local)))
        return prior

    # createNode + addDirichlet + createNodeRef (:337-407, 627-657)
    def create_node_ref(self, zeta, e):
        m = self.m
        Phi = _phi(m.ndim, zeta)
        x = np.zeros(m.ndim)
        U = np.zeros(m.neqn)
        U0 = np.zeros(m.neqn)
        d = np.zeros(m.neqn)
        for N in range(m.nnod):
            g = m.conn[e][N]
            for i in range(m.neqn):
                U[i] += Phi[N] * m.U[g][i]
                U0[i] += Phi[N] * m.U0[g][i]
            for i in range(m.ndim):
                x[i] += Phi[N] * m.x[g][i]
        for N in range(m.nnod):
            g = m.conn[e][N]
            for i in range(m.neqn):
                d[i] += Phi[N] * float(m.dirichlet[g][i])
        dirichlet = np.array([not (d[i] < 1) for i in range(m.neqn)], dtype=bool)
        X = np.array(m.x)
        nd = 2 if self.mode == "parity" else m.ndim
        dist2 = np.zeros(X.shape[0])
        for c in range(nd):
            dc = X[:, c] - x[c]
            dist2 = dist2 + dc * dc
        hit = np.nonzero(np.sqrt(dist2) < self.tol)[0]
        if hit.size:
            return int(hit[0])
        m.x.append(x)
        m.U.append(U)
        m.U0.append(U0)
        m.dirichlet.append(dirichlet)
        return m.n_node - 1

    def create_nodes(self, nodelist, e, new_node):
        m = self.m
        if m.nnod == 4:
            for c in range(4):
                if nodelist[c]:
                    new_node[c] = self.create_node_ref(ZETA_2D[c], e)
            nodelist[4] = True
            new_node[4] = self.create_node_ref(ZETA_2D[4], e)
            return
        for c in range(6):
            if nodelist[c]:
                new_node[c] = self.create_node_ref(ZETA_3D[c], e)
        nodelist[6] = True
        new_node[6] = self.create_node_ref(ZETA_3D[6], e)
        for c, fa, fb in EDGE_FACES_3D:
            if nodelist[fa] and nodelist[fb]:
                nodelist[c] = True
                new_node[c] = self.create_node_ref(ZETA_3D[c], e)

    # createElements (:411-462) with createElements_3d/_2d and set_Element_Node
    def create_elements(self, e, nodelist, new_node):
        m = self.m
        nn = m.nnod
        ne0 = m.n_elem
        kids = list(range(ne0, ne0 + nn - 1))
        for c in kids:
            m.conn.append(list(m.conn[e]))
            m.faces.append([])
            m.hnm.append(np.eye(nn))
            m.adap.append(False)
            m.level.append(0)
        m.hnm[e] = np.eye(nn)
        table = CHILD_NODES_3D if nn == 8 else CHILD_NODES_2D
        rows = HROWS_3D if nn == 8 else HROWS_2D
        for ci, calls in enumerate(table):
            el = e if ci == len(table) - 1 else kids[ci]
            for elenode, cand in calls:
                if nodelist[cand]:
                    m.conn[el][elenode] = new_node[cand]
                else:
                    m.adap[el] = True
                    m.hnm[el][elenode, :] = rows[cand]
        if nn == 8:
            self.add_faces_3d(e, kids)
        else:
            self.add_faces_2d(e, kids)
        m.level[e] += 1
        for c in kids:
            m.level[c] = m.level[e]

    def add_faces_3d(self, e, kids):
        m = self.m
        faces = m.faces[e]
        m.faces[e] = []
        for i in range(len(faces) - 1, -1, -1):
            fc = faces[i]
            for r in FACE_RECEIVERS_3D[fc[0]]:
                m.faces[e if r < 0 else kids[r]].append(fc)

    def add_faces_2d(self, e, kids):
        m = self.m
        fl = m.faces[e]
        for i in range(len(fl) - 1, -1, -1):
            fc = fl[i]
            face = fc[0]
            if face == 0:
                m.faces[kids[0]].append(fc)
            elif face == 1:
                m.faces[kids[0]].append(fc)
                m.faces[kids[1]].append(fc)
                del fl[i]
            elif face == 2:
                m.faces[kids[1]].append(fc)
                m.faces[kids[2]].append(fc)
                del fl[i]
            elif face == 3:
                m.faces[kids[2]].append(fc)

    # update_hanging_nodes (:258-275)
    def update_hanging_nodes(self, el, node, indexi):
        m = self.m
        m.hnm[el][indexi, :] = 0.0
        m.hnm[el][indexi, indexi] = 1.0
        if np.array_equal(m.hnm[el], np.eye(m.nnod)):
            m.adap[el] = False
        m.conn[el][indexi] = node

    def do_hanging_nodes_3d(self, e, new_node, prior):
        m = self.m
        for elno, face, local in prior:
            H = m.hnm[elno]
            for N in range(8):
                if all(H[N, local[q]] == 0.25 for q in range(4)):
                    if new_node[face] != 0:
                        self.update_hanging_nodes(elno, new_node[face], N)
            for N in range(8):
                for ni in range(4):
                    for nj in range(ni + 1, 4):
                        if m.hnm[elno][N, local[ni]] == 0.50 and m.hnm[elno][N, local[nj]] == 0.50:
                            opp = FACE_NODES_3D[OPPOSING_3D[_face_3d(local)]]
                            ni0 = ni + 1 if ni in (0, 2) else ni - 1
                            nj0 = nj + 1 if nj in (0, 2) else nj - 1
                            a, b = opp[ni0], opp[nj0]
                            edge = EDGE_CANDIDATE.get((min(a, b), max(a, b)))
                            if edge is None:
                                raise RefineError("get_edge_from_nodes: no match (get_face.h:288-291)")
                            if new_node[edge] != 0:
                                self.update_hanging_nodes(elno, new_node[edge], N)

    def do_hanging_nodes_2d(self, e, new_node, prior):
        m = self.m
        for elno, face, local in prior:
            node_number = new_node[face]
            if m.level[elno] == m.level[e]:
                indexi = -1
                for j in range(m.nnod):
                    if m.hnm[elno][j, local[0]] == 0.5 and m.hnm[elno][j, local[1]] == 0.5:
                        indexi = j
                if indexi == -1:
                    raise RefineError("do_hanging_nodes_2d: no Hnm match (MeshRefine2D.h:241-248 "
                                      "prints and blocks on stdin)")
                self.update_hanging_nodes(elno, node_number, indexi)

    def refine(self, e, refinelist, max_refine_level):
        m = self.m
        nodelist = self.find_neighbors(e, refinelist)
        prior = self.add_prior_refine(nodelist, e)
        for face, _bc in m.faces[e]:
            nodelist[face] = True
        new_node = [0] * self.nrnod
        self.create_nodes(nodelist, e, new_node)
        self.create_elements(e, nodelist, new_node)
        if m.nnod == 4:
            self.do_hanging_nodes_2d(e, new_node, prior)
        else:
            self.do_hanging_nodes_3d(e, new_node, prior)


def _face_3d(local):
    """get_face_3d (get_face.h:168-183) on a local node list"""
    t = tuple(local)
    for f, nodes in enumerate(FACE_NODES_3D):
        if t == nodes:
            return f
    if t == (4, 5, 6, 7):
        return 5
    raise RefineError("get_face_3d: no match")
