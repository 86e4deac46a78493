"""Host refinement through the C ABI (cfdv_refine_*) against the oracle.

* The golden fixtures (tests/golden/refine_*.npz: two --adap passes of the
  oracle on lid50 and 3dtestlagrange, quirk Q14 node matching) are
  reproduced bit for bit: element list and level breakpoints, connectivity,
  bc, coordinates, U, U0, Dirichlet flags, adap, Hnm and refine levels.
* Generated meshes in both node-matching modes, one and two passes, equal
  the oracle run on the same inputs.
* The closed forms of the oracle's own tests hold for the product too
  (uniform refinement, Q14 node count), and the states where the reference
  blocks or recurses without end come back as CFDV_ERR_REFINE.
No GPU is used: the refinement is host code in the library.
"""
import os

import numpy as np
import pytest

from cfdv_amd import _lib
from cfdv_amd import mesh as M
from cfdv_amd import refine as PR
from oracle import refine as OR

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _oracle_pass(mesh, U, U0, level, flags, mode):
    rm = OR.RMesh.from_arrays(mesh.ndim, mesh.nnod, mesh.neqn, mesh.conn, mesh.bc, mesh.x, U, U0,
                              mesh.dirichlet, mesh.adap, mesh.hnm, level)
    el, rl = OR.adaptive_refine_list(rm, flags)
    OR.MeshRefine(rm, el, rl, mode="parity" if mode == PR.PARITY else "3d")
    conn, bc, x, Un, U0n, d, adap, hnm, lv = rm.arrays()
    return dict(elelist=np.array(el), breaks=np.array(rl), conn=conn, bc=bc, x=x, U=Un, U0=U0n,
                dirichlet=d, adap=adap, hnm=hnm, level=lv)


def _product_pass(mesh, U, U0, level, flags, mode):
    r = PR.RefinableMesh(mesh, U, U0, level)
    el, rl = r.refine_list(flags)
    r.apply(el, rl, mode)
    m2, Un, U0n, lv = r.get()
    adap = m2.adap if m2.adap is not None else np.zeros(m2.n_elem, np.uint8)
    hnm = m2.hnm if m2.hnm is not None else np.tile(np.eye(m2.nnod), (m2.n_elem, 1, 1))
    return dict(elelist=el, breaks=rl, conn=m2.conn, bc=m2.bc, x=m2.x, U=Un, U0=U0n,
                dirichlet=m2.dirichlet, adap=adap, hnm=hnm, level=lv), m2


def _assert_same(a, b, what):
    for k in ("elelist", "breaks", "conn", "bc", "x", "U", "U0", "dirichlet", "adap", "level"):
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), f"{what}: {k} differs"
    sel = np.asarray(a["adap"]).astype(bool)
    assert np.array_equal(np.asarray(a["hnm"])[sel], np.asarray(b["hnm"])[sel]), f"{what}: Hnm differs"


@pytest.mark.parametrize("name", ["lid50", "3dtestlagrange"])
def test_fixture_two_passes(name):
    g = np.load(os.path.join(GOLD, f"refine_{name}.npz"))
    mesh, _, _, _ = M.read_deck(os.path.join(GOLD, "decks", name))
    U, U0, level = g["U"], g["U0"], None
    for p in range(2):
        want = {k: g[f"p{p}_{k}"] for k in ("elelist", "breaks", "conn", "bc", "x", "U", "U0", "dirichlet",
                                           "adap", "hnm", "level")}
        got, m2 = _product_pass(mesh, U, U0, level, g[f"p{p}_flags"], PR.PARITY)
        _assert_same(got, want, f"{name} pass {p}")
        mesh, U, U0, level = m2, got["U"], got["U0"], got["level"]


def _flags(n, every, off):
    return (np.arange(n) % every == off).astype(np.uint8)


@pytest.mark.parametrize("case,mode", [("lid", PR.PARITY), ("box", PR.MODE_3D), ("box", PR.PARITY),
                                       ("plate", PR.MODE_3D)])
def test_matches_oracle_two_passes(case, mode):
    if case == "lid":
        mesh, _, U = M.lid_cavity(8)
    elif case == "box":
        mesh, _, U = M.mixed_bc_box(4, 3, 3, seed=5)
    else:
        mesh, _, U = M.flat_plate(6, 4, 3)
    U0 = U * 0.999
    level = None
    for p, (every, off) in enumerate(((7, 3), (5, 1))):
        flags = _flags(mesh.n_elem, every, off)
        want = _oracle_pass(mesh, U, U0, level, flags, mode)
        got, m2 = _product_pass(mesh, U, U0, level, flags, mode)
        _assert_same(got, want, f"{case} pass {p}")
        mesh, U, U0, level = m2, got["U"], got["U0"], got["level"]


def test_uniform_and_q14_closed_forms():
    m, _, U = M.lid_cavity(3)
    got, m2 = _product_pass(m, U, U, None, np.ones(9, np.uint8), PR.PARITY)
    f, _, _ = M.lid_cavity(6)
    assert m2.n_elem == 36 and m2.n_node == 49 and not got["adap"].any()
    pos = {tuple(np.round(p * 1e9).astype(np.int64)) for p in f.x}
    assert {tuple(np.round(p * 1e9).astype(np.int64)) for p in m2.x} == pos
    x, conn = M.box_3d(1, 1, 1, lx=1.0, ly=1.0, lz=1.0)
    bc = np.full((1, 6), 1, np.int32)
    box = M.Mesh(3, 8, 5, conn, x, bc, M.dirichlet_flags(3, 5, 8, conn, bc, True))
    U = np.ones((8, 5))
    _, mp = _product_pass(box, U, U, None, np.ones(1, np.uint8), PR.PARITY)
    _, m3 = _product_pass(box, U, U, None, np.ones(1, np.uint8), PR.MODE_3D)
    assert mp.n_node == 13 and m3.n_node == 27


def test_endless_recursion_is_an_error():
    m, _, U = M.lid_cavity(3)
    order = [4, 1, 2, 3, 0, 5, 6, 7, 8]
    m = M.Mesh(2, 4, 4, m.conn[order], m.x, m.bc[order], m.dirichlet)
    r = PR.RefinableMesh(m, U)
    r.apply([0], [0, 1])
    level = r.get()[3]
    flags = np.zeros(len(level), np.uint8)
    flags[9] = 1
    with pytest.raises(PR.RefineError) as ei:
        r.refine_list(flags)
    assert ei.value.code == _lib.CFDV_ERR_REFINE
    assert "find_dependancies" in str(ei.value)


def test_bad_arguments():
    m, _, U = M.lid_cavity(2)
    r = PR.RefinableMesh(m, U)
    with pytest.raises(PR.RefineError) as ei:
        r.apply([7], [0, 1])                 # element number out of range
    assert ei.value.code == _lib.CFDV_ERR_ARG
    with pytest.raises(PR.RefineError):
        r.apply([0], [0, 2])                 # breakpoint beyond the list
    with pytest.raises(PR.RefineError):
        r.apply([0], [0, 1], mode=5)
    bad = M.Mesh(2, 4, 4, m.conn + 100, m.x, m.bc, m.dirichlet)
    with pytest.raises(PR.RefineError):
        PR.RefinableMesh(bad, U)
