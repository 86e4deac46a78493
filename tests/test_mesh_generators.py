"""Structured generators (deck inputs for the oracle and the GPU): the hex8
O-grid around a cylinder.  The checks are closed forms of the geometry:
the elements' volumes (the oracle's CalcLength, CalcLength.h:34-110) sum to
the polygonal annulus nt/2 sin(2 pi/nt) (r1^2 - r0^2) lz exactly, every
element has det J > 0, the ring closes without seam nodes, and the bc codes
and Dirichlet flags sit on the faces and nodes the docstring names."""
import numpy as np
import pytest

from cfdv_amd import mesh as M
from oracle import oracle as O


@pytest.mark.parametrize("nt,nr,nz,ratio", [(16, 6, 2, 1.04), (24, 5, 3, 1.2)])
def test_cylinder_ogrid_geometry(nt, nr, nz, ratio):
    r0, r1, lz = 0.5, 10.0, 1.0
    m, th, U = M.cylinder_ogrid(nt, nr, nz, r0=r0, r1=r1, lz=lz, ratio=ratio)
    assert m.n_elem == nt * nr * nz and m.n_node == nt * (nr + 1) * (nz + 1)
    assert np.array_equal(np.unique(m.conn), np.arange(m.n_node))     # every node used, no seam copies
    area, cleng = O.calc_length(m)
    vol = np.asarray(area)
    assert (vol > 0).all()
    exact = nt / 2 * np.sin(2 * np.pi / nt) * (r1 ** 2 - r0 ** 2) * lz
    assert abs(vol.sum() - exact) <= 1e-12 * exact
    # the last ring of cells closes onto the first: element (i = nt-1) shares
    # its eta = +1 face with element (i = 0)'s eta = -1 face
    e_last = (nt - 1) * nr          # j = 0, i = nt-1, k = 0
    e_first = 0
    f3 = set(m.conn[e_last, M.FACE_NODES_3D[3]].tolist())
    f1 = set(m.conn[e_first, M.FACE_NODES_3D[1]].tolist())
    assert f3 == f1
    # bc codes: wall 1 on the inner faces, -1 / 8 on the outer, 2 at the span ends
    codes, counts = np.unique(m.bc, return_counts=True)
    cnt = dict(zip(codes.tolist(), counts.tolist()))
    assert cnt[1] == nt * nz and cnt[2] == 2 * nt * nr and cnt[-1] + cnt[8] == nt * nz
    wall = np.unique(m.conn[m.bc[:, 4] == 1][:, M.FACE_NODES_3D[4]])
    assert np.allclose(np.hypot(m.x[wall, 0], m.x[wall, 1]), r0)
    far = np.unique(m.conn[m.bc[:, 2] == -1][:, M.FACE_NODES_3D[2]])
    assert np.allclose(np.hypot(m.x[far, 0], m.x[far, 1]), r1) and (m.x[far, 0] <= 1e-12).all()
    # Dirichlet: isothermal wall fixes momenta and energy, the farfield all
    assert (m.dirichlet[wall, 1:] == 1).all()
    assert (m.dirichlet[far] == 1).all()
    # state: freestream, wall at rest at the wall temperature
    prim = M.primitives(U, 3, th.Cv, th.gamma)
    assert np.allclose(prim["v"][wall], 0.0) and np.allclose(prim["T"][wall], th.twall)
