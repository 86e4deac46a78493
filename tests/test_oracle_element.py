"""Pins of the oracle's element build (FDVGalerkin.h:121-797) and ApplyBC.

Independent references used here:
- exact finite-element integrals of a box hex8 element, built from 1D
  mass/derivative matrices (tensor products), for the uniform-state closed
  form R = M (x) I + dt^2/4 sum_ij K^ij (x) a_i a_j and the face term;
- invariants that hold for any state: the column-sum identity
  sum_N R(Nr, Ms) = delta_rs int Phi_M (partition of unity kills the g, h
  and numerical-diffusion terms) and conservation of the domain rhs
  (sum_N rhs(N, r) = 0 for an element without boundary faces);
- the ApplyBC row structure (ApplyBC.h:38-76, quirk Q5).
"""
import numpy as np
import pytest

from cfdv_amd import mesh as M
from oracle import oracle as O

# local hex node -> (a, b, c) corner indices (TestFunctions3D.h:12-19)
CORNER = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                   [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])


def _1d(h):
    m = h / 6.0 * np.array([[2.0, 1.0], [1.0, 2.0]])          # int phi_a phi_b
    s = 1.0 / h * np.array([[1.0, -1.0], [-1.0, 1.0]])        # int phi_a' phi_b'
    c = np.array([[-0.5, -0.5], [0.5, 0.5]])                  # int phi_a' phi_b
    return m, s, c


def _box_matrices(hs):
    """Mass M_NM and K^ij_NM = int dPhi_N/dx_i dPhi_M/dx_j on a box."""
    ms = [_1d(h) for h in hs]
    Mm = np.ones((8, 8))
    K = np.ones((3, 3, 8, 8))
    for N in range(8):
        for Mi in range(8):
            for d in range(3):
                a, b = CORNER[N, d], CORNER[Mi, d]
                m1, s1, c1 = ms[d]
                Mm[N, Mi] *= m1[a, b]
                for i in range(3):
                    for j in range(3):
                        if d == i and d == j:
                            K[i, j, N, Mi] *= s1[a, b]
                        elif d == i:
                            K[i, j, N, Mi] *= c1[a, b]          # derivative on N
                        elif d == j:
                            K[i, j, N, Mi] *= c1[b, a]          # derivative on M
                        else:
                            K[i, j, N, Mi] *= m1[a, b]
    return Mm, K


def _box_mesh(hs, bc_center=None, top_code=None):
    """3x3x3 box of equal cells; returns mesh with every bc 0 except optional
    top faces of the center column."""
    x, conn = M.box_3d(3, 3, 3, lx=3 * hs[0], ly=3 * hs[1], lz=3 * hs[2])
    bc = np.zeros((27, 6), np.int32)
    if top_code is not None:
        bc[22, 5] = top_code       # element (1,1,2): top face
    d = np.zeros((x.shape[0], 5), np.uint8)
    return M.Mesh(3, 8, 5, conn, x, bc, d)


def _uniform_state(mesh, th, v=(0.8, 0.3, -0.2), T=1.3, rho=1.1):
    n = mesh.n_node
    return M.conservative(np.full(n, rho), np.tile(np.array(v), (n, 1)), np.full(n, T), th.Cv)


def test_uniform_state_interior_element_closed_form():
    hs = (0.3, 0.2, 0.25)
    mesh = _box_mesh(hs)
    th = M.Thermo(gamma=1.4, mach=2.0, reynolds=500.0)
    U = _uniform_state(mesh, th)
    dt = 0.013
    R, rhs, s = O.build(th, mesh, U, dt, apply_bc=False)
    e = 13                                         # center element, no faces
    assert np.allclose(s[e], [0.0, 0.5, 0.0, 0.5])
    Mm, K = _box_matrices(hs)
    p = O.prim_arrays(O.unpack(th, mesh, U), 3)
    f = O.flow_struct(p["E"][0], p["T"][0], p["v"][0], np.zeros(3), U[0], np.zeros((5, 3)),
                      np.zeros((3, 3)), np.zeros(3), 0.0, 1.0)
    a, _, _ = O.calc_abc(th, 3, f)
    expect = np.kron(Mm, np.eye(5))
    for i in range(3):
        for j in range(3):
            expect += dt * dt / 4.0 * np.kron(K[i, j], a[i] @ a[j])
    assert np.allclose(R[e], expect, rtol=1e-12, atol=1e-14 * np.abs(expect).max())
    # uniform state: every gradient vanishes up to roundoff of the node values
    assert np.abs(rhs[e]).max() < 1e-18


def test_uniform_state_face_term_closed_form():
    """bc 7 top face: R gains sum over face points of
    -dt^2/4 sum_ij (a_i a_j) dpidx(M,j) phi_N en_i, en = +z * area/4 * w."""
    hs = (0.3, 0.2, 0.25)
    th = M.Thermo(gamma=1.4, mach=2.0, reynolds=500.0)
    dt = 0.011
    mesh0 = _box_mesh(hs)
    mesh1 = _box_mesh(hs, top_code=7)
    U = _uniform_state(mesh0, th)
    R0, _, _ = O.build(th, mesh0, U, dt, apply_bc=False)
    R1, rhs1, _ = O.build(th, mesh1, U, dt, apply_bc=False)
    e = 22
    p = O.prim_arrays(O.unpack(th, mesh0, U), 3)
    f = O.flow_struct(p["E"][0], p["T"][0], p["v"][0], np.zeros(3), U[0], np.zeros((5, 3)),
                      np.zeros((3, 3)), np.zeros(3), 0.0, 1.0)
    a, _, _ = O.calc_abc(th, 3, f)
    # top face: local nodes (6, 7, 4, 5); exact surface integral of
    # phi_N d(phi_M)/dx_j over the face times n_z for j = x, y
    m1x, _, c1x = _1d(hs[0])
    m1y, _, c1y = _1d(hs[1])
    face = [6, 7, 4, 5]
    add = np.zeros((40, 40))
    for N in face:
        for Mi in face:
            aN, bN = CORNER[N, 0], CORNER[N, 1]
            aM, bM = CORNER[Mi, 0], CORNER[Mi, 1]
            Ix = c1x[aM, aN] * m1y[bN, bM]       # int phi_N dphi_M/dx
            Iy = m1x[aN, aM] * c1y[bM, bN]       # int phi_N dphi_M/dy
            blk = -(dt * dt / 4.0) * (a[2] @ a[0] * Ix + a[2] @ a[1] * Iy)
            add[5 * N:5 * N + 5, 5 * Mi:5 * Mi + 5] += blk
    diff = R1[e] - R0[e]
    assert np.allclose(diff, add, rtol=1e-11, atol=1e-13 * np.abs(add).max())
    assert np.abs(rhs1[e]).max() < 1e-18


@pytest.mark.parametrize("seed", [1, 2])
def test_column_sum_identity_and_conservation_3d(seed):
    m, th, U = M.mixed_bc_box(4, 4, 4, seed=seed)
    dt = 2e-3
    R, rhs, _ = O.build(th, m, U, dt, apply_bc=False)
    _, area_cl = O.calc_length(m)
    area, _ = O.calc_length(m)
    interior = np.nonzero((m.bc == 0).all(axis=1))[0]
    assert len(interior) > 0
    for e in interior:
        colsum = R[e].reshape(8, 5, 40).sum(axis=0)          # [r][M*5+s]
        intphi = np.zeros(8)
        # int Phi_M by 2x2x2 Gauss (exact for trilinear)
        w, z = O.gauss2()
        xn = m.x[m.conn[e]]
        for zi in z:
            for zj in z:
                for zk in z:
                    phi = np.zeros(8)
                    d = np.zeros(24)
                    O.lib().orc_shape3d(O._p(np.array([zi, zj, zk])), O._p(phi), O._p(d))
                    J = O.lib().orc_jacobian3d(O._p(d), O._p(np.ascontiguousarray(xn)), O._p(np.zeros(9)),
                                               O._p(np.zeros(24)))
                    intphi += phi * J
        expect = np.kron(intphi[None, :], np.eye(5)).reshape(5, 40)
        scale = np.abs(R[e]).max()
        assert np.allclose(colsum, expect, atol=1e-12 * scale)
        assert abs(intphi.sum() - area[e]) < 1e-14 * area[e]
        rs = rhs[e].reshape(8, 5).sum(axis=0)
        assert np.all(np.abs(rs) <= 1e-12 * np.abs(rhs[e]).sum() + 1e-300)


def test_column_sum_identity_2d_with_numerical_diffusion():
    m, th, U = M.lid_cavity(6)
    # moving fluid everywhere so the numerical-diffusion term is active
    rng = np.random.default_rng(5)
    v = 0.3 + 0.1 * rng.standard_normal((m.n_node, 2))
    U = M.conservative(np.ones(m.n_node) + 0.05 * rng.standard_normal(m.n_node), v,
                       1 + 0.1 * rng.random(m.n_node), th.Cv)
    R, rhs, _ = O.build(th, m, U, 3e-3, apply_bc=False)
    area, _ = O.calc_length(m)
    interior = np.nonzero((m.bc == 0).all(axis=1))[0]
    for e in interior:
        colsum = R[e].reshape(4, 4, 16).sum(axis=0)
        # uniform square cells: int Phi_M = area/4
        expect = np.kron(np.full((1, 4), area[e] / 4), np.eye(4)).reshape(4, 16)
        assert np.allclose(colsum, expect, atol=1e-12 * np.abs(R[e]).max())
        rs = rhs[e].reshape(4, 4).sum(axis=0)
        assert np.all(np.abs(rs) <= 1e-12 * np.abs(rhs[e]).sum() + 1e-300)


@pytest.mark.parametrize("adiabatic", [False, True])
def test_apply_bc_rows(adiabatic):
    m, th, U = M.mixed_bc_box(4, 3, 3, adiabatic=adiabatic)
    R0, rhs0, _ = O.build(th, m, U, 1e-3, apply_bc=False)
    R1, rhs1, _ = O.build(th, m, U, 1e-3, apply_bc=True)
    fn = M.FACE_NODES_3D
    n_constrained = 0
    for e in range(m.n_elem):
        mask = np.zeros(40, bool)
        for f in range(6):
            bc = m.bc[e, f]
            if bc == 0:
                continue
            for ja in fn[f]:
                g = m.conn[e, ja]
                for q in range(5):
                    if m.dirichlet[g, q] and not (bc == 1 and q == 4 and not adiabatic):
                        mask[5 * ja + q] = True
        n_constrained += mask.sum()
        for c in range(40):
            if mask[c]:
                row = np.zeros(40)
                row[c] = 0.5
                assert np.array_equal(R1[e, c], row)
                assert rhs1[e, c] == 0
            else:
                assert np.array_equal(R1[e, c], R0[e, c])
                assert rhs1[e, c] == rhs0[e, c]
    assert n_constrained > 0


def test_rotation_covariance_interior_elements():
    """A cyclic permutation of the axes (a rotation, det +1) applied to node
    coordinates and velocities permutes the momentum equations and nothing
    else: R'(Nr', Ms') = R(Nr, Ms), rhs'(Nr') = rhs(Nr).  Any index slip in
    a_i, b_i, c_ij, G_i or in the geometric factors breaks it."""
    m, th, U = M.mixed_bc_box(4, 4, 4, seed=9)
    perm = [1, 2, 0]
    x2 = np.ascontiguousarray(m.x[:, perm])
    U2 = U.copy()
    U2[:, 1:4] = U[:, 1:4][:, perm]
    m2 = M.Mesh(3, 8, 5, m.conn, x2, np.zeros_like(m.bc), m.dirichlet)
    m1 = M.Mesh(3, 8, 5, m.conn, m.x, np.zeros_like(m.bc), m.dirichlet)
    R1, rhs1, s1 = O.build(th, m1, U, 2e-3, apply_bc=False)
    R2, rhs2, s2 = O.build(th, m2, U2, 2e-3, apply_bc=False)
    eq = np.array([0, 1 + perm[0], 1 + perm[1], 1 + perm[2], 4])
    idx = (np.arange(8)[:, None] * 5 + eq[None, :]).ravel()
    assert np.allclose(s1, s2, rtol=1e-13, atol=1e-15)
    for e in range(m.n_elem):
        expect = R1[e][np.ix_(idx, idx)]
        assert np.allclose(R2[e], expect, rtol=1e-11, atol=1e-13 * np.abs(expect).max())
        assert np.allclose(rhs2[e], rhs1[e][idx], rtol=1e-10,
                           atol=1e-12 * np.abs(rhs1[e]).max())


def _flow(rng, th):
    v = rng.standard_normal(3)
    dV = rng.standard_normal((3, 3))
    dT = rng.standard_normal(3)
    Uv = np.concatenate([[1.2], 1.2 * v, [3.0]])
    f = O.flow_struct(3.0, 1.1, v, np.zeros(3), Uv, rng.standard_normal((5, 3)), dV, dT,
                      0.7, 0.9)
    return f


def _textbook_tau(mu, dV):
    return mu * (dV + dV.T) - 2.0 / 3.0 * mu * np.trace(dV) * np.eye(3)


@pytest.mark.parametrize("adiabatic", [False, True])
def test_eval2dboundary_edits(adiabatic):
    """Face-state edits of eval2dboundary.h:14-371: which gradient components
    vanish, which stress rows/columns vanish and when the stress is the
    textbook deviatoric stress of the edited velocity gradient."""
    import ctypes as C
    th = M.Thermo(gamma=1.4, mach=2.0, reynolds=100.0, adiabatic=adiabatic)
    t = O.thermo(th)
    rng = np.random.default_rng(4)
    en = np.array([0.0, 0.0, -0.02])

    def run(bc):
        f = _flow(rng, th)
        dV0 = np.ctypeslib.as_array(f.dV).copy()
        dT0 = np.ctypeslib.as_array(f.dT).copy()
        rc = O.lib().orc_eval2dboundary(C.byref(t), bc, C.byref(f), O._p(en))
        return rc, f, dV0, dT0

    # symmetry planes: the normal derivative vanishes
    for bc, d in ((2, 2), (12, 2), (22, 1)):
        rc, f, dV0, dT0 = run(bc)
        dV = np.ctypeslib.as_array(f.dV)
        dT = np.ctypeslib.as_array(f.dT)
        assert rc == 0
        assert np.all(dV[:, d] == 0) and dT[d] == 0
        keep = [j for j in range(3) if j != d]
        assert np.array_equal(dV[:, keep], dV0[:, keep])
    # walls: tau is recomputed; the wall-normal dT vanishes only if adiabatic
    for bc, d in ((1, 2), (51, 2), (21, 1), (31, 1), (11, 0)):
        rc, f, dV0, dT0 = run(bc)
        tau = np.ctypeslib.as_array(f.tau)
        dT = np.ctypeslib.as_array(f.dT)
        assert np.allclose(tau, _textbook_tau(f.mu, dV0), rtol=1e-14, atol=1e-15)
        assert (dT[d] == 0) == adiabatic or (not adiabatic and dT[d] == dT0[d])
    # outflow/top: dT and stress column of the normal direction vanish.
    # 7 falls through into 77/32 (the source has no break): dT_z and dT_y
    # vanish, and the stress recomputed in 77 restores column z, so only
    # column y stays zero
    for bc, dts, cols in ((6, [0], [0]), (7, [2, 1], [1]), (77, [1], [1]), (32, [1], [1])):
        rc, f, dV0, dT0 = run(bc)
        tau = np.ctypeslib.as_array(f.tau)
        dT = np.ctypeslib.as_array(f.dT)
        assert all(dT[c] == 0 for c in dts)
        for c in cols:
            assert np.all(tau[:, c] == 0)
        rest = [j for j in range(3) if j not in cols]
        full = _textbook_tau(f.mu, dV0)
        assert np.allclose(tau[:, rest], full[:, rest], rtol=1e-14, atol=1e-15)
    # 8 (exit): stress recomputed, nothing else edited
    rc, f, dV0, dT0 = run(8)
    assert np.allclose(np.ctypeslib.as_array(f.tau), _textbook_tau(f.mu, dV0), rtol=1e-14,
                       atol=1e-15)
    assert np.array_equal(np.ctypeslib.as_array(f.dT), dT0)
    # -12 (lid): z-derivatives and w vanish, no stress is formed
    rc, f, dV0, dT0 = run(-12)
    assert np.all(np.ctypeslib.as_array(f.dV)[:, 2] == 0) and f.dT[2] == 0 and f.v[2] == 0
    assert np.all(np.ctypeslib.as_array(f.tau) == 0)
    # 4 (ramp wall): v = 0 and the gradients of u, v, T and of the first
    # three equations lose their component along en, which the source uses
    # unnormalised: with |en| = 1 the projection is exact, otherwise the
    # normal component is scaled by (1 - |en|^2)
    for scale in (1.0, 0.3):
        nrm = np.array([0.6, 0.0, 0.8]) * scale
        f = _flow(rng, th)
        dV0 = np.ctypeslib.as_array(f.dV).copy()
        dT0 = np.ctypeslib.as_array(f.dT).copy()
        dU0 = np.ctypeslib.as_array(f.dU).copy()
        assert O.lib().orc_eval2dboundary(C.byref(t), 4, C.byref(f), O._p(nrm)) == 0
        k = 1 - nrm @ nrm
        dV = np.ctypeslib.as_array(f.dV)
        dU = np.ctypeslib.as_array(f.dU)
        assert np.all(np.ctypeslib.as_array(f.v) == 0)
        assert np.allclose(dV[:2] @ nrm, k * (dV0[:2] @ nrm), atol=1e-14)
        assert np.array_equal(dV[2], dV0[2])
        assert np.allclose(np.ctypeslib.as_array(f.dT) @ nrm, k * (dT0 @ nrm), atol=1e-14)
        assert np.allclose(dU[:3] @ nrm, k * (dU0[:3] @ nrm), atol=1e-14)
        assert np.array_equal(dU[3:], dU0[3:])
    # inflow returns before any stress is formed
    rc, f, _, _ = run(-1)
    assert rc == 0 and np.all(np.ctypeslib.as_array(f.tau) == 0)
    # unknown code is an error
    rc, _, _, _ = run(99)
    assert rc == -1


def test_hnm_permutation_is_a_node_relabelling():
    """adap() (MeshRefine2D.h:877-900) maps Phi -> Phi Hnm.  With Hnm the
    permutation matrix of a cube rotation pi (Hnm(N, M) = 1 for M = pi(N)),
    an interior element behaves as the unadapted element whose reference
    corner k carries global node conn[pi(k)]: R_adapted(M, M') =
    R_relabelled(pi^-1(M), pi^-1(M')) block by block."""
    m, th, U = M.mixed_bc_box(4, 4, 4, seed=21)
    pi = np.array([1, 2, 3, 0, 5, 6, 7, 4])          # quarter turn about z
    inv = np.argsort(pi)
    interior = np.nonzero((m.bc == 0).all(axis=1))[0]
    adap = np.zeros(m.n_elem, np.uint8)
    adap[interior] = 1
    H = np.zeros((m.n_elem, 8, 8))
    H[:] = np.eye(8)
    for e in interior:
        H[e] = 0.0
        H[e, np.arange(8), pi] = 1.0                 # H[e, N, pi(N)] = 1
    ma = M.Mesh(3, 8, 5, m.conn, m.x, m.bc, m.dirichlet, adap=adap, hnm=H)
    conn_r = m.conn.copy()
    conn_r[interior] = m.conn[interior][:, pi]       # conn'[k] = conn[pi(k)]
    mr = M.Mesh(3, 8, 5, conn_r, m.x, m.bc, m.dirichlet)
    dt = 1e-3
    Ra, rhsa, _ = O.build(th, ma, U, dt, apply_bc=False)
    Rr, rhsr, _ = O.build(th, mr, U, dt, apply_bc=False)
    idx = (inv[:, None] * 5 + np.arange(5)).ravel()  # row M of Ra <- row pi^-1(M) of Rr
    for e in interior:
        expect = Rr[e][np.ix_(idx, idx)]
        assert np.allclose(Ra[e], expect, rtol=1e-12, atol=1e-14 * np.abs(expect).max())
        assert np.allclose(rhsa[e], rhsr[e][idx], rtol=1e-10, atol=1e-12 * np.abs(rhsr[e]).max())
    # identity Hnm changes nothing anywhere
    mi = M.Mesh(3, 8, 5, m.conn, m.x, m.bc, m.dirichlet, adap=np.ones(m.n_elem, np.uint8),
                hnm=np.tile(np.eye(8), (m.n_elem, 1, 1)))
    Ri, rhsi, _ = O.build(th, mi, U, dt)
    R0, rhs0, _ = O.build(th, m, U, dt)
    assert np.array_equal(Ri, R0) and np.array_equal(rhsi, rhs0)
