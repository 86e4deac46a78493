"""Pins of the CPU oracle's thermodynamics, basis functions and flux Jacobians.

Every check here is independent of the oracle's own formulas:
closed forms (Gauss points, shape functions, affine Jacobians, Thermo,
NodeUnpack), the Euler flux homogeneity A_i U = F_i(U), and central finite
differences of the textbook inviscid and viscous fluxes for a (dF/dU),
b (dG/dU) and c (dG/d(grad U)) with mu and k held fixed, as the reference's
Jacobians treat them (NavierStokes3D.h:74-477, NavierStokes2D.h:59-222).
"""
import numpy as np
import pytest

from cfdv_amd.mesh import Thermo, conservative, primitives
from oracle import oracle as O


def test_gauss_points_closed_form():
    w, z = O.gauss2()
    assert abs(z[1] - 1 / np.sqrt(3)) <= 2 * np.spacing(z[1])
    assert z[0] == -z[1]
    assert np.all(np.abs(w - 1.0) <= 4 * np.spacing(1.0))


@pytest.mark.parametrize("ndim", [2, 3])
def test_shape_partition_of_unity(ndim):
    rng = np.random.default_rng(0)
    L = O.lib()
    nn = 4 if ndim == 2 else 8
    for _ in range(20):
        zt = rng.uniform(-1, 1, ndim)
        phi = np.zeros(nn)
        d = np.zeros(nn * ndim)
        f = L.orc_shape2d if ndim == 2 else L.orc_shape3d
        f(O._p(zt), O._p(phi), O._p(d))
        assert abs(phi.sum() - 1) < 1e-15
        assert np.all(np.abs(d.reshape(nn, ndim).sum(axis=0)) < 1e-15)
        # nodal interpolation: corner N sits at the sign pattern of its shape function
    corners3 = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                         [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)
    corners = corners3[:nn, :ndim] if ndim == 3 else corners3[:4, :2]
    for n in range(nn):
        phi = np.zeros(nn)
        (L.orc_shape2d if ndim == 2 else L.orc_shape3d)(O._p(corners[n].copy()), O._p(phi),
                                                         O._p(np.zeros(nn * ndim)))
        assert np.allclose(phi, np.eye(nn)[n])


def test_affine_hex_jacobian_and_gradients():
    rng = np.random.default_rng(1)
    L = O.lib()
    corners = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                        [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)
    A = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
    c0 = rng.standard_normal(3)
    xn = corners @ A.T + c0                      # x = A xi + c0, affine
    zt = rng.uniform(-1, 1, 3)
    d = np.zeros(24)
    L.orc_shape3d(O._p(zt), O._p(np.zeros(8)), O._p(d))
    jac = np.zeros(9)
    dphi = np.zeros(24)
    J = L.orc_jacobian3d(O._p(d), O._p(np.ascontiguousarray(xn)), O._p(jac), O._p(dphi))
    assert J == pytest.approx(np.linalg.det(A), rel=1e-13)
    assert np.allclose(jac.reshape(3, 3), A, atol=1e-14)
    # dPhi/dx = dPhi/dxi A^{-1}; an affine field u = g.x is reproduced exactly
    g = rng.standard_normal(3)
    u = xn @ g
    assert np.allclose(dphi.reshape(8, 3).T @ u, g, atol=1e-13)


def test_affine_quad_jacobian():
    L = O.lib()
    xn = np.array([[0, 0], [2.0, 0.5], [2.5, 3.0], [0.5, 2.5]])
    # parallelogram: x = A xi + c
    A = np.array([[1.0, 0.25], [0.25, 1.25]])
    xn = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], float) @ A.T + np.array([1.0, 1.0])
    d = np.zeros(8)
    L.orc_shape2d(O._p(np.array([0.3, -0.2])), O._p(np.zeros(4)), O._p(d))
    dphi = np.zeros(8)
    J = L.orc_jacobian2d(O._p(d), O._p(np.ascontiguousarray(xn)), O._p(np.zeros(4)), O._p(dphi))
    assert J == pytest.approx(np.linalg.det(A), rel=1e-14)
    g = np.array([0.7, -1.3])
    assert np.allclose(dphi.reshape(4, 2).T @ (xn @ g), g, atol=1e-13)


def test_face_normal_axis_aligned():
    """Axis-aligned face: en = outward normal * w0 w1 * area/4, tangential
    gradients exact; the Lambda/Mu renormalisation quirk (calc_TF_norm3D.h:
    208-212) is inert here because Lambda and Mu are axis vectors."""
    L = O.lib()
    hx, hy = 0.3, 0.2
    # bottom face of a hex: local nodes (3, 2, 1, 0) -> x-range/y-range square
    nodeloc = np.array([[0, hy, 0], [hx, hy, 0], [hx, 0, 0], [0, 0, 0]], float)
    w = np.array([1.0, 1.0])
    zt = np.array([0.2, -0.4])
    phib = np.zeros(4)
    en = np.zeros(3)
    dp = np.zeros(12)
    det = L.orc_tfnorm3d(O._p(nodeloc), O._p(w), O._p(zt), O._p(phib), O._p(en), O._p(dp))
    assert det == pytest.approx(hx * hy / 4, rel=1e-14)
    assert np.allclose(en, [0, 0, -hx * hy / 4], atol=1e-16)
    assert abs(phib.sum() - 1) < 1e-15
    f = np.array([1.7, -0.6, 0.0])
    assert np.allclose(dp.reshape(4, 3).T @ (nodeloc @ f), f, atol=1e-13)


def test_face_normal_skewed_quirk_changes_normal_length():
    """On a face whose edge vectors are not axis-aligned, the reference's
    component-wise normalisation leaves Lambda non-unit; the oracle keeps
    that behaviour (its en differs from the unit-normal construction)."""
    L = O.lib()
    nodeloc = np.array([[0, 0, 0], [0.3, 0.1, 0], [0.35, 0.4, 0], [0.05, 0.3, 0]], float)
    phib = np.zeros(4)
    en = np.zeros(3)
    dp = np.zeros(12)
    L.orc_tfnorm3d(O._p(nodeloc), O._p(np.ones(2)), O._p(np.zeros(2)), O._p(phib), O._p(en),
                   O._p(dp))
    lam = nodeloc[1] - nodeloc[0]
    lq = lam.copy()
    for j in range(3):
        lq[j] /= np.linalg.norm(lq)
    assert abs(np.linalg.norm(lq) - 1) > 1e-3          # the quirk is active here
    assert abs(en[0]) < 1e-16 and abs(en[1]) < 1e-16    # still along z


def test_thermo_closed_forms():
    th = Thermo(gamma=1.4, mach=3.0, reynolds=1e5, prandtl=0.72, tinf=216.6666, twall=2.0,
                adiabatic=False)
    t = O.thermo(th)
    assert t.Cv == pytest.approx(1 / (1.4 * 0.4 * 9), rel=1e-15)
    assert t.cgas == pytest.approx(1 / (1.4 * 9), rel=1e-15)
    assert t.csuth == pytest.approx(110 / 216.6666, rel=1e-15)
    L = O.lib()
    assert L.orc_calc_mu(O.thermo(th), 1.0) == pytest.approx(1.0, rel=1e-15)   # mu(T_inf) = 1
    assert L.orc_calc_mu(O.thermo(th), 1e-13) == 1.0
    T = 2.5
    cs = 110 / 216.6666
    assert L.orc_calc_mu(O.thermo(th), T) == pytest.approx((1 + cs) / (T + cs) * T ** 1.5, rel=1e-14)
    mu = 1.7
    assert L.orc_calc_k(O.thermo(th), mu) == pytest.approx(mu / (0.4 * 9 * 0.72), rel=1e-15)


@pytest.mark.parametrize("ndim", [2, 3])
def test_node_unpack_round_trip(ndim):
    rng = np.random.default_rng(ndim)
    th = Thermo(mach=2.0)
    n = 50
    rho = rng.uniform(0.5, 2, n)
    v = rng.uniform(-1, 1, (n, ndim))
    T = rng.uniform(0.5, 3, n)
    U = conservative(rho, v, T, th.Cv)
    from cfdv_amd.mesh import Mesh
    m = Mesh(ndim, 8 if ndim == 3 else 4, ndim + 2, np.zeros((1, 8 if ndim == 3 else 4), np.int32),
             np.zeros((n, ndim)), np.zeros((1, 6 if ndim == 3 else 4), np.int32),
             np.zeros((n, ndim + 2), np.uint8))
    p = O.prim_arrays(O.unpack(th, m, U), ndim)
    assert np.allclose(p["rho"], rho, rtol=1e-15)
    assert np.allclose(p["v"], v, rtol=1e-14, atol=1e-15)
    assert np.allclose(p["T"], T, rtol=1e-13)
    assert np.allclose(p["p"], rho * (th.gamma - 1) * th.Cv * T, rtol=1e-13)
    q = primitives(U, ndim, th.Cv, th.gamma)
    assert np.allclose(q["T"], p["T"], rtol=1e-15)


def _state(ndim, rng, th):
    neq = ndim + 2
    rho = rng.uniform(0.6, 1.6)
    v = rng.uniform(-0.8, 0.9, ndim)
    T = rng.uniform(0.7, 2.0)
    U = np.concatenate([[rho], rho * v, [rho * (th.Cv * T + 0.5 * v @ v)]])
    dU = 0.3 * rng.standard_normal((neq, ndim))
    return U, v, T, dU


def _inviscid_flux(ndim, th, U):
    rho = U[0]
    v = U[1:1 + ndim] / rho
    E = U[ndim + 1] / rho
    p = rho * (th.gamma - 1) * (E - 0.5 * v @ v)
    F = np.zeros((ndim, ndim + 2))
    for i in range(ndim):
        F[i, 0] = rho * v[i]
        F[i, 1:1 + ndim] = rho * v[i] * v
        F[i, 1 + i] += p
        F[i, ndim + 1] = (rho * E + p) * v[i]
    return F


def _viscous_flux(ndim, th, U, dU, mu, k):
    rho = U[0]
    v = U[1:1 + ndim] / rho
    dv = np.array([[(dU[1 + i, j] - v[i] * dU[0, j]) / rho for j in range(ndim)] for i in range(ndim)])
    E = U[ndim + 1] / rho
    dE = np.array([(dU[ndim + 1, j] - E * dU[0, j]) / rho for j in range(ndim)])
    dT = np.array([(dE[j] - v @ dv[:, j]) / th.Cv for j in range(ndim)])
    tau = mu * (dv + dv.T) - (2.0 / 3.0) * mu * np.trace(dv) * np.eye(ndim)
    G = np.zeros((ndim, ndim + 2))
    for i in range(ndim):
        G[i, 1:1 + ndim] = -tau[i] / th.reynolds
        G[i, ndim + 1] = (-(tau[i] @ v) - k * dT[i]) / th.reynolds
    return G, tau, dT, dv


@pytest.mark.parametrize("ndim", [2, 3])
def test_flux_jacobians_against_finite_differences(ndim):
    rng = np.random.default_rng(10 + ndim)
    th = Thermo(gamma=1.4, mach=2.0, reynolds=100.0, prandtl=0.72, tinf=216.6666)
    L = O.lib()
    neq = ndim + 2
    for trial in range(4):
        U, v, T, dU = _state(ndim, rng, th)
        mu = L.orc_calc_mu(O.thermo(th), T)
        k = L.orc_calc_k(O.thermo(th), mu)
        G, tau, dT, dv = _viscous_flux(ndim, th, U, dU, mu, k)
        f = O.flow_struct(U[ndim + 1] / U[0], T, v, np.zeros(3), U, dU, dv, dT, k, mu, tau)
        a, b, c = O.calc_abc(th, ndim, f)
        # Euler homogeneity and FD of the inviscid flux
        F = _inviscid_flux(ndim, th, U)
        for i in range(ndim):
            assert np.allclose(a[i] @ U, F[i], rtol=1e-13, atol=1e-13)
        h = 1e-6
        for s in range(neq):
            e = np.zeros(neq)
            e[s] = h
            dF = (_inviscid_flux(ndim, th, U + e) - _inviscid_flux(ndim, th, U - e)) / (2 * h)
            assert np.allclose(a[:, :, s], dF, rtol=1e-7, atol=1e-8)
            Gp = _viscous_flux(ndim, th, U + e, dU, mu, k)[0]
            Gm = _viscous_flux(ndim, th, U - e, dU, mu, k)[0]
            assert np.allclose(b[:, :, s], (Gp - Gm) / (2 * h), rtol=1e-6, atol=1e-10)
            for j in range(ndim):
                E2 = np.zeros((neq, ndim))
                E2[s, j] = h
                Gp = _viscous_flux(ndim, th, U, dU + E2, mu, k)[0]
                Gm = _viscous_flux(ndim, th, U, dU - E2, mu, k)[0]
                assert np.allclose(c[:, j, :, s], (Gp - Gm) / (2 * h), rtol=1e-6, atol=1e-10)


@pytest.mark.parametrize("ndim", [2, 3])
def test_tau_symmetry_and_trace(ndim):
    rng = np.random.default_rng(3)
    dV = rng.standard_normal((ndim, ndim))
    mu = 1.3
    t = O.tau(ndim, mu, dV)
    assert np.allclose(t, t.T, atol=1e-15)
    # trace: 2 mu div - ndim * (2/3) mu div
    assert np.trace(t) == pytest.approx((2 - ndim * 2.0 / 3.0) * mu * np.trace(dV), abs=1e-14)


def test_viscous_flux_closed_form():
    rng = np.random.default_rng(4)
    th = Thermo(reynolds=250.0)
    tau = np.zeros((3, 3))
    tau[:, :] = rng.standard_normal((3, 3))
    v = rng.standard_normal(3)
    dT = rng.standard_normal(3)
    Gout = np.zeros(15)
    O.lib().orc_calc_G(3, O._p(np.ascontiguousarray(tau)), O._p(v), O._p(dT), 0.37, th.reynolds,
                       O._p(Gout))
    G = Gout.reshape(3, 5)
    for i in range(3):
        assert G[i, 0] == 0
        assert np.allclose(G[i, 1:4], -tau[i] / 250.0, rtol=1e-15)
        assert G[i, 4] == pytest.approx((-(tau[i] @ v) - 0.37 * dT[i]) / 250.0, rel=1e-14)
