"""The wall functionals of the adjoint (cfdv_amd/functionals.py) on the CPU:
their gradients with respect to the state and the Mach number against
central differences, the face geometry against closed forms (the wall area
of an O-grid is the inscribed polygon's perimeter times the span; the
normals point into the cylinder), and a state at rest with uniform
pressure gives zero drag on a closed body."""
import numpy as np
import pytest

from cfdv_amd import mesh as M
from cfdv_amd.functionals import WallFunctional


@pytest.mark.parametrize("kind", ["drag", "heat"])
def test_gradient_against_central_differences(kind):
    mesh, th, U = M.cylinder_ogrid(24, 6, 2)
    rng = np.random.default_rng(3)
    U = U * (1 + 0.01 * rng.standard_normal(U.shape))
    F = WallFunctional(mesh, th, 1, kind)
    J, gU, gM = F.value_and_grad(U)
    assert J == pytest.approx(F.value(U), rel=1e-14)
    h = 1e-6
    for _ in range(3):
        d = rng.standard_normal(U.shape)
        fd = (F.value(U + h * d) - F.value(U - h * d)) / (2 * h)
        assert (gU * d).sum() == pytest.approx(fd, rel=1e-7)
    fdm = (F.value(U, th.mach + h) - F.value(U, th.mach - h)) / (2 * h)
    assert gM == pytest.approx(fdm, rel=1e-7)


def test_wall_geometry_closed_forms():
    nt, nz, r0, lz = 48, 3, 0.5, 1.0
    mesh, th, _ = M.cylinder_ogrid(nt, 4, nz, r0=r0, lz=lz)
    F = WallFunctional(mesh, th, 1, "drag")
    perimeter = nt * 2 * r0 * np.sin(np.pi / nt)
    assert F.dA.sum().item() == pytest.approx(perimeter * lz, rel=1e-13)
    # unit normals, pointing out of the fluid element = into the cylinder
    n = F.n.numpy()
    assert np.allclose(np.linalg.norm(n, axis=2), 1.0)
    X = mesh.x[mesh.conn[F.elems]]
    radial = X.mean(axis=1)[:, :2]
    assert (np.einsum("fgi,fi->fg", n[..., :2], radial) < 0).all()


def test_uniform_state_at_rest_has_no_drag():
    mesh, th, _ = M.cylinder_ogrid(32, 4, 2)
    n = mesh.n_node
    U = M.conservative(np.full(n, 1.2), np.zeros((n, 3)), np.full(n, 1.4), th.Cv)
    F = WallFunctional(mesh, th, 1, "drag")
    scale = (th.gamma - 1.0) * U[0, 4] * F.dA.sum().item()
    assert abs(F.value(U)) < 1e-14 * scale
