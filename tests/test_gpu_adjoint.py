"""Discrete adjoint of one FDV step (hex8) against finite differences.

1. Build tangents: the gradient of Phi(U, M) = lam^T (b - K du) (du the
   step's update, lam random) from cfdv_build_vjp against central
   differences of Phi evaluated from the GPU build and, independently, from
   the CPU oracle's build.
2. Step gradient: J = g^T U1 for one step at a fixed dt; the adjoint
   gradient (transpose GMRES + build tangents) against central differences
   of full GPU steps along 16 seeded directions (random fields, wall-node
   temperatures, the uniform wall temperature, the Mach number); relative
   error <= 1e-5.
3. The torch binding: backward of fdv_step equals the adjoint gradient.

FDVParam's max/min over nodes and its clip at r = 1 make Phi piecewise
smooth: a kink within h of the state makes a central difference average two
slopes (seen once at h = 1e-6 along a random direction on the mixed-bc box,
where h = 1e-7 then agrees with the tangent to 1e-7).  Each check therefore
takes central differences at h and h/10 and requires the closer one to meet
the tolerance; a wrong tangent misses at both.
"""
import dataclasses

import numpy as np
import pytest

from cfdv_amd import mesh as M
from cfdv_amd.adjoint import multi_step_gradient, step_fixed_dt, step_gradient, wall_temperature_gradient
from cfdv_amd.fdv import FDVContext
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "box3d":
        return M.mixed_bc_box(4, 3, 3, seed=5)
    if name == "ramp":
        return M.compression_ramp(6, 4, 3)
    if name == "plate":
        return M.flat_plate(6, 4, 3)
    if name == "cylinder":
        return M.cylinder_ogrid(12, 4, 2)
    raise KeyError(name)


def _phi_gpu(ctx, U, dt, lam, du):
    ctx.set_state(U)
    ctx.build(dt)
    b = ctx.rhs_vector()
    return lam @ b - lam @ ctx.matvec(du)


def _phi_oracle(mesh, th, U, dt, lam, du):
    R, rhs, _ = O.build(th, mesh, U, dt)
    b = O.compose_rhs(mesh, rhs)
    return lam @ b - lam @ O.matvec(mesh, R, du)


def _fd_best(f, an, h=1e-6):
    """relative error of the tangent ``an`` against central differences of
    f at h and h/10 (the closer of the two) and the two differences"""
    fds = [(f(hh) - f(-hh)) / (2 * hh) for hh in (h, h / 10)]
    errs = [abs(an - fd) / max(abs(fd), 1e-300) for fd in fds]
    return min(errs), fds


def _wall_nodes(mesh):
    fn = M.FACE_NODES_3D
    nodes = set()
    for e in range(mesh.n_elem):
        for f in range(6):
            if mesh.bc[e, f] in (1, 4, 11, 21, 31, 41):
                nodes.update(mesh.conn[e, fn[f]].tolist())
    return np.array(sorted(nodes), dtype=np.int64)


@pytest.mark.parametrize("case", ["box3d", "ramp", "plate"])
def test_build_vjp_against_finite_differences(gpu_available, case):
    mesh, th, U = _case(case)
    ctx = FDVContext(mesh, th, U)
    dt = ctx.cfl(0.5)
    rng = np.random.default_rng(3)
    lam = rng.standard_normal(ctx.ndof)
    # du = K^-1 b, so that Phi(U) itself is ~0 and central differences are
    # not swamped by the rounding of a large Phi
    ctx.build(dt)
    ctx.solve("fp64", restarts=200, m=30, rtol=1e-13, precond="nodeblock")
    du = ctx.du()
    gU, gM = ctx.build_vjp(lam, du)
    scale = np.abs(U).max(axis=0)
    for k in range(4):
        d = rng.standard_normal(U.shape) * scale
        an = np.sum(gU * d)
        eg, fdg = _fd_best(lambda h: _phi_gpu(ctx, U + h * d, dt, lam, du), an)
        eo, fdo = _fd_best(lambda h: _phi_oracle(mesh, th, U + h * d, dt, lam, du), an)
        print(case, "dir", k, "adjoint", an, "fd(gpu)", fdg, "fd(oracle)", fdo, "rel err", eg, eo)
        assert eg <= 1e-6 and eo <= 1e-6

    def phiM(h):
        ctx.set_thermo(dataclasses.replace(th, mach=th.mach * (1 + h)))
        return _phi_gpu(ctx, U, dt, lam, du)

    eM, fdM = _fd_best(phiM, gM * th.mach)
    ctx.set_thermo(th)
    print(case, "dPhi/dM adjoint", gM, "fd", [v / th.mach for v in fdM], "rel err", eM)
    assert eM <= 1e-6


@pytest.mark.parametrize("case", ["box3d", "ramp", "cylinder"])
def test_step_gradient_16_directions(gpu_available, case):
    mesh, th, U0 = _case(case)
    ctx = FDVContext(mesh, th, U0)
    dt = ctx.cfl(0.5)
    rng = np.random.default_rng(11)
    g = rng.standard_normal(ctx.ndof)
    gU0, gM, _ = step_gradient(ctx, U0, dt, g, rtol=1e-13)

    # J = g^T U0 + g^T du: the first term is linear in U0 with gradient g, so
    # the differences are taken of g^T du alone (no cancellation against the
    # large g^T U0) and compared with dJ/dU0 - g
    def Jdu(U, mach=th.mach):
        ctx.set_thermo(dataclasses.replace(th, mach=mach))
        _, du = step_fixed_dt(ctx, U, dt, rtol=1e-13)
        return g @ du

    walls = _wall_nodes(mesh)
    scale = np.abs(U0).max(axis=0)
    dirs = []
    for k in range(10):                          # random fields
        dirs.append(("field", rng.standard_normal(U0.shape) * scale))
    for k in range(4):                           # wall-node temperatures (energy of wall nodes)
        d = np.zeros_like(U0)
        if len(walls):
            d[walls, 4] = rng.standard_normal(len(walls)) * scale[4]
        else:
            d = rng.standard_normal(U0.shape) * scale
        dirs.append(("wallT", d))
    d = np.zeros_like(U0)                        # uniform wall temperature
    if len(walls):
        d[walls, 4] = U0[walls, 0] * th.Cv
    else:
        d[:, 4] = U0[:, 0] * th.Cv
    dirs.append(("wallT_uniform", d))
    worst = 0.0
    gdu = gU0 - g.reshape(gU0.shape)
    for name, d in dirs:
        an = np.sum(gdu * d)
        err, fds = _fd_best(lambda h: Jdu(U0 + h * d), an)
        worst = max(worst, err)
        print(case, name, "adjoint", an, "fd", fds, "rel err", err)
        assert err <= 1e-5
    errM, fdM = _fd_best(lambda h: Jdu(U0, th.mach * (1 + h)), gM * th.mach)
    ctx.set_thermo(th)
    print(case, "Mach adjoint", gM, "fd", [v / th.mach for v in fdM], "rel err", errM, "worst field", worst)
    assert errM <= 1e-5
    if len(walls):
        gT = wall_temperature_gradient(gU0, U0, th.Cv, walls)
        terms = gU0 * dirs[14][1]
        assert abs(gT.sum() - terms.sum()) <= 1e-12 * np.abs(terms).sum()


def test_torch_fdv_step_backward(gpu_available):
    torch = pytest.importorskip("torch")
    from cfdv_amd.adjoint import fdv_step
    mesh, th, U0 = _case("box3d")
    ctx = FDVContext(mesh, th, U0)
    dt = ctx.cfl(0.5)
    rng = np.random.default_rng(2)
    g = rng.standard_normal(U0.shape)
    Ut = torch.tensor(U0, dtype=torch.float64, requires_grad=True)
    Mt = torch.tensor(th.mach, dtype=torch.float64, requires_grad=True)
    U1 = fdv_step(Ut, Mt, ctx, dt)
    (U1 * torch.from_numpy(g)).sum().backward()
    gU0, gM, U1n = step_gradient(ctx, U0, dt, g.ravel())
    assert np.allclose(U1.detach().numpy(), U1n, rtol=1e-12, atol=1e-14)
    assert np.allclose(Ut.grad.numpy(), gU0, rtol=1e-10, atol=1e-12)
    assert abs(Mt.grad.item() - gM) <= 1e-10 * max(abs(gM), 1.0)


def test_torch_gradcheck(gpu_available):
    """torch.autograd.gradcheck of fdv_step in float64 on a tiny mesh: the
    full Jacobian of U1 with respect to U0 and the Mach number from the
    adjoint (one backward per output) against central differences of GPU
    steps (one pair per input)."""
    torch = pytest.importorskip("torch")
    from cfdv_amd.adjoint import fdv_step
    mesh, th, U0 = M.mixed_bc_box(2, 2, 2, seed=3)
    ctx = FDVContext(mesh, th, U0)
    dt = ctx.cfl(0.5)
    Ut = torch.tensor(U0, dtype=torch.float64, requires_grad=True)
    Mt = torch.tensor(th.mach, dtype=torch.float64, requires_grad=True)
    ok = torch.autograd.gradcheck(lambda U, m: fdv_step(U, m, ctx, dt), (Ut, Mt), eps=1e-6, atol=1e-7,
                                  rtol=1e-5, fast_mode=False)
    ctx.set_thermo(th)
    assert ok


def test_build_vjp_on_hanging_node_elements(gpu_available):
    rng = np.random.default_rng(8)
    m, th, U = M.mixed_bc_box(4, 3, 3, seed=6)
    adap = (rng.random(m.n_elem) < 0.5).astype(np.uint8)
    A = rng.standard_normal((m.n_elem, 8, 8))
    H = np.tile(np.eye(8), (m.n_elem, 1, 1)) + 0.02 * (A - A.mean(axis=2, keepdims=True))
    mesh = M.Mesh(3, 8, 5, m.conn, m.x, m.bc, m.dirichlet, adap=adap, hnm=H)
    ctx = FDVContext(mesh, th, U)
    dt = ctx.cfl(0.5)
    ctx.build(dt)
    ctx.solve("fp64", restarts=200, m=30, rtol=1e-13, precond="nodeblock")
    du = ctx.du()
    lam = rng.standard_normal(ctx.ndof)
    gU, gM = ctx.build_vjp(lam, du)
    scale = np.abs(U).max(axis=0)
    for k in range(3):
        d = rng.standard_normal(U.shape) * scale
        an = np.sum(gU * d)
        eo, fdo = _fd_best(lambda h: _phi_oracle(mesh, th, U + h * d, dt, lam, du), an)
        print("adapted dir", k, "adjoint", an, "fd(oracle)", fdo, "rel err", eo)
        assert eo <= 1e-6


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_step_gradient_on_node_rows(gpu_available, mode):
    """The adjoint over the assembled rows (K^T lambda = g on the rows
    re-summed as those of K^T, the forward step on K's rows) gives the
    element-by-element adjoint's gradient, to the solver tolerance."""
    mesh, th, U0 = _case("ramp")
    ebe = FDVContext(mesh, th, U0)
    rows = FDVContext(mesh, th, U0, blocks=("fp64", "rows"))
    dt = ebe.cfl(0.5)
    g = np.random.default_rng(5).standard_normal(ebe.ndof)
    gU_e, gM_e, U1_e = step_gradient(ebe, U0, dt, g, rtol=1e-12)
    gU_r, gM_r, U1_r = step_gradient(rows, U0, dt, g, rtol=1e-12, mode=mode)
    eU = np.linalg.norm(gU_r - gU_e) / np.linalg.norm(gU_e)
    eM = abs(gM_r - gM_e) / abs(gM_e)
    e1 = np.linalg.norm(U1_r - U1_e) / np.linalg.norm(U1_e)
    print(mode, "rows vs EBE: dJ/dU0 %.2e dJ/dM %.2e U1 %.2e" % (eU, eM, e1))
    assert eU <= 1e-7 and eM <= 1e-7 and e1 <= 1e-10


def test_multi_step_gradient(gpu_available):
    """The checkpointed adjoint of three steps (dJ/dU0 and dJ/dM of
    J = g^T U3 at fixed dt's) against central differences of the three
    forward steps on 6 random directions and the Mach number.  The case is
    the perturbed box, whose FDVParam ratios r stay clear of the 0.01
    switch (FDVParam.h:15-111) over the three steps; from a uniform
    freestream the developing flow brings some element's r to the switch,
    where J jumps and no difference quotient converges."""
    mesh, th, U0 = _case("box3d")
    ctx = FDVContext(mesh, th, U0)
    dt0 = ctx.cfl(0.5)
    dts = [dt0, 0.9 * dt0, 1.1 * dt0]
    g = np.random.default_rng(17).standard_normal(ctx.ndof)
    gU0, gM, U3 = multi_step_gradient(ctx, U0, dts, g, rtol=1e-13)

    def J(U, mach=th.mach):
        ctx.set_thermo(dataclasses.replace(th, mach=mach))
        Uk = np.asarray(U, dtype=np.float64)
        for dt in dts:
            Uk, _ = step_fixed_dt(ctx, Uk, dt, rtol=1e-13)
        return g @ Uk

    # J = g^T U0 + g^T (U3 - U0): differences of the second term only
    def Jd(U, mach=th.mach):
        return J(U, mach) - g @ np.asarray(U).ravel()

    rng = np.random.default_rng(3)
    scale = np.abs(U0).max(axis=0)
    gd = gU0 - g.reshape(gU0.shape)
    for k in range(6):
        d = rng.standard_normal(U0.shape) * scale
        an = np.sum(gd * d)
        err, fds = _fd_best(lambda h: Jd(U0 + h * d), an)
        print("direction", k, "adjoint", an, "fd", fds, "rel err", err)
        assert err <= 1e-5
    errM, fdM = _fd_best(lambda h: Jd(U0, th.mach * (1 + h)), gM * th.mach)
    ctx.set_thermo(th)
    print("Mach adjoint", gM, "fd", [v / th.mach for v in fdM], "rel err", errM)
    assert errM <= 1e-5
