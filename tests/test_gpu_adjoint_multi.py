"""Eight adjoint right-hand sides at once (cfdv_adjoint_solve_multi).

K^T Lambda_s = G_s for 8 right-hand sides: FP64 refinement per right-hand
side around an FP32 global GMRES whose Arnoldi steps are 8-vector products
K^T X (node rows, or element by element on the FP32 matrix cores).  Each
lambda_s is checked against

* its own FP64 residual, recomputed with the single FP64 transposed product:
  ||g_s - K^T lambda_s|| <= rtol ||g_s||;
* the single-vector FP64 adjoint solve of the same g_s (cfdv_adjoint_solve),
  to the accuracy the conditioning of K^T allows (both solves stop at
  rtol = 1e-12);
* and, for the multi-step gradient of 8 wall functionals, the gradient of
  each functional from the single-functional FP64 adjoint.
A zero right-hand side gives lambda = 0 and a repeated one the same lambda.
"""
import numpy as np
import pytest

from cfdv_amd import mesh as M
from cfdv_amd.adjoint import multi_step_gradient, multi_step_gradient_multi
from cfdv_amd.fdv import FDVContext
from cfdv_amd.functionals import WallFunctional

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "box3d":
        return M.mixed_bc_box(4, 3, 3, seed=5)
    if name == "cylinder":
        return M.cylinder_ogrid(12, 4, 2)
    if name == "plate":
        return M.flat_plate(16, 8, 6)
    raise KeyError(name)


@pytest.mark.parametrize("case,storage", [("box3d", "rows"), ("cylinder", "rows"), ("plate", "rows"),
                                          ("box3d", "ebe"), ("plate", "ebe")])
def test_eight_adjoints_match_single_solves(gpu_available, case, storage):
    m, th, U = _case(case)
    blocks = ("fp64", "rows") if storage == "rows" else ("fp64", "fp32")
    ctx = FDVContext(m, th, U, blocks=blocks)
    ctx.build(ctx.cfl(0.5))
    rng = np.random.default_rng(31)
    G = rng.standard_normal((8, ctx.ndof))
    G[3] = 0.0          # a zero right-hand side
    G[6] = G[1]         # a repeated one
    G[7] *= 1e-6        # a small one: converged relative to its own norm
    lam, rel, st = ctx.adjoint_solve_multi(G, rtol=1e-12, inner_m=30, inner_rtol=1e-4)
    print(f"{case}/{storage}: {st.iters} global Arnoldi steps ({st.matvecs32} 8-vector products), "
          f"{st.matvecs64} FP64 products, {st.restarts} sweeps, rel {np.array2string(rel, precision=1)}")
    assert np.all(rel <= 1e-12)
    assert np.all(lam[3] == 0.0)
    for s in range(8):
        if s == 3:
            continue
        r = G[s] - ctx.matvec(lam[s], transpose=True)
        assert np.linalg.norm(r) <= 1.5e-12 * np.linalg.norm(G[s]), s
        one, st1 = ctx.adjoint_solve(G[s], mode="fp64", rtol=1e-12)
        d = np.linalg.norm(lam[s] - one) / np.linalg.norm(one)
        print(f"  rhs {s}: |lambda - single FP64 solve| / |.| = {d:.1e}")
        assert d <= 1e-9, s
    assert np.linalg.norm(lam[6] - lam[1]) <= 1e-9 * np.linalg.norm(lam[1])


def test_eight_functional_gradients(gpu_available):
    """Gradients of 8 functionals -- the wall heat flux and the drag on four
    streamwise patches of the wall -- after 3 steps, through one
    8-right-hand-side adjoint per step, equal those of the single-functional
    FP64 adjoint."""
    m, th, U0 = M.flat_plate(12, 6, 4)
    ctx = FDVContext(m, th, U0, blocks=("fp64", "rows"))
    dts = [ctx.cfl(0.5)] * 3
    e, _ = np.nonzero(np.asarray(m.bc) == 1)
    xc = np.asarray(m.x)[np.asarray(m.conn)[e], 0].mean(axis=1)
    edges = np.linspace(xc.min(), xc.max() + 1e-9, 5)
    funcs = [WallFunctional(m, th, 1, kind, x_range=(edges[p], edges[p + 1]))
             for p in range(4) for kind in ("heat", "drag")]
    gUs, gMs, U_N, Js = multi_step_gradient_multi(ctx, U0, dts, funcs, rtol=1e-12)
    for s in (0, 1, 5):
        gU, gM, _, J = multi_step_gradient(ctx, U0, dts, functional=funcs[s], mode="fp64", rtol=1e-12)
        eu = np.linalg.norm(gUs[s] - gU) / np.linalg.norm(gU)
        em = abs(gMs[s] - gM) / max(abs(gM), 1e-300)
        print(f"functional {s} ({funcs[s].kind}): J {Js[s]:.6e} vs {J:.6e}; dJ/dU0 rel diff {eu:.1e}, "
              f"dJ/dM rel diff {em:.1e}")
        assert abs(Js[s] - J) <= 1e-12 * abs(J)
        assert eu <= 1e-8 and em <= 1e-8
