"""Node rows built in place (CFDV_BLOCKS_ROWS alone, no element blocks).

The build adds each element's blocks into a staged copy of the FP64 rows,
one color after the other, and lays the FP64 and FP32 rows out from it (the
rows of K^T through the mirrored slots).  The operator is the one a context
with FP64 element blocks assembles from them, summed in another order, so
the two agree to rounding:

* K x, K^T x over the FP64 rows to 1e-13 of |K| |x| (and again after a
  switch of orientation each way), the FP32 rows' K x to FP32 rounding;
* the right-hand side bit for bit (the same element arithmetic);
* FP64 and mixed solves to 1e-10 of the increment, both at rel <= 1e-8;
* two builds of the same state give the same rows bit for bit (the sum
  order is fixed by the colors);
* calls that need element blocks (the reference solve, the block dump) are
  refused.

Meshes cover every boundary kind the face kernel sees (mixed-bc box with
constrained dofs), hanging-node elements with and without faces (a refined
ramp), the O-grid and the flat plate.
"""
import numpy as np
import pytest

from cfdv_amd import mesh as M
from cfdv_amd.fdv import CFDVError, FDVContext
from cfdv_amd.refine import sensor_refine

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "box3d":
        return M.mixed_bc_box(4, 3, 3, seed=5)
    if name == "ramp_refined":
        m, th, U = M.compression_ramp(48, 24, 12, noise=0.02)
        m, U, info = sensor_refine(m, th, U, 1)
        assert m.adap is not None and np.asarray(m.adap).any()
        return m, th, U
    if name == "cylinder":
        return M.cylinder_ogrid(12, 4, 2)
    if name == "plate":
        return M.flat_plate(16, 8, 6)
    raise KeyError(name)


def _scale(ctx, x, tr=False):
    # |K| |x| entry by entry bounds the rounding of any summation order
    return np.abs(ctx.matvec(np.abs(x), transpose=tr)).max() + 1e-300


@pytest.mark.parametrize("case", ["box3d", "ramp_refined", "cylinder", "plate"])
def test_rows_in_place_equal_assembled_rows(gpu_available, case):
    m, th, U = _case(case)
    ref = FDVContext(m, th, U, blocks=("fp64", "rows"))
    drc = FDVContext(m, th, U, blocks=("rows",))
    dt = ref.cfl(0.5)
    assert drc.cfl(0.5) == dt
    ref.build(dt)
    drc.build(dt)
    assert np.array_equal(ref.rhs_vector(), drc.rhs_vector())
    rng = np.random.default_rng(11)
    x = rng.standard_normal(ref.ndof)
    for tr in (False, True, False, True):
        a, b = ref.matvec(x, transpose=tr), drc.matvec(x, transpose=tr)
        err = np.abs(a - b).max() / _scale(ref, x, tr)
        print(f"{case}: K{'^T' if tr else ''} x max diff / |K||x| = {err:.1e}")
        assert err <= 1e-13
    a32, b32 = ref.matvec(x, precision=32), drc.matvec(x, precision=32)
    err32 = np.abs(a32 - b32).max() / _scale(ref, x)
    print(f"{case}: FP32 rows K x max diff / |K||x| = {err32:.1e}")
    assert err32 <= 1e-6
    # the same state built again: the same rows, bit for bit
    y0 = drc.matvec(x)
    drc.build(dt)
    assert np.array_equal(drc.matvec(x), y0)
    for mode in ("fp64", "mixed"):
        s_ref, _ = ref.solve(mode=mode, restarts=50, m=30, rtol=1e-8, precond="nodeblock")
        d_ref = ref.du()
        s_drc, _ = drc.solve(mode=mode, restarts=50, m=30, rtol=1e-8, precond="nodeblock")
        d_drc = drc.du()
        rel = np.linalg.norm(d_drc - d_ref) / np.linalg.norm(d_ref)
        print(f"{case}/{mode}: rel resid {s_ref.rel_resid:.1e} / {s_drc.rel_resid:.1e}, du rel diff {rel:.1e}")
        assert s_drc.rel_resid <= 1e-8
        assert rel <= 1e-7


def test_rows_in_place_adjoint_solve(gpu_available):
    m, th, U = _case("box3d")
    ref = FDVContext(m, th, U, blocks=("fp64", "rows"))
    drc = FDVContext(m, th, U, blocks=("rows",))
    dt = ref.cfl(0.5)
    ref.build(dt)
    drc.build(dt)
    g = np.random.default_rng(3).standard_normal(ref.ndof)
    lr, _ = ref.adjoint_solve(g, mode="fp64", rtol=1e-12)
    ld, _ = drc.adjoint_solve(g, mode="fp64", rtol=1e-12)
    rel = np.linalg.norm(ld - lr) / np.linalg.norm(lr)
    print(f"adjoint solve: rel diff {rel:.1e}")
    assert rel <= 1e-9


def test_rows_in_place_refuse_element_block_calls(gpu_available):
    m, th, U = _case("box3d")
    drc = FDVContext(m, th, U, blocks=("rows",))
    drc.build(drc.cfl(0.5))
    with pytest.raises(CFDVError):
        drc.blocks()
    with pytest.raises(CFDVError):
        drc.solve(mode="reference", restarts=1, m=20)
    m2, th2, U2 = M.lid_cavity(4)
    with pytest.raises(CFDVError):
        FDVContext(m2, th2, U2, blocks=("rows",))
