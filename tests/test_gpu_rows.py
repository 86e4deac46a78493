"""Assembled node rows (CFDV_BLOCKS_ROWS) against the oracle's EBE product.

The rows are the element blocks summed per node pair.  The reference has no
such storage: it multiplies element by element (MPI_GMRES.h:442-566), so
K x from the rows equals the oracle's EBE product up to summation order, to
1e-13 relative in FP64.  The FP32 rows are the FP64 sums rounded once; K32 x
agrees to 2e-6 relative (FP32 inputs and accumulation).  Meshes: hex8 box
with mixed bc, flat plate, quad4 lid, a box with shuffled node numbers
(ragged chunks, neighbour slots that differ lane by lane) and a refined box
with hanging-node elements.  Solves on the rows (FP64, node-block
preconditioned FP64 and mixed) reach the dense solution to 1e-8, the
reference solve still multiplies element by element (bit-identical to a
context without rows), and K^T x stays on the element blocks.
"""
import numpy as np
import pytest

from cfdv_amd import mesh as M
from cfdv_amd import refine as PR
from cfdv_amd.fdv import CFDVError, FDVContext
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _shuffled_box(seed=9):
    m, th, U = M.mixed_bc_box(5, 4, 3, seed=2)
    rng = np.random.default_rng(seed)
    p = rng.permutation(m.n_node)            # old node -> new node
    inv = np.argsort(p)
    conn = p[m.conn]
    x = m.x[inv]
    d = m.dirichlet[inv]
    return M.Mesh(3, 8, 5, conn, x, m.bc, d), th, U[inv]


def _refined_box():
    m, th, U = M.mixed_bc_box(4, 3, 3, seed=5)
    r = PR.RefinableMesh(m, U)
    for every, off in ((7, 3), (5, 1)):
        ne = r.sizes()[0]
        el, rl = r.refine_list((np.arange(ne) % every == off).astype(np.uint8))
        r.apply(el, rl, PR.MODE_3D)
    m2, U2, _, _ = r.get()
    return m2, th, U2


def _case(name):
    if name == "box3d":
        return M.mixed_bc_box()
    if name == "plate":
        return M.flat_plate(12, 6, 5)
    if name == "lid":
        return M.lid_cavity(8)
    if name == "shuffled":
        return _shuffled_box()
    if name == "refined":
        return _refined_box()
    raise KeyError(name)


def _dense(m, R):
    n = m.n_node * m.neqn
    A = np.zeros((n, n))
    for e in range(m.n_elem):
        dofs = (m.conn[e][:, None] * m.neqn + np.arange(m.neqn)).ravel()
        A[np.ix_(dofs, dofs)] += R[e]
    return A


@pytest.mark.parametrize("case", ["box3d", "plate", "lid", "shuffled", "refined"])
def test_rows_product_parity(gpu_available, case):
    m, th, U = _case(case)
    ctx = FDVContext(m, th, U, blocks=("fp64", "rows"))
    ebe = FDVContext(m, th, U, blocks=("fp64", "fp32"))
    dt = ctx.cfl(0.5)
    ctx.build(dt)
    ebe.build(dt)
    R, _, _ = ctx.blocks()
    assert np.array_equal(R, ebe.blocks()[0])
    rng = np.random.default_rng(4)
    x = rng.standard_normal(m.n_node * m.neqn)
    yo = O.matvec(m, R, x)
    nrm = np.linalg.norm(yo)
    y = ctx.matvec(x)
    err64 = np.linalg.norm(y - yo) / nrm
    assert np.linalg.norm(y - ebe.matvec(x)) <= 1e-13 * nrm
    y32 = ctx.matvec(x, precision=32)
    y3 = ctx.matvec(x, precision=3)
    err32 = np.linalg.norm(y32 - yo) / nrm
    err3 = np.linalg.norm(y3 - yo) / nrm
    print(case, "nodes", m.n_node, "rows FP64 rel err %.2e FP32 %.2e FP32 rows/FP64 vectors %.2e" % (err64, err32, err3))
    assert err64 <= 1e-13
    assert err32 <= 2e-6 and err3 <= 2e-6
    # FP32 rows are rounded FP64 sums: the product with FP64 vectors is that
    # of the rounded dense K
    A = _dense(m, R)
    y3d = A.astype(np.float32).astype(np.float64) @ x
    assert np.linalg.norm(y3 - y3d) <= 1e-13 * nrm
    # K^T x over the rows re-summed transposed (same slots, same element
    # order), in FP64 and FP32; K x afterwards re-sums K's rows, bit for bit
    yt = ctx.matvec(x, transpose=True)
    assert np.linalg.norm(yt - O.matvec(m, R, x, transpose=True)) <= 1e-13 * nrm
    y3t = ctx.matvec(x, precision=3, transpose=True)
    assert np.linalg.norm(y3t - A.astype(np.float32).astype(np.float64).T @ x) <= 1e-13 * nrm
    assert np.array_equal(ctx.matvec(x), y)


@pytest.mark.parametrize("case", ["box3d", "lid", "shuffled", "refined"])
def test_rows_solves(gpu_available, case):
    m, th, U = _case(case)
    ctx = FDVContext(m, th, U, blocks=("fp64", "rows"))
    dt = ctx.cfl(0.5)
    ctx.build(dt)
    R, rhs, _ = ctx.blocks()
    b = O.compose_rhs(m, rhs)
    A = _dense(m, R)
    xs = np.linalg.solve(A, b)
    nx, nb = np.linalg.norm(xs), np.linalg.norm(b)
    # a relative residual of 1e-10 bounds the solution error by cond(A) 1e-10
    tol_x = 2e-10 * np.linalg.cond(A)

    def check(st, AA, xref):
        du = ctx.du()
        assert st.rel_resid <= 1e-10
        assert np.linalg.norm(AA @ du - b) <= 1.5e-10 * nb      # the true residual, on the CPU
        assert np.linalg.norm(du - xref) <= tol_x * np.linalg.norm(xref)

    st, _ = ctx.solve("fp64", restarts=60, m=30, rtol=1e-10)
    check(st, A, xs)
    st, _ = ctx.solve("fp64", restarts=60, m=30, rtol=1e-10, precond="nodeblock")
    check(st, A, xs)
    st, _ = ctx.solve("mixed", restarts=20, m=30, rtol=1e-10, inner_m=20, inner_rtol=1e-4,
                      inner_restarts=3, precond="nodeblock")
    print(case, "cond %.1e" % np.linalg.cond(A), "mixed on rows: sweeps", st.restarts,
          "FP32 products", st.matvecs32, "FP64", st.matvecs64)
    assert st.matvecs32 > 0
    check(st, A, xs)
    # the transposed solves run on the rows of K^T, the node-block
    # preconditioner stays that of K (applied transposed)
    xt = np.linalg.solve(A.T, b)
    st, _ = ctx.solve("fp64", restarts=60, m=30, rtol=1e-10, transpose=True, precond="nodeblock")
    check(st, A.T, xt)
    assert nx > 0
    st, _ = ctx.solve("mixed", restarts=20, m=30, rtol=1e-10, inner_m=20, inner_rtol=1e-4,
                      inner_restarts=3, transpose=True, precond="nodeblock")
    assert st.matvecs32 > 0
    check(st, A.T, xt)
    # and K x = b again after them
    st, _ = ctx.solve("fp64", restarts=60, m=30, rtol=1e-10, precond="nodeblock")
    check(st, A, xs)


def test_reference_solve_stays_element_by_element(gpu_available):
    """The reference solve of a context with node rows multiplies element by
    element like a context without: the same iterations and the same step.
    The rows context numbers its nodes in its own (locality) order, so the
    solve's dot products sum in another order: the steps agree to rounding,
    not bit for bit."""
    m, th, U = _case("box3d")
    a = FDVContext(m, th, U, blocks=("fp64", "rows"))
    b = FDVContext(m, th, U)
    sa = a.step(0.5, mode="reference", restarts=1, m=20)
    sb = b.step(0.5, mode="reference", restarts=1, m=20)
    assert sa.solve.iters == sb.solve.iters
    Ua, _ = a.get_state()
    Ub, _ = b.get_state()
    d = np.linalg.norm(Ua - Ub) / np.linalg.norm(Ub - U)
    print(f"rows context vs EBE context, reference step: |dU| rel diff {d:.1e}")
    assert d <= 1e-10


def test_mixed_step_on_rows(gpu_available):
    m, th, U = M.compression_ramp(12, 5, 6)
    ctx = FDVContext(m, th, U, blocks=("fp64", "rows"))
    ref = FDVContext(m, th, U, blocks=("fp64", "fp32"))
    kw = dict(restarts=10, m=30, rtol=1e-10, inner_m=30, inner_rtol=1e-4, inner_restarts=3, precond="nodeblock")
    st = ctx.step(0.5, mode="mixed", **kw)
    sr = ref.step(0.5, mode="mixed", **kw)
    print("rows step: build %.3f ms assemble %.3f ms solve %.3f ms, FP32 products %d" %
          (st.ms_build, st.ms_assemble, st.ms_solve, st.n_matvec32))
    assert st.ms_assemble > 0 and sr.ms_assemble == 0
    assert st.n_matvec32 > 0 and st.solve.rel_resid <= 1e-10
    U1, U0 = ctx.get_state()
    R1, _ = ref.get_state()
    du, dr = U1 - U0, R1 - U0
    assert np.linalg.norm(du - dr) <= 1e-8 * np.linalg.norm(dr)
    # rows with FP32 element blocks need the FP64 ones (rows alone build in
    # place: tests/test_gpu_rows_direct.py)
    with pytest.raises(CFDVError):
        FDVContext(m, th, U, blocks=("fp32", "rows"))


@pytest.mark.parametrize("ws", [1, 2, 3])
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_warm_start_reaches_the_same_steps(gpu_available, mode, ws):
    """warm_start 1 begins each solve from the last update's du scaled by
    dt / dt_prev, warm_start 2 from dt (2 r_1 - r_2), 3 from
    dt (3 r_1 - 3 r_2 + r_3), r_k = du_k / dt_k: four steps reach the same
    states as cold-started steps (both stop at rtol 1e-10 of the true
    residual), the first solve after create or set_state ignores the option
    (bit-identical to a cold one), the second falls back to 1 (bit-identical
    to it), and the warm solves need fewer products."""
    m, th, U = M.compression_ramp(12, 5, 6)
    kw = dict(restarts=40, m=30, rtol=1e-10, inner_m=30, inner_rtol=1e-4, inner_restarts=3, precond="nodeblock")
    cold = FDVContext(m, th, U, blocks=("fp64", "rows"))
    warm = FDVContext(m, th, U, blocks=("fp64", "rows"))
    one = FDVContext(m, th, U, blocks=("fp64", "rows"))
    nc = nw = 0
    for k in range(4):
        sc = cold.step(0.5, mode=mode, **kw)
        sw = warm.step(0.5, mode=mode, warm_start=ws, **kw)
        assert sc.solve.rel_resid <= 1e-10 and sw.solve.rel_resid <= 1e-10
        Uc, U0c = cold.get_state()
        Uw, U0w = warm.get_state()
        if k == 0:
            assert np.array_equal(Uc, Uw)
        else:
            nc += sc.solve.matvecs32 + sc.solve.matvecs64
            nw += sw.solve.matvecs32 + sw.solve.matvecs64
        if k < 2:
            one.step(0.5, mode=mode, warm_start=1, **kw)
            assert np.array_equal(one.get_state()[0], Uw)
        assert np.linalg.norm((Uw - U0w) - (Uc - U0c)) <= 1e-7 * np.linalg.norm(Uc - U0c)
    print(mode, "warm_start", ws, "products over steps 2-4: cold", nc, "warm", nw)
    assert nw <= nc
    # set_state forgets the increments: the next solve is cold again
    warm.set_state(U)
    cold.set_state(U)
    sw = warm.step(0.5, mode=mode, warm_start=ws, **kw)
    sc = cold.step(0.5, mode=mode, **kw)
    assert np.array_equal(warm.get_state()[0], cold.get_state()[0])


@pytest.mark.parametrize("case", ["box3d", "shuffled", "refined"])
def test_rows_eight_vector_product(gpu_available, case):
    """The 8-vector product over the FP32 rows is, vector by vector, the
    single FP32 row product bit for bit (same sums in the same order), and
    within 1e-6 of the oracle's EBE product; likewise K^T X over the rows
    re-summed transposed."""
    m, th, U = _case(case)
    ctx = FDVContext(m, th, U, blocks=("fp64", "rows"))
    ctx.build(ctx.cfl(0.5))
    R, _, _ = ctx.blocks()
    rng = np.random.default_rng(21)
    X = rng.standard_normal((ctx.ndof, 8)).astype(np.float32)
    Y = ctx.matvec_multi(X)
    for v in range(8):
        x = X[:, v].astype(np.float64)
        y1 = ctx.matvec(x, precision=32)
        assert np.array_equal(Y[:, v], y1.astype(np.float32))
        yo = O.matvec(m, R, x)
        assert np.linalg.norm(Y[:, v] - yo) <= 2e-6 * np.linalg.norm(yo)
    Yt = ctx.matvec_multi(X, transpose=True)
    for v in range(8):
        x = X[:, v].astype(np.float64)
        assert np.array_equal(Yt[:, v], ctx.matvec(x, precision=32, transpose=True).astype(np.float32))
        yo = O.matvec(m, R, x, transpose=True)
        assert np.linalg.norm(Yt[:, v] - yo) <= 2e-6 * np.linalg.norm(yo)
    assert np.array_equal(ctx.matvec_multi(X), Y)


def test_rows_with_an_unused_node(gpu_available):
    """A node no element references gets a row of its own column only (a
    zero block): K x is zero there, as element by element."""
    m, th, U = M.mixed_bc_box(3, 3, 2, seed=3)
    x = np.vstack([m.x, m.x[-1:] + 0.5])
    d = np.vstack([m.dirichlet, np.zeros((1, m.neqn), m.dirichlet.dtype)])
    m2 = M.Mesh(3, 8, 5, m.conn, x, m.bc, d)
    U2 = np.vstack([U, U[-1:]])
    rows = FDVContext(m2, th, U2, blocks=("fp64", "rows"))
    ebe = FDVContext(m2, th, U2)
    dt = rows.cfl(0.5)
    rows.build(dt)
    ebe.build(dt)
    v = np.random.default_rng(5).standard_normal(m2.n_node * m2.neqn)
    yr, ye = rows.matvec(v), ebe.matvec(v)
    assert np.all(yr[-m2.neqn:] == 0.0)
    assert np.linalg.norm(yr - ye) <= 1e-13 * np.linalg.norm(ye)


def test_preconditioner_from_transposed_rows(gpu_available):
    """The node-block preconditioner formed while the rows hold K^T (a K^T x
    came first after the build) is K's, read transposed from the diagonal
    slots: the transposed solve is bit-identical to one whose
    preconditioner was formed from K's rows."""
    m, th, U = _case("box3d")
    a = FDVContext(m, th, U, blocks=("fp64", "rows"))
    b = FDVContext(m, th, U, blocks=("fp64", "rows"))
    dt = a.cfl(0.5)
    a.build(dt)
    b.build(dt)
    x = np.random.default_rng(9).standard_normal(a.ndof)
    a.matvec(x, transpose=True)           # a: rows now hold K^T, P not formed yet
    b.solve("fp64", restarts=60, m=30, rtol=1e-10, precond="nodeblock")   # b: P formed from K's rows
    sa, _ = a.solve("fp64", restarts=60, m=30, rtol=1e-10, transpose=True, precond="nodeblock")
    sb, _ = b.solve("fp64", restarts=60, m=30, rtol=1e-10, transpose=True, precond="nodeblock")
    assert sa.iters == sb.iters
    assert np.array_equal(a.du(), b.du())
