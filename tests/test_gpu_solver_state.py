"""Solver state across calls of the C ABI (cfdv_build / cfdv_solve /
cfdv_update): results must not depend on call history.

* A second build at another dt forms new node-block inverses: its
  preconditioned solve equals a fresh context's bit for bit.
* A repeated solve of one build (a retry) starts from the same warm-start
  guess as the first and returns the same du; the next step then equals a
  run with one solve per build, bit for bit.
* The FP64 solve meets the requested tolerance on the true residual.
"""
import numpy as np
import pytest

from cfdv_amd import mesh as M
from cfdv_amd.fdv import FDVContext

pytestmark = pytest.mark.gpu


def test_preconditioner_follows_the_build(gpu_available):
    m, th, U = M.mixed_bc_box(6, 5, 4)
    a = FDVContext(m, th, U)
    dt = a.cfl(0.5)
    a.build(dt)
    a.solve("fp64", restarts=50, m=20, rtol=1e-10, precond="nodeblock")
    a.build(0.5 * dt)
    a.solve("fp64", restarts=50, m=20, rtol=1e-10, precond="nodeblock")
    b = FDVContext(m, th, U)
    b.build(0.5 * dt)
    b.solve("fp64", restarts=50, m=20, rtol=1e-10, precond="nodeblock")
    print("second-build du equal to a fresh context's:", np.array_equal(a.du(), b.du()), flush=True)
    assert np.array_equal(a.du(), b.du())


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_repeated_solve_keeps_the_warm_start(gpu_available, mode):
    m, th, U = M.mixed_bc_box(6, 5, 4)
    kw = dict(restarts=50, m=20, rtol=1e-10, precond="nodeblock", warm_start=2, inner_m=20,
              inner_rtol=1e-4, inner_restarts=4)
    runs = []
    for repeat in (False, True):
        c = FDVContext(m, th, U, blocks=("fp64", "fp32"))
        dus = []
        for step in range(4):
            c.build(c.cfl(0.5))
            c.solve(mode, **kw)
            d1 = c.du()
            if repeat:
                c.solve(mode, **kw)
                assert np.array_equal(c.du(), d1), f"step {step}: the repeated solve changed du"
            dus.append(d1)
            c.update()
        runs.append(dus)
    for k, (x, y) in enumerate(zip(*runs)):
        print(f"{mode} step {k}: equal {np.array_equal(x, y)}", flush=True)
        assert np.array_equal(x, y)


def test_fp64_solve_meets_rtol_on_the_true_residual(gpu_available):
    m, th, U = M.mixed_bc_box(6, 5, 4)
    c = FDVContext(m, th, U)
    c.build(c.cfl(0.5))
    for rtol in (1e-8, 1e-10, 1e-12):
        st, rc = c.solve("fp64", restarts=400, m=10, rtol=rtol)
        print(f"rtol {rtol:.0e}: rel resid {st.rel_resid:.3e}, restarts {st.restarts}", flush=True)
        assert rc == 0 and st.rel_resid <= rtol
