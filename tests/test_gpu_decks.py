"""GPU parity on three of the reference's example decks against the oracle's
committed golden outputs (tests/golden/*.npz, written by make_golden.py).

Per step, the FP64 HIP path stays within 1e-10 relative of the oracle: dt,
s1..s4, the ApplyBC'd element blocks of every bc code, the element rhs, and
the reference-mode GMRES update with the deck's own gmresiter/gmresrestart.

Where a deck's step is ill-conditioned the bound is measured, not assumed:
make_golden.py perturbs the oracle's own element blocks and element rhs by
1e-16 relative noise and stores how far du moves.  On 3dtestlagrange the
reference's GMRES (MPI_GMRES.h:756-780 reads the rotated Hessenberg
transposed) turns that into a change of du of up to ~0.2 at step 0; on
lid50 the assembled b cancels heavily in the fluid at rest, and du moves by
~1e-10.  No implementation with another summation order can do better than
that spread, so the du tolerance is max(1e-10, 100 x spread); the iteration
count must still match exactly.

The solver itself is then pinned apart from the build
(test_deck_reference_solver_trace): the oracle's reference GMRES runs on the
GPU's own element blocks and rhs, and per restart the GPU's Hessenberg
matrix (as the Arnoldi process leaves it), its rotated right-hand side g,
the Arnoldi step count and du must equal the oracle's column by column to
max(1e-12, 100 x the spread the oracle's own values show under 1e-16
relative noise on those blocks and rhs), measured in the test.  The first
Arnoldi columns agree to ~1e-15; later ones inherit the Krylov process's
own amplification of rounding (on 3dtestlagrange step 0 its last column
moves by 7e-9 under 1e-16 noise), which the measured bound follows.  The same systems solved with
textbook GMRES (FP64 and mixed paths) are well conditioned and are checked
against the oracle's 1e-12 solution at 1e-9.
"""
import os

import numpy as np
import pytest

from cfdv_amd import mesh as M
from cfdv_amd.fdv import FDVContext
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DECKS = ["3dtestlagrange", "3dtestramp_new", "lid50"]


def _load(name):
    mesh, th, U, _ = M.read_deck(os.path.join(GOLD, "decks", name))
    return mesh, th, U, np.load(os.path.join(GOLD, f"{name}.npz"))


@pytest.mark.parametrize("name", DECKS)
def test_deck_build_parity(gpu_available, name):
    mesh, th, U, g = _load(name)
    print(f"deck {name}: {mesh.n_elem} elements", flush=True)
    ctx = FDVContext(mesh, th, U, blocks=("fp64", "fp32"))
    dt = ctx.cfl(float(g["cfl"]))
    assert dt == pytest.approx(float(g["dt"]), rel=1e-14)
    ctx.build(float(g["dt"]))
    R, rhs, s = ctx.blocks()
    assert np.allclose(s, g["s"], rtol=1e-13, atol=0)
    sel = g["sel"]
    num = np.linalg.norm((R[sel] - g["R_sel"]).reshape(len(sel), -1), axis=1)
    den = np.linalg.norm(g["R_sel"].reshape(len(sel), -1), axis=1)
    print("R_sel rel err max", (num / den).max())
    assert (num / den).max() < 1e-11
    fro = np.linalg.norm(R.reshape(len(R), -1), axis=1)
    assert np.allclose(fro, g["R_fro"], rtol=1e-11)
    rerr = np.linalg.norm(rhs - g["rhs"]) / np.linalg.norm(g["rhs"])
    print("rhs rel err", rerr)
    assert rerr < 1e-10
    b = ctx.rhs_vector()
    assert np.linalg.norm(b - g["b"]) <= 1e-10 * np.linalg.norm(g["b"])


@pytest.mark.parametrize("name", DECKS)
def test_deck_reference_steps(gpu_available, name):
    mesh, th, U, g = _load(name)
    ctx = FDVContext(mesh, th, U)
    starts = [g["U0"], g["U_steps"][0]]
    for k in range(2):
        ctx.set_state(starts[k])
        st = ctx.step(float(g["cfl"]), mode="reference", restarts=int(g["restart"]),
                      m=int(g["mdim"]))
        U1, U0 = ctx.get_state()
        du = (U1 - U0).ravel()
        err = np.linalg.norm(du - g["du_steps"][k]) / np.linalg.norm(g["du_steps"][k])
        print(f"deck {name} step {k}: iters {st.solve.iters} du rel err {err:.3e}", flush=True)
        assert st.dt == pytest.approx(float(g["dt_steps"][k]), rel=1e-13)
        assert st.solve.iters == int(g["iters_steps"][k])
        tol = max(1e-10, 100 * float(g["sens_steps"][k]))
        assert err < tol
        # Urms = sqrt(sum U^2 / sum U0^2) - 1 cancels against 1: its absolute
        # floor is the rounding of a sum of n squares in another order (~n eps)
        assert np.allclose(st.urms[:mesh.neqn], g["rms_steps"][k], rtol=max(1e-8, tol),
                           atol=mesh.n_node * 2.3e-16)


def _col_err(a, b, scale):
    """max |a - b| per column (last axis = Arnoldi column) over max |b|"""
    return np.abs(a - b).max(axis=0) / scale


@pytest.mark.parametrize("name", ["3dtestlagrange", "lid50"])
def test_deck_reference_solver_trace(gpu_available, name):
    mesh, th, U, g = _load(name)
    restart, mdim = int(g["restart"]), int(g["mdim"])
    ctx = FDVContext(mesh, th, U)
    rng = np.random.default_rng(20261204)
    for k, Uk in enumerate((g["U0"], g["U_steps"][0])):
        ctx.set_state(Uk)
        ctx.build(ctx.cfl(float(g["cfl"])))
        st, _ = ctx.solve("reference", restarts=restart, m=mdim)
        du = ctx.du()
        H, gg, mm = ctx.reference_trace()
        R, rhs, _ = ctx.blocks()
        du_o, st_o, H_o, g_o, mm_o = O.gmres_quirk_trace(mesh, R, rhs, restart, mdim)
        assert np.array_equal(mm, mm_o) and st.iters == st_o.iters
        # the oracle's own spread under 1e-16 relative noise on its inputs
        sH = np.zeros_like(H_o)
        sg = np.zeros_like(g_o)
        sdu = 0.0
        for nR, nr in ((1e-16, 0.0), (0.0, 1e-16)):
            Rp = R * (1 + nR * rng.standard_normal(R.shape))
            rp = rhs * (1 + nr * rng.standard_normal(rhs.shape))
            du_p, _, H_p, g_p, _ = O.gmres_quirk_trace(mesh, Rp, rp, restart, mdim)
            sH = np.maximum(sH, np.abs(H_p - H_o))
            sg = np.maximum(sg, np.abs(g_p - g_o))
            sdu = max(sdu, np.linalg.norm(du_p - du_o) / np.linalg.norm(du_o))
        for r in range(len(mm)):
            n = int(mm[r])
            hs = np.abs(H_o[r]).max()
            e_h = _col_err(H[r][:, :n], H_o[r][:, :n], hs)
            t_h = np.maximum(1e-12, 100 * sH[r][:, :n].max(axis=0) / hs)
            gs = np.abs(g_o[r]).max()
            e_g = np.abs(gg[r][:n + 1] - g_o[r][:n + 1]) / gs
            t_g = np.maximum(1e-12, 100 * sg[r][:n + 1] / gs)
            print(f"deck {name} step {k} restart {r}: {n} Arnoldi steps", flush=True)
            print("  H column err " + " ".join(f"{e:.1e}" for e in e_h))
            print("  H column tol " + " ".join(f"{t:.0e}" for t in t_h))
            print(f"  g err max {e_g.max():.1e} (tol min {t_g.min():.0e})", flush=True)
            assert np.all(e_h <= t_h), "Hessenberg column beyond the oracle's own rounding spread"
            assert np.all(e_g <= t_g), "rotated g beyond the oracle's own rounding spread"
        err = np.linalg.norm(du - du_o) / np.linalg.norm(du_o)
        tol = max(1e-12, 100 * sdu)
        print(f"deck {name} step {k}: du rel err vs the oracle on the same blocks {err:.2e} (tol {tol:.0e})",
              flush=True)
        assert err <= tol


@pytest.mark.parametrize("name", DECKS)
def test_deck_fp64_and_mixed_solves(gpu_available, name):
    mesh, th, U, g = _load(name)
    ctx = FDVContext(mesh, th, U, blocks=("fp64", "fp32"))
    ctx.build(float(g["dt"]))
    ref = g["du_true"]
    st, _ = ctx.solve("fp64", restarts=400, m=30, rtol=1e-12)
    e64 = np.linalg.norm(ctx.du() - ref) / np.linalg.norm(ref)
    st2, _ = ctx.solve("mixed", restarts=40, m=30, rtol=1e-12, inner_m=30, inner_rtol=1e-4,
                       inner_restarts=4)
    emx = np.linalg.norm(ctx.du() - ref) / np.linalg.norm(ref)
    print(f"deck {name}: fp64 rel resid {st.rel_resid:.2e} err {e64:.2e}; "
          f"mixed rel resid {st2.rel_resid:.2e} err {emx:.2e}", flush=True)
    assert st.rel_resid <= 1e-12 and st2.rel_resid <= 1e-12
    assert e64 < 1e-9 and emx < 1e-9
