"""Writes the golden fixtures under tests/golden/.

1. Copies three of the reference's example decks (data files only: control,
   aero, binarynodes, binaryelements; plus the ASCII nodes/elements/faces of
   the small 3D deck, which pin the binary record layout and the face node
   tables) from the reference checkout into tests/golden/decks/:
     3dtestlagrange  324 hex8, bc -1 1 2 22 6 7 32, isothermal (Twall 2.8)
     3dtestramp_new  5814 hex8, bc -1 1 2 4 8 21 22, adiabatic
     lid50           Lid/50x50, 2401 quad4, bc -1 1
2. Runs the CPU oracle on each deck and stores, per deck, an .npz with
   s1..s4 per element, the ApplyBC'd element blocks of a selection of
   elements (every bc code, interior elements) and the Frobenius norm of
   every block, the element rhs, dt at the deck's CFL, the composed b, and
   two reference-mode GMRES steps with the deck's gmresiter/gmresrestart
   (du, U after each step, RMS rows).  For each step it also stores how far
   du moves when the element blocks and/or the element rhs are perturbed by
   1e-16 relative noise: the reference's back substitution (quirk:
   transposed read) can amplify rounding enormously, and the assembled b
   cancels heavily where the flow is uniform, so a parity tolerance below
   that spread would test summation order, not the algorithm.  Finally the textbook-GMRES solution
   of the first step's system to a 1e-12 relative residual.

3. Refinement (``refine_<deck>.npz``, ``--refine`` writes only these): two
   --adap passes of the refinement oracle (oracle/refine.py:
   AdaptiveRefineList + MeshRefine2D, node match on x and y as the reference,
   quirk Q14) on lid50 and 3dtestlagrange from the state after the two golden
   steps.  The flags of each pass are the FDVParam s1 > 0.8 pass of the
   oracle at that pass's state, restricted to a window of the domain so that
   the fixtures stay small; they are stored with the outputs (elelist, level
   breakpoints, connectivity, bc, coordinates, U, U0, Dirichlet flags,
   adap, Hnm, refine levels).

4. The Tecplot writer's fixture (``--tecplot`` writes only this): the deck
   backstep_refined_new (control, aero, gzipped binary records) and the
   reference's own output on it at t = 0, tecplot_0.plt (gzipped).

5. Configuration C1 (``--c1`` writes only this, ``lid32_c1.npz``): the 2D
   lid-driven cavity on 32 x 32 quad4 (cfdv_amd.mesh.lid_cavity(32): M 0.1,
   Re 100, Pr 0.72, gamma 1.4, at rest with rho = T = 1), 100 time steps at
   CFL 0.5 with the reference GMRES as the Lid deck runs it (gmresiter 20,
   gmresrestart 1): U after every step, dt, Arnoldi steps, RMS rows, and per
   step how far du moves under 1e-16 relative noise on the element blocks
   and rhs (as for the decks above).

Run from the repository root:
    python tests/golden/make_golden.py [--refine | --tecplot | --c1]
"""
from __future__ import annotations

import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cfdv_amd import mesh as M  # noqa: E402
from oracle import oracle as O  # noqa: E402
from oracle import refine as OR  # noqa: E402

SRC = "/root/reference/example"
DECKS = {
    "3dtestlagrange": ("3dtestlagrange", True),
    "3dtestramp_new": ("3dtestramp_new", False),
    "lid50": ("Lid/50x50", False),
}
BIN_FILES = ("control", "aero", "binarynodes", "binaryelements")
ASCII_FILES = ("nodes", "elements", "faces")


TECPLOT_DECK = "backstep_refined_new"   # ships tecplot_0.plt, the writer's own output at t = 0


def copy_tecplot_fixture():
    """backstep_refined_new (18400 quad4): its control, aero and binary
    records (gzipped) and the reference's own tecplot_0.plt (gzipped)"""
    import gzip
    src = os.path.join(SRC, TECPLOT_DECK)
    dst = os.path.join(HERE, "decks", TECPLOT_DECK)
    os.makedirs(dst, exist_ok=True)
    for f in ("control", "aero"):
        shutil.copyfile(os.path.join(src, f), os.path.join(dst, f))
        os.chmod(os.path.join(dst, f), 0o644)
    for f, out in (("binarynodes", os.path.join(dst, "binarynodes.gz")),
                   ("binaryelements", os.path.join(dst, "binaryelements.gz")),
                   ("tecplot_0.plt", os.path.join(HERE, "tecplot_backstep_0.plt.gz"))):
        with open(os.path.join(src, f), "rb") as fi, gzip.open(out, "wb", compresslevel=9) as fo:
            fo.write(fi.read())


def copy_decks():
    for name, (rel, with_ascii) in DECKS.items():
        dst = os.path.join(HERE, "decks", name)
        os.makedirs(dst, exist_ok=True)
        files = BIN_FILES + (ASCII_FILES if with_ascii else ())
        for f in files:
            shutil.copyfile(os.path.join(SRC, rel, f), os.path.join(dst, f))
            os.chmod(os.path.join(dst, f), 0o644)


def select_elements(mesh, per_code=3, n_interior=8):
    sel = []
    for code in sorted(set(mesh.bc.ravel().tolist()) - {0}):
        idx = np.nonzero((mesh.bc == code).any(axis=1))[0]
        sel.extend(idx[:per_code].tolist())
        sel.extend(idx[-1:].tolist())
    interior = np.nonzero((mesh.bc == 0).all(axis=1))[0]
    if len(interior):
        sel.extend(interior[np.linspace(0, len(interior) - 1, n_interior).astype(int)].tolist())
    return np.array(sorted(set(sel)), dtype=np.int64)


def make(name):
    mesh, th, U, ctl = M.read_deck(os.path.join(HERE, "decks", name))
    cfl = float(ctl["cfl"])
    mdim = int(float(ctl["gmresiter"]))
    restart = int(float(ctl["gmresrestart"]))
    dt = O.cfl(th, mesh, U, cfl)
    s, ties = O.fdvparam_all(th, mesh, U)
    R, rhs, _ = O.build(th, mesh, U, dt)
    sel = select_elements(mesh)
    b = O.compose_rhs(mesh, rhs)
    steps_U, steps_du, steps_rms, steps_dt, steps_it, steps_sens = [], [], [], [], [], []
    Uc = U.copy()
    rng = np.random.default_rng(20261203)
    for _ in range(2):
        dtc = O.cfl(th, mesh, Uc, cfl)
        Rc, rhsc, _ = O.build(th, mesh, Uc, dtc)
        du_a, _ = O.gmres_quirk(mesh, Rc, rhsc, restart, mdim)
        spread = 0.0
        for noise_R, noise_rhs in ((1e-16, 0.0), (0.0, 1e-16), (1e-16, 1e-16)):
            Rp = Rc * (1 + noise_R * rng.standard_normal(Rc.shape))
            rp = rhsc * (1 + noise_rhs * rng.standard_normal(rhsc.shape))
            du_b, _ = O.gmres_quirk(mesh, Rp, rp, restart, mdim)
            spread = max(spread, np.linalg.norm(du_b - du_a) / np.linalg.norm(du_a))
        steps_sens.append(spread)
        U1, U0, st = O.step(th, mesh, Uc, cfl=cfl, restart=restart, mdim=mdim, mode=0)
        steps_U.append(U1)
        steps_du.append((U1 - U0).ravel())
        steps_rms.append(np.array(st.urms)[:mesh.neqn])
        steps_dt.append(st.dt)
        steps_it.append(st.gm.iters)
        Uc = U1
    du_true, gst = O.gmres_true(mesh, R, b, 30, 400, 1e-12)
    out = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(
        out, cfl=cfl, mdim=mdim, restart=restart, dt=dt, s=s, ties=ties,
        sel=sel, R_sel=R[sel], R_fro=np.linalg.norm(R.reshape(len(R), -1), axis=1),
        rhs=rhs, b=b, U0=U, U_steps=np.array(steps_U), du_steps=np.array(steps_du),
        rms_steps=np.array(steps_rms), dt_steps=np.array(steps_dt),
        iters_steps=np.array(steps_it), sens_steps=np.array(steps_sens), du_true=du_true)
    print(f"{name}: {mesh.n_elem} elements, {mesh.n_node} nodes, dt {dt:.6e}, "
          f"{len(sel)} selected blocks, ties {ties}, step sensitivity {steps_sens}, "
          f"true GMRES iters {gst.iters}, -> {os.path.getsize(out)} bytes")


# refinement windows per pass: (lower corner, upper corner) of element centroids
REFINE_WINDOWS = {
    "lid50": (((0.75, 0.75), (1.0, 1.0)), ((0.88, 0.88), (1.0, 1.0))),
    "3dtestlagrange": (((0.0, 0.0, 0.0), (0.5, 0.25, 0.25)), ((0.0, 0.0, 0.0), (0.5, 0.25, 0.25))),
}


def make_refine(name):
    mesh, th, _, _ = M.read_deck(os.path.join(HERE, "decks", name))
    g = np.load(os.path.join(HERE, f"{name}.npz"))
    U, U0 = g["U_steps"][-1], g["U_steps"][-2]
    rm = OR.RMesh.from_arrays(mesh.ndim, mesh.nnod, mesh.neqn, mesh.conn, mesh.bc, mesh.x, U, U0,
                              mesh.dirichlet)
    out = {}
    cur = mesh
    for p, (lo, hi) in enumerate(REFINE_WINDOWS[name]):
        Uc = np.array(rm.U)
        s, _ = O.fdvparam_all(th, cur, Uc)
        cent = np.array(rm.x)[np.array(rm.conn)].mean(axis=1)
        inside = np.all((cent >= np.array(lo)) & (cent <= np.array(hi)), axis=1)
        flags = (s[:, 0] > 0.8) & inside
        el, rl = OR.adaptive_refine_list(rm, flags)
        OR.MeshRefine(rm, el, rl, mode="parity")
        conn, bc, x, Un, U0n, d, adap, hnm, lv = rm.arrays()
        cur = M.Mesh(mesh.ndim, mesh.nnod, mesh.neqn, conn, x, bc, d, adap=adap, hnm=hnm)
        out.update({f"p{p}_flags": flags.astype(np.uint8), f"p{p}_elelist": np.array(el, np.int32),
                    f"p{p}_breaks": np.array(rl, np.int32), f"p{p}_conn": conn, f"p{p}_bc": bc,
                    f"p{p}_x": x, f"p{p}_U": Un, f"p{p}_U0": U0n, f"p{p}_dirichlet": d,
                    f"p{p}_adap": adap, f"p{p}_hnm": hnm, f"p{p}_level": lv})
        print(f"refine {name} pass {p}: {int(flags.sum())} flagged, elelist {len(el)}, levels {rl}, "
              f"-> {rm.n_elem} elements, {rm.n_node} nodes, {int(adap.sum())} adapted")
    path = os.path.join(HERE, f"refine_{name}.npz")
    np.savez_compressed(path, U=U, U0=U0, **out)
    print(f"  -> {os.path.getsize(path)} bytes")


C1_STEPS = 100


def make_c1():
    mesh, th, U = M.lid_cavity(32)
    cfl, mdim, restart = 0.5, 20, 1
    Us, dts, its, rms, sens = [U], [], [], [], []
    rng = np.random.default_rng(20261204)
    Uc = U
    for k in range(C1_STEPS):
        dt = O.cfl(th, mesh, Uc, cfl)
        R, rhs, _ = O.build(th, mesh, Uc, dt)
        du_a, _ = O.gmres_quirk(mesh, R, rhs, restart, mdim)
        spread = 0.0
        for noise_R, noise_rhs in ((1e-16, 0.0), (0.0, 1e-16), (1e-16, 1e-16)):
            Rp = R * (1 + noise_R * rng.standard_normal(R.shape))
            rp = rhs * (1 + noise_rhs * rng.standard_normal(rhs.shape))
            du_b, _ = O.gmres_quirk(mesh, Rp, rp, restart, mdim)
            spread = max(spread, np.linalg.norm(du_b - du_a) / np.linalg.norm(du_a))
        U1, U0, st = O.step(th, mesh, Uc, cfl=cfl, restart=restart, mdim=mdim, mode=0)
        assert np.array_equal(Uc.ravel() + du_a, U1.ravel())
        Us.append(U1)
        dts.append(st.dt)
        its.append(st.gm.iters)
        rms.append(np.array(st.urms)[:mesh.neqn])
        sens.append(spread)
        Uc = U1
    path = os.path.join(HERE, "lid32_c1.npz")
    np.savez_compressed(path, cfl=cfl, mdim=mdim, restart=restart, U=np.array(Us), dt=np.array(dts),
                        iters=np.array(its), rms=np.array(rms), sens=np.array(sens))
    print(f"C1 lid 32x32: {C1_STEPS} steps, dt {dts[0]:.4e}..{dts[-1]:.4e}, iters {min(its)}..{max(its)}, "
          f"sensitivity {min(sens):.1e}..{max(sens):.1e} -> {os.path.getsize(path)} bytes")


def main():
    if "--tecplot" in sys.argv:
        copy_tecplot_fixture()
        return
    if "--c1" in sys.argv:
        make_c1()
        return
    if "--refine" not in sys.argv:
        if os.path.isdir(SRC):
            copy_decks()
            copy_tecplot_fixture()
        for name in DECKS:
            make(name)
    for name in REFINE_WINDOWS:
        make_refine(name)


if __name__ == "__main__":
    main()
