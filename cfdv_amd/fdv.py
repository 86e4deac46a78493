"""Host-side mirror of the reference interface for the FDV time step.

The reference drives one step as (src/main.cpp:162-206)::

    for_each(elements, ClearElement); for_each(nodes, NodeUnpack);
    CFL(elements, cfl, dt);
    FDVGalerkin fdv(nnod, neqn, ndim, nbnod, dt, NS); for_each(elements, ref(fdv));
    for_each(elements, ApplyBC(neqn, nnod, ndim, nbnod));
    gmres->iterate(elements); gmres->update(nodes); RMSErr(t, iter);

``FDVContext`` owns the device-resident mesh and state.  ``FDVGalerkin`` and
``ApplyBC`` keep the reference names: applying FDVGalerkin to the element set
builds every element block on the GPU (FDVParam, domain and boundary
integrals) and ApplyBC is fused into that same kernel.  ``HipGMRES`` has the
``GMRESbase`` shape (iterate/update, src/Utility/Solvers/GMRESbase.h:8-23).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from .mesh import Mesh, Thermo



_PRECOND = {"none": _lib.CFDV_PRECOND_NONE, "nodeblock": _lib.CFDV_PRECOND_NODE_BLOCK}

class CFDVError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"cfdv error {code}: {msg}")
        self.code = code


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class FDVContext:
    """One device-resident FDV problem (a ``cfdv_ctx``)."""

    def __init__(self, mesh: Mesh, thermo: Thermo, U: Optional[np.ndarray] = None,
                 blocks=("fp64",)):
        self.L = _lib.load()
        self.mesh = mesh
        self.thermo = thermo
        self._conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
        self._x = np.ascontiguousarray(mesh.x, dtype=np.float64)
        self._bc = np.ascontiguousarray(mesh.bc, dtype=np.int32)
        self._d = np.ascontiguousarray(mesh.dirichlet, dtype=np.uint8)
        self._adap = None if mesh.adap is None else np.ascontiguousarray(mesh.adap, dtype=np.uint8)
        self._hnm = None
        if mesh.hnm is not None:
            # Hnm(N, M) at N + nnod*M (the reference Tensor's column-major order)
            self._hnm = np.ascontiguousarray(np.transpose(mesh.hnm, (0, 2, 1)), dtype=np.float64)
        m = _lib.Mesh(mesh.ndim, mesh.nnod, mesh.neqn, mesh.n_elem, mesh.n_node,
                      self._conn.ctypes.data_as(C.POINTER(C.c_int32)), _dp(self._x),
                      self._bc.ctypes.data_as(C.POINTER(C.c_int32)),
                      self._d.ctypes.data_as(C.POINTER(C.c_uint8)),
                      self._adap.ctypes.data_as(C.POINTER(C.c_uint8)) if self._adap is not None
                      else C.cast(None, C.POINTER(C.c_uint8)),
                      _dp(self._hnm) if self._hnm is not None else C.cast(None, C.POINTER(C.c_double)))
        t = _lib.ThermoS(thermo.gamma, thermo.mach, thermo.reynolds, thermo.prandtl,
                         thermo.tinf, thermo.twall, int(thermo.adiabatic))
        flags = (_lib.CFDV_BLOCKS_FP64 if "fp64" in blocks else 0) | \
                (_lib.CFDV_BLOCKS_FP32 if "fp32" in blocks else 0) | \
                (_lib.CFDV_BLOCKS_ROWS if "rows" in blocks else 0)
        h = C.c_void_p()
        rc = self.L.cfdv_create(C.byref(m), C.byref(t), flags, C.byref(h))
        if rc != 0:
            raise CFDVError(rc, "cfdv_create failed")
        self.h = h
        self.ndof = mesh.n_node * mesh.neqn
        if U is not None:
            self.set_state(U)

    def close(self):
        if getattr(self, "h", None):
            self.L.cfdv_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise CFDVError(rc, self.L.cfdv_last_error(self.h).decode())

    # ---- state ----
    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.mesh.n_node, self.mesh.neqn)
        self._check(self.L.cfdv_set_state(self.h, _dp(U)))

    def get_state(self):
        U = np.zeros((self.mesh.n_node, self.mesh.neqn))
        U0 = np.zeros_like(U)
        self._check(self.L.cfdv_get_state(self.h, _dp(U), _dp(U0)))
        return U, U0

    def info(self):
        ne, nn, nc, ns = C.c_int64(), C.c_int64(), C.c_int(), C.c_int64()
        self._check(self.L.cfdv_mesh_info(self.h, C.byref(ne), C.byref(nn), C.byref(nc),
                                          C.byref(ns)))
        return {"n_elem": ne.value, "n_node": nn.value, "ncolors": nc.value, "n_slots": ns.value}

    def rows_info(self):
        """(node blocks, stored block slots) of the assembled rows"""
        nb, ns = C.c_int64(), C.c_int64()
        self._check(self.L.cfdv_rows_info(self.h, C.byref(nb), C.byref(ns)))
        return nb.value, ns.value

    # ---- step pieces ----
    def cfl(self, cfl: float) -> float:
        dt = C.c_double()
        self._check(self.L.cfdv_cfl(self.h, cfl, C.byref(dt)))
        return dt.value

    def build(self, dt: float):
        self._check(self.L.cfdv_build(self.h, dt))

    def blocks(self, want_R=True):
        """R [n_e][k][k] with R[e, i, j] = R(i, j); rhs [n_e][k]; s [n_e][4]."""
        k = self.mesh.k
        Rcm = np.zeros((self.mesh.n_elem, k, k)) if want_R else None
        rhs = np.zeros((self.mesh.n_elem, k))
        s = np.zeros((self.mesh.n_elem, 4))
        self._check(self.L.cfdv_get_blocks(self.h, _dp(Rcm) if want_R else None, _dp(rhs),
                                           _dp(s)))
        R = np.ascontiguousarray(np.transpose(Rcm, (0, 2, 1))) if want_R else None
        return R, rhs, s

    def blocks32(self):
        k = self.mesh.k
        Rcm = np.zeros((self.mesh.n_elem, k, k), dtype=np.float32)
        self._check(self.L.cfdv_get_blocks32(self.h, Rcm.ctypes.data_as(C.POINTER(C.c_float))))
        return np.ascontiguousarray(np.transpose(Rcm, (0, 2, 1)))

    def rhs_vector(self):
        b = np.zeros(self.ndof)
        self._check(self.L.cfdv_get_rhs_vector(self.h, _dp(b)))
        return b

    def solve(self, mode="reference", restarts=1, m=20, rtol=1e-8, inner_m=30,
              inner_rtol=1e-3, inner_restarts=4, transpose=False, check=True, precond="none",
              warm_start=False):
        """``precond``: "none" or "nodeblock" (node-block Jacobi, right-applied;
        FP64 and mixed modes).  ``warm_start`` (FP64 and mixed modes): 1 or
        True, start from the last update's du scaled by dt / dt_prev; 2 or 3,
        from the increment rate of the last two or three updates extrapolated
        linearly or quadratically."""
        md = {"reference": _lib.CFDV_SOLVE_REFERENCE, "fp64": _lib.CFDV_SOLVE_FP64,
              "mixed": _lib.CFDV_SOLVE_MIXED}[mode]
        o = _lib.SolveOpts(md, restarts, m, rtol, inner_m, inner_rtol, inner_restarts,
                           int(transpose), _PRECOND[precond], int(warm_start))
        st = _lib.SolveStats()
        rc = self.L.cfdv_solve(self.h, C.byref(o), C.byref(st))
        if check:
            self._check(rc)
        return st, rc

    def du(self):
        d = np.zeros(self.ndof)
        self._check(self.L.cfdv_get_du(self.h, _dp(d)))
        return d

    def reference_trace(self):
        """Per restart of the last reference-mode solve (MPI_GMRES.h:125-220):
        H [restarts][m+1][m] before the rotations, the rotated g
        [restarts][m+1] and the Arnoldi steps taken [restarts]."""
        m, nr = C.c_int32(), C.c_int32()
        self._check(self.L.cfdv_reference_trace(self.h, C.byref(m), C.byref(nr), None, None, None))
        H = np.zeros((nr.value, m.value, m.value + 1))
        g = np.zeros((nr.value, m.value + 1))
        mm = np.zeros(nr.value, dtype=np.int32)
        self._check(self.L.cfdv_reference_trace(self.h, None, None, _dp(H), _dp(g),
                                                mm.ctypes.data_as(C.POINTER(C.c_int32))))
        return H.transpose(0, 2, 1).copy(), g, mm

    def update(self):
        self._check(self.L.cfdv_update(self.h))

    def rms(self):
        u = np.zeros(5)
        self._check(self.L.cfdv_rms(self.h, _dp(u)))
        return u[:self.mesh.neqn]

    def step(self, cfl, mode="reference", **kw):
        md = {"reference": _lib.CFDV_SOLVE_REFERENCE, "fp64": _lib.CFDV_SOLVE_FP64,
              "mixed": _lib.CFDV_SOLVE_MIXED}[mode]
        o = _lib.SolveOpts(md, kw.get("restarts", 1), kw.get("m", 20), kw.get("rtol", 1e-8),
                           kw.get("inner_m", 30), kw.get("inner_rtol", 1e-3),
                           kw.get("inner_restarts", 4), 0, _PRECOND[kw.get("precond", "none")],
                           int(kw.get("warm_start", False)))
        st = _lib.StepStats()
        self._check(self.L.cfdv_step(self.h, cfl, C.byref(o), C.byref(st)))
        return st

    def refine_flags(self):
        f = np.zeros(self.mesh.n_elem, dtype=np.uint8)
        s1 = np.zeros(self.mesh.n_elem)
        nf, nt = C.c_int32(), C.c_int32()
        self._check(self.L.cfdv_refine_flags(self.h, f.ctypes.data_as(C.POINTER(C.c_uint8)),
                                             _dp(s1), C.byref(nf), C.byref(nt)))
        return f, s1, nt.value

    def refine_flags_noalloc(self, want_s1=True):
        """refine_flags through cfdv_refine_flags_noalloc (no device
        allocation: the library's fixed per-GPU scratch)."""
        f = np.zeros(self.mesh.n_elem, dtype=np.uint8)
        s1 = np.zeros(self.mesh.n_elem) if want_s1 else None
        nf, nt = C.c_int32(), C.c_int32()
        self._check(self.L.cfdv_refine_flags_noalloc(self.h, f.ctypes.data_as(C.POINTER(C.c_uint8)),
                                                     _dp(s1) if want_s1 else None, C.byref(nf), C.byref(nt)))
        return f, s1, nt.value, nf.value

    def color(self):
        c = np.zeros(self.mesh.n_elem, dtype=np.int32)
        n = C.c_int32()
        self._check(self.L.cfdv_color(self.h, c.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(n)))
        return c, n.value

    def matvec(self, x, precision=64, transpose=False):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.ndof)
        self._check(self.L.cfdv_matvec(self.h, _dp(x), _dp(y), precision, int(transpose)))
        return y

    def profile_matvec(self, precision=64, reps=10):
        ms = C.c_double()
        self._check(self.L.cfdv_profile_matvec(self.h, precision, reps, C.byref(ms)))
        return ms.value

    def matvec_multi(self, X, transpose=False):
        """Y = K X for X [ndof][8] (float32) in one pass over the FP32 blocks."""
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(self.ndof, 8)
        Y = np.zeros_like(X)
        self._check(self.L.cfdv_matvec_multi(self.h, X.ctypes.data_as(C.POINTER(C.c_float)),
                                             Y.ctypes.data_as(C.POINTER(C.c_float)), 8,
                                             int(transpose)))
        return Y

    def profile_matvec_multi(self, reps=10):
        ms = C.c_double()
        self._check(self.L.cfdv_profile_matvec_multi(self.h, reps, C.byref(ms)))
        return ms.value

    # ---- discrete adjoint of one step ----
    def set_thermo(self, thermo: Thermo):
        """Replace the Thermo constants (e.g. the Mach number); rebuild after."""
        t = _lib.ThermoS(thermo.gamma, thermo.mach, thermo.reynolds, thermo.prandtl,
                         thermo.tinf, thermo.twall, int(thermo.adiabatic))
        self._check(self.L.cfdv_set_thermo(self.h, C.byref(t)))
        self.thermo = thermo

    def adjoint_solve(self, g, mode="fp64", restarts=200, m=30, rtol=1e-12, inner_m=30,
                      inner_rtol=1e-4, inner_restarts=4, precond="nodeblock"):
        """lambda with K^T lambda = g for the blocks of the last build."""
        md = {"fp64": _lib.CFDV_SOLVE_FP64, "mixed": _lib.CFDV_SOLVE_MIXED}[mode]
        o = _lib.SolveOpts(md, restarts, m, rtol, inner_m, inner_rtol, inner_restarts, 1,
                           _PRECOND[precond])
        g = np.ascontiguousarray(g, dtype=np.float64).ravel()
        lam = np.zeros(self.ndof)
        st = _lib.SolveStats()
        self._check(self.L.cfdv_adjoint_solve(self.h, C.byref(o), _dp(g), _dp(lam), C.byref(st)))
        return lam, st

    def adjoint_solve_multi(self, G, restarts=20, rtol=1e-12, inner_m=30, inner_rtol=1e-4,
                            precond="nodeblock"):
        """Lambda [8][ndof] with K^T Lambda[s] = G[s] for 8 right-hand sides
        at once (G [8][ndof]): FP64 refinement per right-hand side around a
        FP32 global GMRES whose every Arnoldi step is one 8-vector product.
        Returns (Lambda, rel_resid[8], stats)."""
        G = np.ascontiguousarray(G, dtype=np.float64).reshape(8, self.ndof)
        o = _lib.SolveOpts(_lib.CFDV_SOLVE_MIXED, restarts, 30, rtol, inner_m, inner_rtol, 1, 1,
                           _PRECOND[precond])
        lam = np.zeros((8, self.ndof))
        rel = np.zeros(8)
        st = _lib.SolveStats()
        self._check(self.L.cfdv_adjoint_solve_multi(self.h, C.byref(o), _dp(G), _dp(lam), 8, _dp(rel),
                                                    C.byref(st)))
        return lam, rel, st

    def build_vjp(self, lam, du=None):
        """Gradient of Phi = lam^T (b - K du) with respect to the node state U
        and the Mach number, at the state and dt of the last build."""
        lam = np.ascontiguousarray(lam, dtype=np.float64).ravel()
        gU = np.zeros(self.ndof)
        gM = C.c_double()
        dup = None
        if du is not None:
            du = np.ascontiguousarray(du, dtype=np.float64).ravel()
            dup = _dp(du)
        self._check(self.L.cfdv_build_vjp(self.h, _dp(lam), dup, _dp(gU), C.byref(gM)))
        return gU.reshape(self.mesh.n_node, self.mesh.neqn), gM.value


def field_differences(ndim, U0, Ua, Ub):
    """Per field group (rho, momentum, energy) the difference of two states
    reached from U0, ||Ub_f - Ua_f|| / ||Ua_f - U0_f|| (relative to Ua's
    change of that field), and max |Ub_f - Ua_f|."""
    U0, Ua, Ub = (np.asarray(u, dtype=np.float64).reshape(-1, ndim + 2) for u in (U0, Ua, Ub))
    groups = {"rho": [0], "momentum": list(range(1, 1 + ndim)), "energy": [1 + ndim]}
    out = {}
    for name, cols in groups.items():
        d = np.linalg.norm(Ub[:, cols] - Ua[:, cols])
        ch = np.linalg.norm(Ua[:, cols] - U0[:, cols])
        out[name] = float(d / ch) if ch > 0 else None
        out[name + "_max_abs"] = float(np.abs(Ub[:, cols] - Ua[:, cols]).max())
    return out


class FDVGalerkin:
    """``FDVGalerkin<Element*>`` (src/FDV/FDVGalerkin.h:52-154) over the whole
    element set: ``FDVGalerkin(ctx, dt)(all)`` builds every element block on the
    device.  ApplyBC (src/FDV/Function/ApplyBC.h) is applied in the same pass."""

    def __init__(self, ctx: FDVContext, dt: float):
        self.ctx = ctx
        self.dt = dt

    def __call__(self, elements=None):
        self.ctx.build(self.dt)


class HipGMRES:
    """``GMRESbase`` shape (src/Utility/Solvers/GMRESbase.h:8-23).

    ``iterate`` solves with the element blocks built for this step;
    ``update`` applies U0 = U, U += du on the device (MPI_GMRES.h:224-301).
    mode "reference" is MPI_GMRES::iterate as written (gmresrestart
    restarts of gmresiter iterations); "fp64" and "mixed" iterate to ``rtol``.
    """

    def __init__(self, ctx: FDVContext, restart: int, m: int, mode="reference", **kw):
        self.ctx, self.restart, self.m, self.mode, self.kw = ctx, restart, m, mode, kw
        self.stats = None

    def iterate(self, elements=None):
        self.stats, _ = self.ctx.solve(self.mode, restarts=self.restart, m=self.m, **self.kw)
        return self.stats

    def update(self, nodes=None):
        self.ctx.update()
