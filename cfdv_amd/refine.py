"""Adaptive refinement on the host, through the C ABI (cfdv_refine_*).

Mirrors the reference's refinement entry (src/main.cpp:143-151)::

    AdaptiveRefineList<Element,Node>(elelist, refine_level, elements, nodes, nnod, nface);
    MeshRefine2D<Element,Node>(elements, nodes, elelist, refine_level, ndim, neqn, nnod, iter, nface);

``AdaptiveRefineList(mesh, U, flags, levels)`` returns ``(elelist,
refine_level)`` and ``MeshRefine2D(...)`` returns the refined mesh, node
states and element levels.  The flags are the s1 > 0.8 pass
(``FDVContext.refine_flags`` on the GPU).  Element and node numbering is
append order, exactly as in the reference, so a refined mesh reproduces the
reference's numbering.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from .mesh import Mesh

PARITY = _lib.CFDV_REFINE_PARITY
MODE_3D = _lib.CFDV_REFINE_3D


class RefineError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"cfdv refine error {code}: {msg}")
        self.code = code


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class RefinableMesh:
    """A ``cfdv_refine`` object: the element and node vectors that the
    reference's refinement reads and appends to."""

    def __init__(self, mesh: Mesh, U: np.ndarray, U0: Optional[np.ndarray] = None,
                 refine_level: Optional[np.ndarray] = None):
        self.L = _lib.load()
        self.ndim, self.nnod, self.neqn = mesh.ndim, mesh.nnod, mesh.neqn
        self.nface = mesh.nface
        conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
        x = np.ascontiguousarray(mesh.x, dtype=np.float64)
        bc = np.ascontiguousarray(mesh.bc, dtype=np.int32)
        d = np.ascontiguousarray(mesh.dirichlet, dtype=np.uint8)
        adap = None if mesh.adap is None else np.ascontiguousarray(mesh.adap, dtype=np.uint8)
        hnm = None
        if mesh.hnm is not None:
            hnm = np.ascontiguousarray(np.transpose(mesh.hnm, (0, 2, 1)), dtype=np.float64)
        m = _lib.Mesh(mesh.ndim, mesh.nnod, mesh.neqn, mesh.n_elem, mesh.n_node,
                      _ptr(conn, C.c_int32), _ptr(x, C.c_double), _ptr(bc, C.c_int32),
                      _ptr(d, C.c_uint8),
                      _ptr(adap, C.c_uint8) if adap is not None else C.cast(None, C.POINTER(C.c_uint8)),
                      _ptr(hnm, C.c_double) if hnm is not None else C.cast(None, C.POINTER(C.c_double)))
        Uc = np.ascontiguousarray(U, dtype=np.float64)
        U0c = None if U0 is None else np.ascontiguousarray(U0, dtype=np.float64)
        lv = None if refine_level is None else np.ascontiguousarray(refine_level, dtype=np.int32)
        h = C.c_void_p()
        rc = self.L.cfdv_refine_create(C.byref(m), _ptr(Uc, C.c_double),
                                       _ptr(U0c, C.c_double) if U0c is not None else None,
                                       _ptr(lv, C.c_int32) if lv is not None else None, C.byref(h))
        if rc != 0:
            raise RefineError(rc, "cfdv_refine_create rejected the mesh")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            self.L.cfdv_refine_destroy(self.h)
            self.h = None

    def _check(self, rc):
        if rc != 0:
            raise RefineError(rc, self.L.cfdv_refine_last_error(self.h).decode())

    def sizes(self):
        ne, nn, ml = C.c_int64(), C.c_int64(), C.c_int32()
        self._check(self.L.cfdv_refine_size(self.h, C.byref(ne), C.byref(nn), C.byref(ml)))
        return ne.value, nn.value, ml.value

    def refine_list(self, flags: np.ndarray):
        """AdaptiveRefineList (src/adap/RefineListAdaptive.h:8-66)"""
        ne, _, ml = self.sizes()
        f = np.ascontiguousarray(flags, dtype=np.uint8)
        if f.shape != (ne,):
            raise ValueError("flags must have one entry per element")
        el = np.zeros(ne, dtype=np.int32)
        rl = np.zeros(ml + 2, dtype=np.int32)
        nl, nr = C.c_int32(), C.c_int32()
        self._check(self.L.cfdv_refine_list(self.h, _ptr(f, C.c_uint8), _ptr(el, C.c_int32), C.byref(nl),
                                            _ptr(rl, C.c_int32), C.byref(nr)))
        return el[:nl.value].copy(), rl[:nr.value].copy()

    def apply(self, elelist, refine_level, mode: int = PARITY, tol: float = 0.0):
        """MeshRefine2D (src/adap/MeshRefine2D.h:41-75)"""
        el = np.ascontiguousarray(elelist, dtype=np.int32)
        rl = np.ascontiguousarray(refine_level, dtype=np.int32)
        self._check(self.L.cfdv_refine_apply(self.h, _ptr(el, C.c_int32), len(el), _ptr(rl, C.c_int32),
                                             len(rl), mode, tol))

    def get(self):
        """(mesh, U, U0, refine_level) of the current state"""
        ne, nn, _ = self.sizes()
        conn = np.zeros((ne, self.nnod), dtype=np.int32)
        bc = np.zeros((ne, self.nface), dtype=np.int32)
        x = np.zeros((nn, self.ndim))
        U = np.zeros((nn, self.neqn))
        U0 = np.zeros((nn, self.neqn))
        d = np.zeros((nn, self.neqn), dtype=np.uint8)
        adap = np.zeros(ne, dtype=np.uint8)
        hcm = np.zeros((ne, self.nnod, self.nnod))
        lv = np.zeros(ne, dtype=np.int32)
        self._check(self.L.cfdv_refine_get(self.h, _ptr(conn, C.c_int32), _ptr(bc, C.c_int32),
                                           _ptr(x, C.c_double), _ptr(U, C.c_double), _ptr(U0, C.c_double),
                                           _ptr(d, C.c_uint8), _ptr(adap, C.c_uint8), _ptr(hcm, C.c_double),
                                           _ptr(lv, C.c_int32)))
        hnm = np.ascontiguousarray(np.transpose(hcm, (0, 2, 1)))   # [e][N][M]
        mesh = Mesh(self.ndim, self.nnod, self.neqn, conn, x, bc, d,
                    adap=adap if adap.any() else None, hnm=hnm if adap.any() else None)
        return mesh, U, U0, lv


def AdaptiveRefineList(mesh: Mesh, U: np.ndarray, flags: np.ndarray, refine_level=None):
    """(elelist, refine_level breakpoints) for the flagged elements"""
    return RefinableMesh(mesh, U, refine_level=refine_level).refine_list(flags)


def MeshRefine2D(mesh: Mesh, U: np.ndarray, elelist, breaks, U0=None, refine_level=None,
                 mode: int = PARITY, tol: float = 0.0):
    """The refined (mesh, U, U0, refine_level)"""
    r = RefinableMesh(mesh, U, U0, refine_level)
    r.apply(elelist, breaks, mode, tol)
    return r.get()


def refine_pass(mesh: Mesh, U: np.ndarray, flags: np.ndarray, U0=None, refine_level=None,
                mode: int = PARITY, tol: float = 0.0):
    """One --adap pass (main.cpp:143-151): list, then refine.  Returns
    (mesh, U, U0, refine_level, elelist, breaks)."""
    r = RefinableMesh(mesh, U, U0, refine_level)
    el, rl = r.refine_list(flags)
    r.apply(el, rl, mode, tol)
    m, U1, U01, lv = r.get()
    return m, U1, U01, lv, el, rl


def sensor_refine(mesh: Mesh, thermo, U: np.ndarray, levels: int, mode: int = MODE_3D, tol: float = 1e-9,
                  keep_passes: bool = False):
    """``levels`` --adap passes as a run of the reference would make them
    (main.cpp:143-157), each on the mesh and state the previous one left:
    the s1 > 0.8 flags of the current state on the GPU
    (FDVContext.refine_flags, RefineListAdaptive.h:28-38), then
    AdaptiveRefineList and MeshRefine2D on the host.  ``tol`` is the
    repeated-node distance (the reference's absolute 1e-3 merges distinct
    nodes once cells are thinner than that; 1e-9 still catches the nodes
    that neighbouring refinements interpolate to the same point).

    Returns (mesh, U, info): info holds per pass the flagged and refined
    element counts, the element and node counts after it, the GPU flag-pass
    time and the host list/apply times; with keep_passes also each pass's
    input (mesh, U) and flags under info["passes"]."""
    import time
    from .fdv import FDVContext
    cur, Uc, U0c, lv = mesh, U, None, None
    info = {"elements": [int(mesh.n_elem)], "nodes": [int(mesh.n_node)], "flagged": [], "refined": [],
            "flags_ms": [], "host_list_s": [], "host_apply_s": []}
    passes = []
    for _ in range(levels):
        ctx = FDVContext(cur, thermo, Uc, blocks=("fp32",))
        t0 = time.perf_counter()
        flags, _, _ = ctx.refine_flags()
        info["flags_ms"].append((time.perf_counter() - t0) * 1e3)
        ctx.close()
        if keep_passes:
            passes.append((cur, Uc, flags))
        r = RefinableMesh(cur, Uc, U0c, lv)
        t0 = time.perf_counter()
        el, rl = r.refine_list(flags)
        t1 = time.perf_counter()
        r.apply(el, rl, mode, tol)
        t2 = time.perf_counter()
        cur, Uc, U0c, lv = r.get()
        del r
        info["flagged"].append(int(flags.sum()))
        info["refined"].append(int(len(el)))
        info["host_list_s"].append(t1 - t0)
        info["host_apply_s"].append(t2 - t1)
        info["elements"].append(int(cur.n_elem))
        info["nodes"].append(int(cur.n_node))
    info["adap_elements"] = int(cur.adap.sum()) if cur.adap is not None else 0
    info["levels"] = np.bincount(lv).tolist() if lv is not None else [int(mesh.n_elem)]
    if keep_passes:
        info["passes"] = passes
    return cur, Uc, info
