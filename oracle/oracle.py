"""ctypes binding of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE.

Imported only by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg.  The product package (cfdv_amd) never imports this.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
up = C.POINTER(C.c_ubyte)


class Thermo(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("gamma", "gamm1", "gm1d2", "gamm3", "gm3d2",
                                          "Pr", "csuth", "Cv", "cgas", "cmach", "creyn",
                                          "Twall")] + [("adiabatic", C.c_int)]


class Prim(C.Structure):
    _fields_ = [("rho", C.c_double), ("v", C.c_double * 3), ("E", C.c_double),
                ("e", C.c_double), ("p", C.c_double), ("T", C.c_double)]


class Flow(C.Structure):
    _fields_ = [("E", C.c_double), ("T", C.c_double), ("v", C.c_double * 3),
                ("x", C.c_double * 3), ("U", C.c_double * 5),
                ("dU", (C.c_double * 3) * 5), ("dV", (C.c_double * 3) * 3),
                ("dT", C.c_double * 3), ("k", C.c_double), ("mu", C.c_double),
                ("tau", (C.c_double * 3) * 3)]


class MeshS(C.Structure):
    _fields_ = [("ndim", C.c_int), ("nnod", C.c_int), ("neqn", C.c_int),
                ("nbnod", C.c_int), ("nface", C.c_int), ("n_elem", C.c_int),
                ("n_node", C.c_int), ("conn", ip), ("x", dp), ("bc", ip),
                ("dirichlet", up), ("adap", up), ("hnm", dp)]


class GmresStats(C.Structure):
    _fields_ = [("tol", C.c_double), ("bnorm", C.c_double), ("last_resid", C.c_double),
                ("true_resid", C.c_double), ("matvecs", C.c_int), ("iters", C.c_int),
                ("restarts", C.c_int), ("breakdown", C.c_int)]


class StepOpts(C.Structure):
    _fields_ = [("cfl", C.c_double), ("restart", C.c_int), ("mdim", C.c_int),
                ("mode", C.c_int), ("rtol", C.c_double), ("maxrestart", C.c_int)]


class StepStats(C.Structure):
    _fields_ = [("dt", C.c_double), ("urms", C.c_double * 5), ("gm", GmresStats)]


def build_lib() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build_lib()
        _LIB = _bind(C.CDLL(path))
    return _LIB


class using_library:
    """Run the oracle functions on another build of cfdv_oracle.c (tests that
    plant a mistake with -DORC_MUTATE=k):  with using_library(path): ..."""

    def __init__(self, path):
        self.L = _bind(C.CDLL(path))

    def __enter__(self):
        global _LIB
        self.prev, _LIB = lib(), self.L
        return self.L

    def __exit__(self, *exc):
        global _LIB
        _LIB = self.prev
        return False


def _bind(L: C.CDLL) -> C.CDLL:
    L.orc_thermo_set.argtypes = [C.POINTER(Thermo)] + [C.c_double] * 6 + [C.c_int]
    L.orc_calc_mu.restype = C.c_double
    L.orc_calc_mu.argtypes = [C.POINTER(Thermo), C.c_double]
    L.orc_calc_k.restype = C.c_double
    L.orc_calc_k.argtypes = [C.POINTER(Thermo), C.c_double]
    L.orc_gauss.argtypes = [C.c_int, dp, dp]
    L.orc_shape3d.argtypes = [dp, dp, dp]
    L.orc_shape2d.argtypes = [dp, dp, dp]
    L.orc_jacobian3d.restype = C.c_double
    L.orc_jacobian3d.argtypes = [dp, dp, dp, dp]
    L.orc_jacobian2d.restype = C.c_double
    L.orc_jacobian2d.argtypes = [dp, dp, dp, dp]
    L.orc_tfnorm3d.restype = C.c_double
    L.orc_tfnorm3d.argtypes = [dp, dp, dp, dp, dp, dp]
    L.orc_node_unpack_all.argtypes = [C.POINTER(Thermo), C.c_int, C.c_int, C.c_int, dp,
                                      C.POINTER(Prim)]
    L.orc_calc_length.argtypes = [C.POINTER(MeshS), dp, dp]
    L.orc_cfl.argtypes = [C.POINTER(Thermo), C.POINTER(MeshS), C.POINTER(Prim), dp,
                          C.c_double, dp]
    L.orc_fdvparam.argtypes = [C.POINTER(Thermo), C.POINTER(MeshS), C.POINTER(Prim),
                               C.c_int, dp]
    L.orc_tau.argtypes = [C.c_int, C.c_double, dp, dp]
    for f in ("orc_calc_a", "orc_calc_b", "orc_calc_c"):
        getattr(L, f).argtypes = [C.POINTER(Thermo), C.c_int, C.POINTER(Flow), dp]
    L.orc_calc_G.argtypes = [C.c_int, dp, dp, dp, C.c_double, C.c_double, dp]
    L.orc_calc_F2d.argtypes = [dp, C.c_double, dp]
    L.orc_eval2dboundary.argtypes = [C.POINTER(Thermo), C.c_int, C.POINTER(Flow), dp]
    L.orc_build_element.argtypes = [C.POINTER(Thermo), C.POINTER(MeshS), dp,
                                    C.POINTER(Prim), C.c_int, C.c_double, dp, dp, dp]
    L.orc_build.argtypes = [C.POINTER(Thermo), C.POINTER(MeshS), dp, C.POINTER(Prim),
                            C.c_double, dp, dp, dp]
    L.orc_apply_bc.argtypes = [C.POINTER(Thermo), C.POINTER(MeshS), dp, dp]
    L.orc_compose_rhs.argtypes = [C.POINTER(MeshS), dp, dp]
    L.orc_matvec.argtypes = [C.POINTER(MeshS), dp, dp, dp]
    L.orc_matvec_t.argtypes = [C.POINTER(MeshS), dp, dp, dp]
    L.orc_gmres_quirk.argtypes = [C.POINTER(MeshS), dp, dp, C.c_int, C.c_int, dp,
                                  C.POINTER(GmresStats)]
    L.orc_gmres_quirk_trace.argtypes = [C.POINTER(MeshS), dp, dp, C.c_int, C.c_int, dp,
                                        C.POINTER(GmresStats), dp, dp, ip]
    L.orc_gmres_true.argtypes = [C.POINTER(MeshS), dp, dp, C.c_int, C.c_int, C.c_double,
                                 C.c_int, dp, C.POINTER(GmresStats)]
    L.orc_update.argtypes = [C.c_int, dp, dp, dp]
    L.orc_rms.argtypes = [C.c_int, C.c_int, dp, dp, dp]
    L.orc_refine_flags.argtypes = [C.POINTER(Thermo), C.POINTER(MeshS), C.POINTER(Prim), up]
    L.orc_color_greedy.argtypes = [C.POINTER(MeshS), ip]
    L.orc_color_greedy_keys.argtypes = [C.POINTER(MeshS), C.POINTER(C.c_uint64), ip]
    L.orc_step.argtypes = [C.POINTER(Thermo), C.POINTER(MeshS), dp, dp,
                           C.POINTER(StepOpts), C.POINTER(StepStats)]
    L.orc_step_par.argtypes = [C.POINTER(Thermo), C.POINTER(MeshS), dp, dp, C.POINTER(StepOpts), ip, ip,
                               C.c_int, C.c_int, C.POINTER(StepStats), dp, C.c_int]
    L.orc_psi.argtypes = [C.c_int, C.c_int, C.POINTER(Flow), dp, dp, dp, dp]
    L.orc_eval1dboundary.argtypes = [C.c_int, C.POINTER(Flow), dp, dp, dp]
    return L


def _p(a, t=dp):
    return a.ctypes.data_as(t)


def thermo(th) -> Thermo:
    t = Thermo()
    lib().orc_thermo_set(C.byref(t), th.gamma, th.mach, th.reynolds, th.prandtl, th.tinf,
                         th.twall, int(th.adiabatic))
    return t


class OMesh:
    """Keeps the numpy arrays alive behind an orc_mesh struct."""

    def __init__(self, mesh):
        self.mesh = mesh
        self.conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
        self.x = np.ascontiguousarray(mesh.x, dtype=np.float64)
        self.bc = np.ascontiguousarray(mesh.bc, dtype=np.int32)
        self.dirichlet = np.ascontiguousarray(mesh.dirichlet, dtype=np.uint8)
        self.adap = None if mesh.adap is None else np.ascontiguousarray(mesh.adap, dtype=np.uint8)
        self.hnm = None
        if mesh.hnm is not None:
            # Hnm(N, M) at N + nnod*M per element (column-major, as Tensor)
            self.hnm = np.ascontiguousarray(np.transpose(mesh.hnm, (0, 2, 1)), dtype=np.float64)
        s = MeshS()
        s.ndim, s.nnod, s.neqn = mesh.ndim, mesh.nnod, mesh.neqn
        s.nbnod, s.nface = mesh.nbnod, mesh.nface
        s.n_elem, s.n_node = mesh.n_elem, mesh.n_node
        s.conn = _p(self.conn, ip)
        s.x = _p(self.x)
        s.bc = _p(self.bc, ip)
        s.dirichlet = _p(self.dirichlet, up)
        s.adap = _p(self.adap, up) if self.adap is not None else C.cast(None, up)
        s.hnm = _p(self.hnm) if self.hnm is not None else C.cast(None, dp)
        self.s = s

    @property
    def ref(self):
        return C.byref(self.s)


def gauss2():
    w = np.zeros(2)
    z = np.zeros(2)
    lib().orc_gauss(2, _p(w), _p(z))
    return w, z


def unpack(th, mesh, U):
    prim = (Prim * mesh.n_node)()
    U = np.ascontiguousarray(U, dtype=np.float64)
    lib().orc_node_unpack_all(C.byref(thermo(th)), mesh.ndim, mesh.neqn, mesh.n_node, _p(U),
                              prim)
    return prim


def prim_arrays(prim, ndim):
    a = np.ctypeslib.as_array(C.cast(prim, dp), shape=(len(prim), 8))
    return {"rho": a[:, 0].copy(), "v": a[:, 1:1 + ndim].copy(), "E": a[:, 4].copy(),
            "e": a[:, 5].copy(), "p": a[:, 6].copy(), "T": a[:, 7].copy()}


def calc_length(mesh):
    om = OMesh(mesh)
    area = np.zeros(mesh.n_elem)
    cleng = np.zeros(mesh.n_elem)
    lib().orc_calc_length(om.ref, _p(area), _p(cleng))
    return area, cleng


def cfl(th, mesh, U, cfl_number):
    om = OMesh(mesh)
    prim = unpack(th, mesh, U)
    _, cleng = calc_length(mesh)
    dt = C.c_double(0.0)
    rc = lib().orc_cfl(C.byref(thermo(th)), om.ref, prim, _p(cleng), cfl_number, C.byref(dt))
    if rc != 0:
        raise RuntimeError("no element set dt")
    return dt.value


def fdvparam_all(th, mesh, U):
    """s1..s4 per element with the functor's carried state; returns (s, ties)."""
    om = OMesh(mesh)
    prim = unpack(th, mesh, U)
    t = thermo(th)
    s = np.zeros(4)
    out = np.zeros((mesh.n_elem, 4))
    ties = 0
    for e in range(mesh.n_elem):
        ties += lib().orc_fdvparam(C.byref(t), om.ref, prim, e, _p(s))
        out[e] = s
    return out, ties


def build(th, mesh, U, dt, apply_bc=True):
    """Element blocks R [n_e][k][k] (R[e, i, j] = R(i, j)), rhs [n_e][k], s [n_e][4]."""
    om = OMesh(mesh)
    prim = unpack(th, mesh, U)
    k = mesh.k
    Rcm = np.zeros((mesh.n_elem, k, k))     # column-major per element: [e][j][i]
    rhs = np.zeros((mesh.n_elem, k))
    s = np.zeros((mesh.n_elem, 4))
    U = np.ascontiguousarray(U, dtype=np.float64)
    t = thermo(th)
    rc = lib().orc_build(C.byref(t), om.ref, _p(U), prim, dt, _p(Rcm), _p(rhs), _p(s))
    if rc < 0:
        raise RuntimeError(f"orc_build failed: {rc}")
    if apply_bc:
        lib().orc_apply_bc(C.byref(t), om.ref, _p(Rcm), _p(rhs))
    R = np.ascontiguousarray(np.transpose(Rcm, (0, 2, 1)))
    return R, rhs, s


def _cm(R):
    return np.ascontiguousarray(np.transpose(R, (0, 2, 1)))


def compose_rhs(mesh, rhs):
    om = OMesh(mesh)
    b = np.zeros(mesh.n_node * mesh.neqn)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    lib().orc_compose_rhs(om.ref, _p(rhs), _p(b))
    return b


def matvec(mesh, R, x, transpose=False):
    om = OMesh(mesh)
    y = np.zeros(mesh.n_node * mesh.neqn)
    Rc = _cm(R)
    x = np.ascontiguousarray(x, dtype=np.float64)
    f = lib().orc_matvec_t if transpose else lib().orc_matvec
    f(om.ref, _p(Rc), _p(x), _p(y))
    return y


def gmres_quirk(mesh, R, rhs, restart, mdim):
    om = OMesh(mesh)
    du = np.zeros(mesh.n_node * mesh.neqn)
    st = GmresStats()
    Rc = _cm(R)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    rc = lib().orc_gmres_quirk(om.ref, _p(Rc), _p(rhs), restart, mdim, _p(du), C.byref(st))
    if rc:
        raise RuntimeError("orc_gmres_quirk failed")
    return du, st


def gmres_quirk_trace(mesh, R, rhs, restart, mdim):
    """gmres_quirk plus, per restart, the Hessenberg matrix as the Arnoldi
    process leaves it (H [restart][mdim+1][mdim], before the rotations), the
    rotated right-hand side g [restart][mdim+1] and the Arnoldi step count."""
    om = OMesh(mesh)
    du = np.zeros(mesh.n_node * mesh.neqn)
    st = GmresStats()
    Rc = _cm(R)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    H = np.zeros((restart, mdim, mdim + 1))
    g = np.zeros((restart, mdim + 1))
    mm = np.zeros(restart, dtype=np.int32)
    rc = lib().orc_gmres_quirk_trace(om.ref, _p(Rc), _p(rhs), restart, mdim, _p(du), C.byref(st),
                                     _p(H), _p(g), mm.ctypes.data_as(C.POINTER(C.c_int)))
    if rc:
        raise RuntimeError("orc_gmres_quirk_trace failed")
    return du, st, H.transpose(0, 2, 1).copy(), g, mm


def gmres_true(mesh, R, b, mdim, maxrestart, rtol, transpose=False):
    om = OMesh(mesh)
    du = np.zeros(mesh.n_node * mesh.neqn)
    st = GmresStats()
    Rc = _cm(R)
    b = np.ascontiguousarray(b, dtype=np.float64)
    rc = lib().orc_gmres_true(om.ref, _p(Rc), _p(b), mdim, maxrestart, rtol, int(transpose),
                              _p(du), C.byref(st))
    if rc:
        raise RuntimeError("orc_gmres_true failed")
    return du, st


def refine_flags(th, mesh, U):
    om = OMesh(mesh)
    prim = unpack(th, mesh, U)
    f = np.zeros(mesh.n_elem, dtype=np.uint8)
    lib().orc_refine_flags(C.byref(thermo(th)), om.ref, prim, _p(f, up))
    return f


def color_greedy(mesh, keys=None):
    """First-fit greedy coloring in element order, or in increasing key order."""
    om = OMesh(mesh)
    c = np.zeros(mesh.n_elem, dtype=np.int32)
    if keys is None:
        n = lib().orc_color_greedy(om.ref, _p(c, ip))
    else:
        k = np.ascontiguousarray(keys, dtype=np.uint64)
        n = lib().orc_color_greedy_keys(om.ref, k.ctypes.data_as(C.POINTER(C.c_uint64)), _p(c, ip))
    return c, n


def jp_keys(n_elem):
    """Priority keys of the device coloring: splitmix64(e) with the low 27
    bits replaced by e (unique, hash-ordered)."""
    e = np.arange(n_elem, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = e + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return ((z >> np.uint64(27)) << np.uint64(27)) | e


def step(th, mesh, U, *, cfl=0.5, restart=1, mdim=20, mode=0, rtol=1e-8, maxrestart=50):
    """One FDV step in place on a copy; returns (U_new, U0, stats)."""
    om = OMesh(mesh)
    U = np.array(U, dtype=np.float64, copy=True)
    U0 = np.zeros_like(U)
    o = StepOpts(cfl, restart, mdim, mode, rtol, maxrestart)
    st = StepStats()
    rc = lib().orc_step(C.byref(thermo(th)), om.ref, _p(U), _p(U0), C.byref(o), C.byref(st))
    if rc:
        raise RuntimeError(f"orc_step failed: {rc}")
    return U, U0, st


def step_par(th, mesh, U, *, threads, cfl=0.5, mdim=30, rtol=1e-8, maxrestart=50, colors=None, verbose=False):
    """One FDV step (true GMRES(mdim) to rtol) on ``threads`` OpenMP threads,
    the timing variant of step(mode=1); returns (U_new, U0, stats, phases)
    with phases the seconds of set-up, build, ApplyBC + rhs, GMRES + update."""
    om = OMesh(mesh)
    if colors is None:
        colors, _ = color_greedy(mesh)
    colors = np.asarray(colors, dtype=np.int32)
    nc = int(colors.max()) + 1
    celem = np.argsort(colors, kind="stable").astype(np.int32)
    cptr = np.zeros(nc + 1, dtype=np.int32)
    cptr[1:] = np.cumsum(np.bincount(colors, minlength=nc))
    U = np.array(U, dtype=np.float64, copy=True)
    U0 = np.zeros_like(U)
    o = StepOpts(cfl, 1, mdim, 1, rtol, maxrestart)
    st = StepStats()
    ph = np.zeros(4)
    rc = lib().orc_step_par(C.byref(thermo(th)), om.ref, _p(U), _p(U0), C.byref(o), _p(cptr, ip),
                            _p(celem, ip), nc, int(threads), C.byref(st), _p(ph), int(verbose))
    if rc:
        raise RuntimeError(f"orc_step_par failed: {rc}")
    return U, U0, st, ph


def flow_struct(E, T, v, x, U, dU, dV, dT, k, mu, tau=None) -> Flow:
    f = Flow()
    f.E, f.T, f.k, f.mu = E, T, k, mu
    for i in range(len(v)):
        f.v[i] = v[i]
    for i in range(len(x)):
        f.x[i] = x[i]
    for i in range(len(U)):
        f.U[i] = U[i]
    dU = np.asarray(dU)
    for i in range(dU.shape[0]):
        for j in range(dU.shape[1]):
            f.dU[i][j] = dU[i, j]
    dV = np.asarray(dV)
    for i in range(dV.shape[0]):
        for j in range(dV.shape[1]):
            f.dV[i][j] = dV[i, j]
    for i in range(len(dT)):
        f.dT[i] = dT[i]
    if tau is not None:
        tau = np.asarray(tau)
        for i in range(tau.shape[0]):
            for j in range(tau.shape[1]):
                f.tau[i][j] = tau[i, j]
    return f


def flow_tau(f: Flow):
    return np.array([[f.tau[i][j] for j in range(3)] for i in range(3)])


def calc_abc(th, ndim, f: Flow):
    neqn = ndim + 2
    a = np.zeros(ndim * neqn * neqn)
    b = np.zeros(ndim * neqn * neqn)
    c = np.zeros(ndim * ndim * neqn * neqn)
    t = thermo(th)
    lib().orc_calc_a(C.byref(t), ndim, C.byref(f), _p(a))
    lib().orc_calc_b(C.byref(t), ndim, C.byref(f), _p(b))
    lib().orc_calc_c(C.byref(t), ndim, C.byref(f), _p(c))
    return (a.reshape(ndim, neqn, neqn), b.reshape(ndim, neqn, neqn),
            c.reshape(ndim, ndim, neqn, neqn))


def calc_G(ndim, tau3, v, dT, k, creyn):
    """G [ndim][neqn] (NavierStokes3D::calc_G / NavierStokes2D) from the
    Gauss-point tau, dT, k and a node's velocity"""
    neqn = ndim + 2
    t = np.ascontiguousarray(tau3, dtype=np.float64).reshape(3, 3)
    vv = np.zeros(3)
    vv[:ndim] = v
    dd = np.zeros(3)
    dd[:ndim] = dT
    G = np.zeros(ndim * neqn)
    lib().orc_calc_G(ndim, _p(t), _p(vv), _p(dd), float(k), float(creyn), _p(G))
    return G.reshape(ndim, neqn)


def psi(nd, f: Flow, xn, zeta, dphi):
    """Psi [nnod] of one Gauss point (orc_psi)"""
    xn = np.ascontiguousarray(xn, dtype=np.float64)
    z = np.ascontiguousarray(zeta, dtype=np.float64)
    d = np.ascontiguousarray(dphi, dtype=np.float64)
    nn = xn.shape[0]
    out = np.zeros(nn)
    lib().orc_psi(nd, nn, C.byref(f), _p(xn), _p(z), _p(d), _p(out))
    return out


def eval1dboundary(bc, f: Flow, xn0, xn1, en):
    """orc_eval1dboundary on f in place; returns its status (0 or -1)"""
    return lib().orc_eval1dboundary(bc, C.byref(f), _p(np.ascontiguousarray(xn0, dtype=np.float64)),
                                    _p(np.ascontiguousarray(xn1, dtype=np.float64)),
                                    _p(np.ascontiguousarray(en, dtype=np.float64)))


def tau(ndim, mu, dV):
    dv = np.zeros((3, 3))
    dv[:ndim, :ndim] = dV
    out = np.zeros((3, 3))
    lib().orc_tau(ndim, mu, _p(dv), _p(out))
    return out[:ndim, :ndim]
