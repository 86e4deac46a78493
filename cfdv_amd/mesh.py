"""Meshes, decks and initial states for the FDV time step.

A ``Mesh`` carries exactly what the reference's ``LoadBinaryData`` builds
(src/Utility/IO/ASCIItoBinary/LoadBinary.h): connectivity, coordinates, the
per-face boundary codes, the node Dirichlet flags that ``add_face`` derives
from those codes, and the optional hanging-node matrices ``Hnm``.

Structured generators write the same element and face numbering as the
reference's 3D decks (node 0..7 of a hex as in ``get_nodes_from_face_3d``,
src/FiniteElement/get_face.h:106-165; faces 0 z-min, 1 y-min, 2 x-max,
3 y-max, 4 x-min, 5 z-max) and of the 2D decks (face f = nodes f, f+1).
"""
from __future__ import annotations

import dataclasses
import os
from typing import Optional

import numpy as np

# face -> local nodes, get_face.h:80-89 (2D) and :106-165 (3D)
FACE_NODES_3D = np.array([[3, 2, 1, 0], [0, 1, 5, 4], [1, 2, 6, 5],
                          [2, 3, 7, 6], [3, 0, 4, 7], [6, 7, 4, 5]], dtype=np.int32)
FACE_NODES_2D = np.array([[0, 1], [1, 2], [2, 3], [3, 0]], dtype=np.int32)


@dataclasses.dataclass
class Thermo:
    """Inputs of ``Thermo::set`` (src/FDV/Thermo.h:22-33), the ``aero`` deck keys."""
    gamma: float = 1.4
    mach: float = 3.0
    reynolds: float = 1000.0
    prandtl: float = 0.72
    tinf: float = 216.6666
    twall: float = 1.0
    adiabatic: bool = True

    @property
    def Cv(self) -> float:
        return 1.0 / self.gamma / (self.gamma - 1.0) / (self.mach * self.mach)

    @property
    def cgas(self) -> float:
        return 1.0 / self.gamma / (self.mach * self.mach)


@dataclasses.dataclass
class Mesh:
    ndim: int
    nnod: int
    neqn: int
    conn: np.ndarray            # int32 [n_elem][nnod]
    x: np.ndarray               # f64 [n_node][ndim]
    bc: np.ndarray              # int32 [n_elem][nface]; 0 = no face
    dirichlet: np.ndarray       # uint8 [n_node][neqn]
    adap: Optional[np.ndarray] = None   # uint8 [n_elem]
    hnm: Optional[np.ndarray] = None    # f64 [n_elem][nnod][nnod], Hnm[e, N, M]
    shape: Optional[tuple] = None       # structured (nx, ny[, nz]) if generated
    refine_level: Optional[np.ndarray] = None   # int32 [n_elem] (Element::refine_level)

    @property
    def nbnod(self) -> int:
        return 4 if self.ndim == 3 else 2

    @property
    def nface(self) -> int:
        return 6 if self.ndim == 3 else 4

    @property
    def n_elem(self) -> int:
        return int(self.conn.shape[0])

    @property
    def n_node(self) -> int:
        return int(self.x.shape[0])

    @property
    def k(self) -> int:
        return self.nnod * self.neqn

    def face_nodes(self) -> np.ndarray:
        return FACE_NODES_3D if self.ndim == 3 else FACE_NODES_2D


def dirichlet_flags(ndim: int, neqn: int, n_node: int, conn: np.ndarray,
                    bc: np.ndarray, adiabatic: bool) -> np.ndarray:
    """Node Dirichlet flags set by ``add_face`` (LoadBinary.h:264-318).

    bc -1/-12: every equation; bc 2/9: momentum ``ndim``; bc 22: momentum 2;
    walls 1/11/21/31/41/4: all momenta, plus energy unless adiabatic.
    """
    fn = FACE_NODES_3D if ndim == 3 else FACE_NODES_2D
    d = np.zeros((n_node, neqn), dtype=np.uint8)
    for f in range(fn.shape[0]):
        codes = bc[:, f]
        for code in np.unique(codes):
            if code == 0:
                continue
            sel = conn[codes == code][:, fn[f]].ravel()
            if code in (-1, -12):
                d[sel, :] = 1
            elif code in (2, 9):
                d[sel, ndim] = 1
            elif code == 22:
                d[sel, 2] = 1
            elif code in (1, 11, 21, 31, 41, 4):
                d[sel, 1:1 + ndim] = 1
                if not adiabatic:
                    d[sel, 1 + ndim] = 1
    return d


def conservative(rho: np.ndarray, v: np.ndarray, T: np.ndarray, Cv: float) -> np.ndarray:
    """Primitive -> conservative map of ``loadNodes`` (LoadBinary.h:179-196)."""
    n, ndim = v.shape
    U = np.empty((n, ndim + 2))
    U[:, 0] = rho
    U[:, 1:1 + ndim] = v
    U[:, ndim + 1] = T * Cv
    vs = np.zeros(n)
    for j in range(ndim):
        vs = vs + U[:, j + 1] * U[:, j + 1]
        U[:, j + 1] = U[:, j + 1] * U[:, 0]
    U[:, ndim + 1] = U[:, ndim + 1] + 0.5 * vs
    U[:, ndim + 1] = U[:, ndim + 1] * U[:, 0]
    return U


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser, vectorised over uint64."""
    with np.errstate(over="ignore"):
        z = (x.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def density_noise(n_node: int, seed: int, amp: float) -> np.ndarray:
    """Deterministic U[-amp, amp] per node from splitmix64(seed ^ node id)."""
    ids = np.arange(n_node, dtype=np.uint64) ^ np.uint64(seed)
    u = (splitmix64(ids) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return amp * (2.0 * u - 1.0)


# ---------------------------------------------------------------------------
# structured generators
# ---------------------------------------------------------------------------
def _hex_conn(nx: int, ny: int, nz: int) -> np.ndarray:
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i = i.transpose(2, 1, 0).ravel()
    j = j.transpose(2, 1, 0).ravel()
    k = k.transpose(2, 1, 0).ravel()
    px, pxy = nx + 1, (nx + 1) * (ny + 1)
    n0 = i + px * j + pxy * k
    conn = np.stack([n0, n0 + 1, n0 + 1 + px, n0 + px,
                     n0 + pxy, n0 + pxy + 1, n0 + pxy + 1 + px, n0 + pxy + px], axis=1)
    return conn.astype(np.int32)


def _stretch(n: int, ratio: float) -> np.ndarray:
    if ratio == 1.0:
        return np.linspace(0.0, 1.0, n + 1)
    s = (ratio ** np.arange(n + 1) - 1.0) / (ratio ** n - 1.0)
    s[-1] = 1.0
    return s


def box_3d(nx: int, ny: int, nz: int, *, lx=1.0, ly=0.5, lz=0.25, ratio=1.0,
           ramp_start: Optional[float] = None, ramp_deg: float = 15.0,
           skew: float = 0.0) -> tuple:
    """Node coordinates and connectivity of a structured hex8 block.

    ``ratio`` stretches the z (wall-normal) spacing geometrically; with
    ``ramp_start`` the bottom surface rises at ``ramp_deg`` after that x.
    ``skew`` tilts x-lines by a linear function of y (non-affine elements).
    """
    xs = np.linspace(0.0, lx, nx + 1)
    ys = np.linspace(0.0, ly, ny + 1)
    s = _stretch(nz, ratio)
    X = np.empty((nz + 1, ny + 1, nx + 1, 3))
    X[..., 0] = xs[None, None, :] + skew * ys[None, :, None] * s[:, None, None]
    X[..., 1] = ys[None, :, None]
    zb = np.zeros(nx + 1)
    if ramp_start is not None:
        zb = np.where(xs > ramp_start, (xs - ramp_start) * np.tan(np.radians(ramp_deg)), 0.0)
    X[..., 2] = zb[None, None, :] + (lz - zb[None, None, :]) * s[:, None, None]
    return X.reshape(-1, 3), _hex_conn(nx, ny, nz)


def _face_bc_3d(nx, ny, nz, *, zmin, ymin, xmax, ymax, xmin, zmax):
    """bc array [n_elem][6]; each argument is a code or an (nx,) array for zmin."""
    ne = nx * ny * nz
    bc = np.zeros((ne, 6), dtype=np.int32)
    e = np.arange(ne)
    i = e % nx
    j = (e // nx) % ny
    k = e // (nx * ny)
    zcodes = np.broadcast_to(np.asarray(zmin, dtype=np.int32), (nx,))
    bc[k == 0, 0] = zcodes[i[k == 0]]
    bc[j == 0, 1] = ymin
    bc[i == nx - 1, 2] = xmax
    bc[j == ny - 1, 3] = ymax
    bc[i == 0, 4] = xmin
    bc[k == nz - 1, 5] = zmax
    return bc


def flat_plate(nx=128, ny=64, nz=32, *, ratio=1.08, mach=3.0, reynolds=1e5,
               twall_ratio=2.0, noise=1e-3, seed=20261203, leading_edge=0.125,
               adiabatic=False) -> tuple:
    """Supersonic laminar flat plate on a hex8 block (x streamwise, z normal).

    Inflow -1 at x-min, outflow 6 at x-max, top 7, spanwise 22 (y-min) and 32
    (y-max); bottom is symmetry 2 ahead of ``leading_edge`` and wall 1 after,
    as in the reference's 3D plate decks.  Isothermal wall at
    ``twall_ratio``; uniform freestream plus seeded density noise.
    """
    x, conn = box_3d(nx, ny, nz, lx=1.0, ly=0.5, lz=0.25, ratio=ratio)
    xc = 0.5 * (np.linspace(0, 1, nx + 1)[:-1] + np.linspace(0, 1, nx + 1)[1:])
    zmin = np.where(xc < leading_edge, 2, 1)
    bc = _face_bc_3d(nx, ny, nz, zmin=zmin, ymin=22, xmax=6, ymax=32, xmin=-1, zmax=7)
    th = Thermo(gamma=1.4, mach=mach, reynolds=reynolds, prandtl=0.72, tinf=216.6666,
                twall=twall_ratio, adiabatic=adiabatic)
    d = dirichlet_flags(3, 5, x.shape[0], conn, bc, adiabatic)
    nn = x.shape[0]
    rho = 1.0 + density_noise(nn, seed, noise) if noise else np.ones(nn)
    v = np.zeros((nn, 3))
    v[:, 0] = 1.0
    T = np.ones(nn)
    wall = np.zeros(nn, dtype=bool)
    wall[conn[bc[:, 0] == 1][:, FACE_NODES_3D[0]].ravel()] = True
    v[wall] = 0.0
    T[wall] = twall_ratio
    U = conservative(rho, v, T, th.Cv)
    mesh = Mesh(3, 8, 5, conn, x, bc, d, shape=(nx, ny, nz))
    return mesh, th, U


def compression_ramp(nx=256, ny=128, nz=64, *, ratio=1.05, mach=6.0, reynolds=5e5,
                     ramp_deg=15.0, corner=0.5, noise=0.0, seed=20261203) -> tuple:
    """Hypersonic compression ramp, adiabatic wall (x streamwise, z normal).

    Inflow -1, exit 8, top 7, spanwise symmetry 22 on both sides; bottom wall
    is 1 ahead of the corner and 4 (compression corner) on the ramp.
    """
    x, conn = box_3d(nx, ny, nz, lx=1.0, ly=0.5, lz=0.5, ratio=ratio,
                     ramp_start=corner, ramp_deg=ramp_deg)
    xs = np.linspace(0, 1, nx + 1)
    xc = 0.5 * (xs[:-1] + xs[1:])
    zmin = np.where(xc < corner, 1, 4)
    bc = _face_bc_3d(nx, ny, nz, zmin=zmin, ymin=22, xmax=8, ymax=22, xmin=-1, zmax=7)
    th = Thermo(gamma=1.4, mach=mach, reynolds=reynolds, prandtl=0.72, tinf=216.6666,
                twall=1.0, adiabatic=True)
    d = dirichlet_flags(3, 5, x.shape[0], conn, bc, True)
    nn = x.shape[0]
    rho = 1.0 + density_noise(nn, seed, noise) if noise else np.ones(nn)
    v = np.zeros((nn, 3))
    v[:, 0] = 1.0
    T = np.ones(nn)
    wall = np.zeros(nn, dtype=bool)
    wall[conn[bc[:, 0] != 0][:, FACE_NODES_3D[0]].ravel()] = True
    v[wall] = 0.0
    U = conservative(rho, v, T, th.Cv)
    mesh = Mesh(3, 8, 5, conn, x, bc, d, shape=(nx, ny, nz))
    return mesh, th, U


def cylinder_ogrid(nt=256, nr=128, nz=32, *, r0=0.5, r1=10.0, lz=1.0, ratio=1.04, mach=2.0,
                   reynolds=1e4, twall_ratio=1.8, noise=0.0, seed=20261203) -> tuple:
    """Cylinder in crossflow on a hex8 O-grid (x streamwise, z spanwise).

    nt cells around the cylinder (periodic: no seam nodes, the last ring of
    cells closes onto the first), nr cells radially from the wall r0 to the
    farfield r1 (geometrically stretched by ``ratio``), nz spanwise over lz.
    Element axes: xi radial, eta circumferential, zeta spanwise, which keeps
    det J > 0.  Isothermal wall 1 at ``twall_ratio`` on the cylinder (faces
    xi = -1), farfield -1 on the outer faces whose centre has x < 0 and
    exit 8 on the rest, spanwise symmetry 2 (z-normal faces) at both ends;
    uniform freestream u = 1 with the wall at rest.  Nodes are numbered
    circumferential index fastest, then radial, then spanwise.
    """
    th_ = 2.0 * np.pi * np.arange(nt) / nt
    rs = r0 + (r1 - r0) * _stretch(nr, ratio)
    zs = np.linspace(0.0, lz, nz + 1)
    X = np.empty((nz + 1, nr + 1, nt, 3))
    X[..., 0] = rs[None, :, None] * np.cos(th_)[None, None, :]
    X[..., 1] = rs[None, :, None] * np.sin(th_)[None, None, :]
    X[..., 2] = zs[:, None, None]
    x = X.reshape(-1, 3)

    def node(i, j, k):
        return (i % nt) + nt * (j + (nr + 1) * k)

    j, i, k = np.meshgrid(np.arange(nr), np.arange(nt), np.arange(nz), indexing="ij")
    j = j.transpose(2, 1, 0).ravel()      # element order: radial fastest
    i = i.transpose(2, 1, 0).ravel()
    k = k.transpose(2, 1, 0).ravel()
    conn = np.stack([node(i, j, k), node(i, j + 1, k), node(i + 1, j + 1, k), node(i + 1, j, k),
                     node(i, j, k + 1), node(i, j + 1, k + 1), node(i + 1, j + 1, k + 1),
                     node(i + 1, j, k + 1)], axis=1).astype(np.int32)
    ne = conn.shape[0]
    bc = np.zeros((ne, 6), dtype=np.int32)
    bc[k == 0, 0] = 2
    bc[k == nz - 1, 5] = 2
    bc[j == 0, 4] = 1
    outer = j == nr - 1
    xc = 0.5 * (np.cos(th_[i]) + np.cos(th_[(i + 1) % nt])) * r1
    bc[outer, 2] = np.where(xc[outer] < 0.0, -1, 8)
    th = Thermo(gamma=1.4, mach=mach, reynolds=reynolds, prandtl=0.72, tinf=216.6666,
                twall=twall_ratio, adiabatic=False)
    d = dirichlet_flags(3, 5, x.shape[0], conn, bc, False)
    nn = x.shape[0]
    rho = 1.0 + density_noise(nn, seed, noise) if noise else np.ones(nn)
    v = np.zeros((nn, 3))
    v[:, 0] = 1.0
    T = np.ones(nn)
    wall = np.zeros(nn, dtype=bool)
    wall[conn[bc[:, 4] == 1][:, FACE_NODES_3D[4]].ravel()] = True
    v[wall] = 0.0
    T[wall] = twall_ratio
    U = conservative(rho, v, T, th.Cv)
    mesh = Mesh(3, 8, 5, conn, x, bc, d, shape=(nr, nt, nz))
    return mesh, th, U


def lid_cavity(n=32, *, mach=0.1, reynolds=100.0) -> tuple:
    """2D lid-driven cavity on the unit square, uniform n x n quad4.

    Lid bc -1 on the top faces with u = 1 on interior top nodes (corners at
    rest), walls bc 1 elsewhere, adiabatic; at rest with rho = 1, T = 1.
    """
    px = n + 1
    xs = np.linspace(0.0, 1.0, px)
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    x = np.stack([X.ravel(), Y.ravel()], axis=1)
    e = np.arange(n * n)
    i = e % n
    j = e // n
    n0 = i + px * j
    conn = np.stack([n0, n0 + 1, n0 + 1 + px, n0 + px], axis=1).astype(np.int32)
    bc = np.zeros((n * n, 4), dtype=np.int32)
    bc[j == 0, 0] = 1
    bc[i == n - 1, 1] = 1
    bc[j == n - 1, 2] = -1
    bc[i == 0, 3] = 1
    th = Thermo(gamma=1.4, mach=mach, reynolds=reynolds, prandtl=0.72, tinf=216.6666,
                twall=1.0, adiabatic=True)
    d = dirichlet_flags(2, 4, x.shape[0], conn, bc, True)
    nn = x.shape[0]
    v = np.zeros((nn, 2))
    top = np.isclose(x[:, 1], 1.0) & (x[:, 0] > 0.0) & (x[:, 0] < 1.0)
    v[top, 0] = 1.0
    U = conservative(np.ones(nn), v, np.ones(nn), th.Cv)
    mesh = Mesh(2, 4, 4, conn, x, bc, d, shape=(n, n))
    return mesh, th, U


def mixed_bc_box(nx=6, ny=5, nz=4, *, adiabatic=False, seed=7, skew=0.15,
                 ratio=1.1) -> tuple:
    """Small skewed, stretched hex8 block touching many bc codes.

    Used by parity tests: bottom alternates 1/2/4/21/11/31 along x, and the
    other faces carry -1, 6, 7, 22, 32, 8, 12 so that every 3D boundary edit
    and the ApplyBC energy exception are exercised.  The state is a smooth
    non-uniform field plus noise.
    """
    x, conn = box_3d(nx, ny, nz, lx=0.6, ly=0.4, lz=0.3, ratio=ratio, skew=skew)
    cycle = np.array([1, 2, 4, 21, 11, 31], dtype=np.int32)
    zmin = cycle[np.arange(nx) % cycle.size]
    bc = _face_bc_3d(nx, ny, nz, zmin=zmin, ymin=22, xmax=6, ymax=32, xmin=-1, zmax=7)
    ne = conn.shape[0]
    e = np.arange(ne)
    k = e // (nx * ny)
    i = e % nx
    # a few top faces use 8 and 12, a few y-max faces use 77
    bc[(k == nz - 1) & (i == 1), 5] = 8
    bc[(k == nz - 1) & (i == 2), 5] = 12
    th = Thermo(gamma=1.4, mach=3.0, reynolds=1000.0, prandtl=0.72, tinf=216.6666,
                twall=2.8, adiabatic=adiabatic)
    d = dirichlet_flags(3, 5, x.shape[0], conn, bc, adiabatic)
    rng = np.random.default_rng(seed)
    nn = x.shape[0]
    rho = 1.0 + 0.2 * np.sin(3 * x[:, 0]) * np.cos(2 * x[:, 1]) + 0.02 * rng.standard_normal(nn)
    v = np.stack([1.0 + 0.3 * x[:, 2] - 0.2 * x[:, 1],
                  0.1 * np.sin(5 * x[:, 0]),
                  0.05 * np.cos(4 * x[:, 1] + x[:, 2])], axis=1) + 0.02 * rng.standard_normal((nn, 3))
    T = 1.0 + 0.5 * x[:, 2] + 0.1 * np.cos(3 * x[:, 0]) + 0.01 * rng.standard_normal(nn)
    U = conservative(rho, v, T, th.Cv)
    mesh = Mesh(3, 8, 5, conn, x, bc, d, shape=(nx, ny, nz))
    return mesh, th, U


# ---------------------------------------------------------------------------
# reference deck I/O (binarynodes / binaryelements / control / aero)
# ---------------------------------------------------------------------------
def parse_keyfile(path: str) -> dict:
    """``key = value`` files; lines starting with '!' or '#' are skipped
    (AsciiLoad.h:166-190 tolerates them; boost::program_options would not)."""
    out = {}
    with open(path) as fh:
        for line in fh:
            line = line.strip()
            if not line or line[0] in "!#" or "=" not in line:
                continue
            k, v = (t.strip() for t in line.split("=", 1))
            out[k] = v
    return out


def read_deck(path: str, node_range=None) -> tuple:
    """Read a reference deck directory (LoadBinary.h:156-245)."""
    control = parse_keyfile(os.path.join(path, "control"))
    aero = parse_keyfile(os.path.join(path, "aero"))
    ndim = int(control["ndim"])
    nnod = int(control["nnod"])
    neqn = int(control["neqn"])
    th = Thermo(gamma=float(aero["gamma"]), mach=float(aero["M"]),
                reynolds=float(aero["Re"]), prandtl=float(aero["Pr"]),
                tinf=float(aero["Tinf"]), twall=float(aero["Twall"]),
                adiabatic=bool(int(float(aero["Adiabatic"]))))
    nface = {3: 6, 2: 4, 1: 2}[ndim]
    raw_n = np.fromfile(os.path.join(path, "binarynodes"), dtype=np.float64)
    raw_n = raw_n.reshape(-1, 2 * ndim + 2)
    raw_e = np.fromfile(os.path.join(path, "binaryelements"), dtype=np.int32)
    raw_e = raw_e.reshape(-1, nnod + nface)
    nn = int(control.get("nodes", raw_n.shape[0]))
    ne = int(control.get("elements", raw_e.shape[0]))
    raw_n = raw_n[:nn]
    raw_e = raw_e[:ne]
    x = raw_n[:, :ndim].copy()
    rho = raw_n[:, ndim].copy()
    v = raw_n[:, ndim + 1:2 * ndim + 1].copy()
    T = raw_n[:, 2 * ndim + 1].copy()
    conn = raw_e[:, :nnod].astype(np.int32).copy()
    bc = raw_e[:, nnod:].astype(np.int32).copy()
    U = conservative(rho, v, T, th.Cv)
    d = dirichlet_flags(ndim, neqn, nn, conn, bc, th.adiabatic)
    mesh = Mesh(ndim, nnod, neqn, conn, x, bc, d)
    if int(float(control.get("adap", 0))):
        mesh.adap, mesh.hnm, mesh.refine_level = read_adap(os.path.join(path, "adap"), ne, nnod)
    return mesh, th, U, control


def read_adap(path: str, n_elem: int, nnod: int):
    """The adap file (LoadBinary.h:247-261): per record ``elno adap
    refine_level`` then the serialised Hnm tensor (TensorStream.h:95-175:
    type, rank, sizes, data in the Tensor's column-major order).  Elements
    without a record have adap = 0, level 0, Hnm = identity.  Parsing stops
    at the end of the last complete record (the reference's ``while
    (!file.eof())`` loop would read one record past it)."""
    adap = np.zeros(n_elem, dtype=np.uint8)
    hnm = np.tile(np.eye(nnod), (n_elem, 1, 1))
    level = np.zeros(n_elem, dtype=np.int32)
    tok = open(path).read().split()
    i = 0
    while i + 7 <= len(tok):
        e, a, lv = int(tok[i]), int(tok[i + 1]), int(tok[i + 2])
        rank, ni, nj = int(tok[i + 4]), int(tok[i + 5]), int(tok[i + 6])
        if rank != 2 or ni != nnod or nj != nnod or not 0 <= e < n_elem:
            raise ValueError(f"bad adap record for element {e} in {path}")
        vals = np.array([float(t) for t in tok[i + 7:i + 7 + ni * nj]])
        if vals.size != ni * nj:
            raise ValueError(f"short Hnm for element {e} in {path}")
        adap[e] = 1 if a else 0
        level[e] = lv
        hnm[e] = vals.reshape(nnod, nnod, order="F")
        i += 7 + ni * nj
    return adap, hnm, level


def write_deck(path: str, mesh: Mesh, th: Thermo, rho, v, T, control_extra=None) -> None:
    """Write a deck readable by both ``read_deck`` and the reference loader.

    Records are the reference's node_3d_t/node_2d_t (x, rho, v, T doubles)
    and ele_3d_t/ele_2d_t (node ids then bc codes, int32).
    """
    os.makedirs(path, exist_ok=True)
    nd = mesh.ndim
    rec = np.concatenate([mesh.x, np.asarray(rho)[:, None], np.asarray(v),
                          np.asarray(T)[:, None]], axis=1).astype(np.float64)
    rec.tofile(os.path.join(path, "binarynodes"))
    erec = np.concatenate([mesh.conn, mesh.bc], axis=1).astype(np.int32)
    erec.tofile(os.path.join(path, "binaryelements"))
    ctl = {"itermax": 100, "cfl": 0.5, "gmresrestart": 1, "gmresiter": 20,
           "ndim": nd, "nnod": mesh.nnod, "neqn": mesh.neqn, "nbnod": mesh.nbnod,
           "nface": mesh.nface, "nodes": mesh.n_node, "elements": mesh.n_elem,
           "iter": 0, "adap": int(mesh.adap is not None and bool(mesh.adap.any())), "axi": 0}
    if control_extra:
        ctl.update(control_extra)
    with open(os.path.join(path, "control"), "w") as fh:
        for k, v_ in ctl.items():
            fh.write(f"{k} = {v_}\n")
    with open(os.path.join(path, "aero"), "w") as fh:
        fh.write(f"M = {th.mach!r}\nPr = {th.prandtl!r}\nRe = {th.reynolds!r}\n"
                 f"Tinf = {th.tinf!r}\nTwall = {th.twall!r}\n"
                 f"Adiabatic = {int(th.adiabatic)}\ngamma = {th.gamma!r}\n")
    if mesh.adap is not None and mesh.adap.any():
        # SaveBinary::write_adap (SaveBinary.h:153-168): adapted elements only
        lv = mesh.refine_level if mesh.refine_level is not None else np.zeros(mesh.n_elem, np.int32)
        with open(os.path.join(path, "adap"), "w") as fh:
            for e in np.nonzero(mesh.adap)[0]:
                data = " ".join(repr(float(v)) for v in np.asarray(mesh.hnm[e]).ravel(order="F"))
                fh.write(f"{e} 1 {int(lv[e])} double 2 {mesh.nnod} {mesh.nnod} {data}\n")
    with open(os.path.join(path, "MPIBreakdown"), "w") as fh:
        fh.write("1\nelemin, elemax, nodemin, nodemax, nedgenode_l, nedgenode_r, "
                 "NodeIteratorMin, NodeIteratorMax\n")
        fh.write(f"0 {mesh.n_elem} 0 {mesh.n_node - 1} 0 0 0 {mesh.n_node - 1}\n")


def primitives(U: np.ndarray, ndim: int, Cv: float, gamma: float) -> dict:
    """NodeUnpack (src/FDV/Function/NodeUnpack.h:9-29), vectorised."""
    rho = U[:, 0]
    v = U[:, 1:1 + ndim] / U[:, :1]
    vs = np.zeros(U.shape[0])
    for i in range(ndim):
        vs = vs + v[:, i] * v[:, i]
    E = U[:, ndim + 1] / U[:, 0]
    e = E - 0.5 * vs
    return {"rho": rho, "v": v, "E": E, "e": e, "T": e / Cv, "p": rho * (gamma - 1.0) * e}
