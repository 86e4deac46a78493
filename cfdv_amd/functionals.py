"""Wall functionals of a hex8 state for the adjoint: integrated drag and
integrated wall heat flux, with their gradients (torch autograd, float64 on
the host).

Both are surface integrals over the element faces of one bc code (the
wall), at the 2 x 2 Gauss points of each face, of the state the face's
element interpolates -- the primitive variables at the element's nodes
(NodeUnpack, FDV/Function/NodeUnpack.h:9-29) interpolated with the
trilinear test functions and differentiated with the element's own
Jacobian (TestFunctions3D.h:7-134), as FDVGalerkin::sum forms the Gauss-point
state (FDV/FDVGalerkin.h:761-797):

    drag  D = sum (p n_x - tau_xj n_j / Re) dA,
            p = (gamma - 1) rho e,  tau = mu (grad v + grad v^T - 2/3 div v I)
    heat  Q = sum k (grad T . n) dA / Re,   k = mu / ((gamma - 1) M^2 Pr)

with n the unit normal out of the element (into the body), dA the face's
area element and mu Sutherland's law (Thermo.h:53-75).  The Mach number
enters through Cv = 1 / (gamma (gamma - 1) M^2) (T = e / Cv) and k, so
each functional is J(U, M).
"""
from __future__ import annotations

import numpy as np

# local hex node -> corner signs (TestFunctions3D.h:12-19)
SIGNS = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)
# the fixed local coordinate of each face (get_face.h:106-165 node lists)
FACE_AXIS = np.array([2, 1, 0, 1, 0, 2])
FACE_SIDE = np.array([-1.0, -1.0, 1.0, 1.0, -1.0, 1.0])


def _face_tables():
    """Phi [6][4][8] and dPhi/dzeta [6][4][8][3] at the 2 x 2 Gauss points
    of each face, and the two in-face axes of each face."""
    z = 1.0 / np.sqrt(3.0)
    phi = np.zeros((6, 4, 8))
    dphi = np.zeros((6, 4, 8, 3))
    free = np.zeros((6, 2), dtype=np.int64)
    for f in range(6):
        ax = FACE_AXIS[f]
        fa = [d for d in range(3) if d != ax]
        free[f] = fa
        for g, (za, zb) in enumerate(((-z, -z), (z, -z), (z, z), (-z, z))):
            zeta = np.zeros(3)
            zeta[ax] = FACE_SIDE[f]
            zeta[fa[0]], zeta[fa[1]] = za, zb
            fac = 1.0 + SIGNS * zeta
            phi[f, g] = np.prod(fac, axis=1) / 8.0
            for d in range(3):
                o = [q for q in range(3) if q != d]
                dphi[f, g, :, d] = SIGNS[:, d] * fac[:, o[0]] * fac[:, o[1]] / 8.0
    return phi, dphi, free


class WallFunctional:
    """J(U, M) over the faces with bc code ``code``; ``kind`` "drag" or
    "heat"; ``x_range`` = (x0, x1) keeps the faces whose element centroid
    has x0 <= x < x1 (a streamwise patch of the wall).
    ``value_and_grad(U, mach)`` returns (J, dJ/dU [n_node][5], dJ/dM)."""

    def __init__(self, mesh, thermo, code, kind="drag", x_range=None):
        import torch
        if mesh.ndim != 3:
            raise ValueError("wall functionals are hex8 only")
        if kind not in ("drag", "heat"):
            raise ValueError(kind)
        self.kind = kind
        self.th = thermo
        e, f = np.nonzero(np.asarray(mesh.bc) == code)
        if x_range is not None:
            xc = np.asarray(mesh.x)[np.asarray(mesh.conn)[e], 0].mean(axis=1)
            keep = (xc >= x_range[0]) & (xc < x_range[1])
            e, f = e[keep], f[keep]
        if len(e) == 0:
            raise ValueError(f"no face with bc code {code}")
        self.elems, self.faces = e, f
        phi, dphi, _ = _face_tables()
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(torch.float64)   # noqa: E731
        self.conn = torch.from_numpy(np.asarray(mesh.conn)[e].astype(np.int64))     # [F][8]
        X = np.asarray(mesh.x)[np.asarray(mesh.conn)[e]]                            # [F][8][3]
        self.phi = t(phi[f])                                                        # [F][4][8]
        dz = dphi[f]                                                                # [F][4][8][3]
        jac = np.einsum("fgnd,fni->fgid", dz, X)                                    # dx_i / dzeta_d
        jinv = np.linalg.inv(jac)                                                   # dzeta_d / dx_i
        self.dphi = t(np.einsum("fgnd,fgdi->fgni", dz, jinv))                       # dPhi_N / dx_i
        # area element and unit normal out of the element
        ax = FACE_AXIS[f]
        fa = np.array([[d for d in range(3) if d != a] for a in ax])
        ta = np.take_along_axis(jac, fa[:, None, None, 0:1].repeat(4, 1).repeat(3, 2), axis=3)[..., 0]
        tb = np.take_along_axis(jac, fa[:, None, None, 1:2].repeat(4, 1).repeat(3, 2), axis=3)[..., 0]
        nr = np.cross(ta, tb)
        outward = np.take_along_axis(jac, ax[:, None, None, None].repeat(4, 1).repeat(3, 2), axis=3)[..., 0]
        outward = outward * FACE_SIDE[f][:, None, None]
        sgn = np.sign(np.einsum("fgi,fgi->fg", nr, outward))
        dA = np.linalg.norm(nr, axis=2)
        self.n = t(nr / dA[..., None] * sgn[..., None])                             # [F][4][3]
        self.dA = t(dA)                                                             # Gauss weights 1

    def _J(self, U, mach):
        import torch
        th = self.th
        g = th.gamma
        Ue = U[self.conn]                                                           # [F][8][5]
        rho = Ue[..., 0]
        v = Ue[..., 1:4] / rho[..., None]
        e = Ue[..., 4] / rho - 0.5 * (v * v).sum(-1)
        Cv = 1.0 / (g * (g - 1.0) * mach * mach)
        T = e / Cv
        cs = 110.0 / th.tinf
        Tg = torch.einsum("fgn,fn->fg", self.phi, T)
        mu = (1.0 + cs) / (Tg + cs) * Tg * torch.sqrt(Tg)
        if self.kind == "drag":
            p = (g - 1.0) * torch.einsum("fgn,fn->fg", self.phi, rho * e)
            dV = torch.einsum("fgnj,fni->fgij", self.dphi, v)                        # dv_i / dx_j
            div = dV.diagonal(dim1=2, dim2=3).sum(-1)
            eye = torch.eye(3, dtype=U.dtype)
            tau = mu[..., None, None] * (dV + dV.transpose(2, 3) - 2.0 / 3.0 * div[..., None, None] * eye)
            tx = (tau[:, :, 0, :] * self.n).sum(-1)
            integrand = p * self.n[..., 0] - tx / th.reynolds
        else:
            k = mu / ((g - 1.0) * mach * mach * th.prandtl)
            dT = torch.einsum("fgnj,fn->fgj", self.dphi, T)
            integrand = k * (dT * self.n).sum(-1) / th.reynolds
        return (integrand * self.dA).sum()

    def value(self, U, mach=None):
        import torch
        m = self.th.mach if mach is None else mach
        with torch.no_grad():
            return float(self._J(torch.from_numpy(np.asarray(U, dtype=np.float64).reshape(-1, 5)),
                                 torch.tensor(float(m), dtype=torch.float64)))

    def value_and_grad(self, U, mach=None):
        import torch
        m = self.th.mach if mach is None else mach
        Ut = torch.from_numpy(np.array(U, dtype=np.float64).reshape(-1, 5)).requires_grad_(True)
        mt = torch.tensor(float(m), dtype=torch.float64, requires_grad=True)
        J = self._J(Ut, mt)
        gU, gM = torch.autograd.grad(J, [Ut, mt])
        return float(J.detach()), gU.numpy(), float(gM)
