"""Discrete adjoint of one FDV time step and its torch binding.

One step at a fixed dt: U1 = U0 + du with K(U0, M) du = b(U0, M) (the
element blocks of FDVGalerkin + ApplyBC, solved with the FP64 or mixed
GMRES).  For a scalar J(U1) with g = dJ/dU1:

    K^T lam = g,    dJ/dU0 = g + d/dU0 [lam^T (b - K du)],
                    dJ/dM  =     d/dM  [lam^T (b - K du)],

the bracket differentiated with lam and du held fixed (cfdv_build_vjp, the
forward-mode tangents of the element build on the device).  M is the Mach
number of the Thermo constants (Thermo::set, FDV/Thermo.h:22-33) with the
conservative state held fixed.  The wall temperature enters the step only
through the state of the wall nodes (ApplyBC.h's Twall branch is commented
out in the reference), so a wall-temperature gradient is the chain rule
through U: dU_E/dT = rho Cv at those nodes (``wall_temperature_gradient``).
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .fdv import FDVContext


def step_fixed_dt(ctx: FDVContext, U0, dt, mode="fp64", rtol=1e-12, precond="nodeblock", **kw):
    """U1 = U0 + du for one step at the given dt; returns (U1, du)."""
    ctx.set_state(U0)
    ctx.build(dt)
    kw.setdefault("restarts", 200 if mode == "fp64" else 40)
    ctx.solve(mode, m=kw.pop("m", 30), rtol=rtol, precond=precond, **kw)
    du = ctx.du()
    return np.asarray(U0).ravel() + du, du


def step_gradient(ctx: FDVContext, U0, dt, g, mode="fp64", rtol=1e-12, precond="nodeblock"):
    """(J-independent) gradient of J = g^T U1 with respect to U0 and the Mach
    number; returns (dJ/dU0 [n_node][neqn], dJ/dM, U1)."""
    U1, du = step_fixed_dt(ctx, U0, dt, mode=mode, rtol=rtol, precond=precond)
    lam, _ = ctx.adjoint_solve(g, mode=mode, rtol=rtol, precond=precond)
    gU, gM = ctx.build_vjp(lam, du)
    n = ctx.mesh.n_node
    return np.asarray(g).reshape(n, -1) + gU, gM, U1.reshape(n, -1)


def multi_step_gradient(ctx: FDVContext, U0, dts, g=None, mode="fp64", rtol=1e-12, precond="nodeblock",
                        functional=None, progress=None):
    """Gradient of J after N = len(dts) steps at the given dt's with respect
    to U0 and the Mach number: J = g^T U_N, or J = functional(U_N, M) (an
    object with value_and_grad(U, mach) -> (J, dJ/dU, dJ/dM), such as
    functionals.WallFunctional).  Returns (dJ/dU0 [n_node][neqn], dJ/dM,
    U_N), and J as a fourth value with a functional.

    Checkpointed: the forward sweep keeps only the N states U_0..U_{N-1}
    (host arrays); the reverse sweep re-runs step n from U_n and applies its
    adjoint, g_n = dJ/dU_n = g_{n+1} + d/dU_n [lam_n^T (b - K du)] with
    K^T lam_n = g_{n+1}, and dJ/dM sums the steps' Mach terms (plus the
    functional's own).  One forward solve and one forward plus one adjoint
    solve per step in reverse: 2N forward and N adjoint solves in all.
    ``progress(phase, k)`` is called after each step of either sweep.
    """
    if (g is None) == (functional is None):
        raise ValueError("give either g or functional")
    n = ctx.mesh.n_node
    states = []
    U = np.asarray(U0, dtype=np.float64).reshape(n, -1)
    for k, dt in enumerate(dts):
        states.append(U)
        U1, _ = step_fixed_dt(ctx, U, dt, mode=mode, rtol=rtol, precond=precond)
        U = U1.reshape(n, -1)
        if progress:
            progress("forward", k)
    gM = 0.0
    J = None
    if functional is not None:
        J, gU_N, gM = functional.value_and_grad(U, ctx.thermo.mach)
        gk = gU_N.ravel()
    else:
        gk = np.asarray(g, dtype=np.float64).ravel()
    for k, (Un, dt) in enumerate(zip(reversed(states), reversed(list(dts)))):
        gU, gMn, _ = step_gradient(ctx, Un, dt, gk, mode=mode, rtol=rtol, precond=precond)
        gk = gU.ravel()
        gM += gMn
        if progress:
            progress("reverse", len(dts) - 1 - k)
    if functional is not None:
        return gk.reshape(n, -1), gM, U, J
    return gk.reshape(n, -1), gM, U


def multi_step_gradient_multi(ctx: FDVContext, U0, dts, functionals, rtol=1e-12, precond="nodeblock",
                              inner_rtol=1e-4, progress=None):
    """Gradients of up to 8 functionals after N = len(dts) steps at once:
    the forward sweep as multi_step_gradient (FP64 steps, states kept), the
    reverse sweep with one 8-right-hand-side adjoint solve per step
    (cfdv_adjoint_solve_multi: FP64 refinement per functional around one
    FP32 global GMRES whose Arnoldi steps are 8-vector products K^T X) and
    one build tangent per functional.  Returns (list of dJ_s/dU0
    [n_node][neqn], list of dJ_s/dM, U_N, list of J_s)."""
    nf = len(functionals)
    if not 1 <= nf <= 8:
        raise ValueError("1 to 8 functionals")
    n = ctx.mesh.n_node
    states = []
    U = np.asarray(U0, dtype=np.float64).reshape(n, -1)
    for k, dt in enumerate(dts):
        states.append(U)
        U1, _ = step_fixed_dt(ctx, U, dt, mode="fp64", rtol=rtol, precond=precond)
        U = U1.reshape(n, -1)
        if progress:
            progress("forward", k)
    G = np.zeros((8, ctx.ndof))
    gM = [0.0] * nf
    Js = []
    for s, F in enumerate(functionals):
        J, gU_N, gMs = F.value_and_grad(U, ctx.thermo.mach)
        Js.append(J)
        G[s] = gU_N.ravel()
        gM[s] = gMs
    for k, (Un, dt) in enumerate(zip(reversed(states), reversed(list(dts)))):
        _, du = step_fixed_dt(ctx, Un, dt, mode="fp64", rtol=rtol, precond=precond)
        lam, _, _ = ctx.adjoint_solve_multi(G, rtol=rtol, inner_rtol=inner_rtol, precond=precond)
        for s in range(nf):
            gU, gMn = ctx.build_vjp(lam[s], du)
            G[s] = G[s] + gU.ravel()
            gM[s] += gMn
        if progress:
            progress("reverse", len(dts) - 1 - k)
    return [G[s].reshape(n, -1) for s in range(nf)], gM, U, Js


def functional_after_steps(ctx: FDVContext, U0, dts, functional, mach=None, mode="fp64", rtol=1e-12,
                           precond="nodeblock", progress=None):
    """functional(U_N, M) after N steps at fixed dt's from U0, with the Thermo
    Mach number set to ``mach`` for the steps and the functional (the
    context keeps that Mach number afterwards)."""
    if mach is not None:
        ctx.set_thermo(dataclasses.replace(ctx.thermo, mach=float(mach)))
    n = ctx.mesh.n_node
    U = np.asarray(U0, dtype=np.float64).reshape(n, -1)
    for k, dt in enumerate(dts):
        U1, _ = step_fixed_dt(ctx, U, dt, mode=mode, rtol=rtol, precond=precond)
        U = U1.reshape(n, -1)
        if progress:
            progress("forward", k)
    return functional.value(U, ctx.thermo.mach)


def wall_temperature_gradient(dJdU0, U0, Cv, wall_nodes):
    """dJ/dT_n at the given nodes, T = (E/rho - |v|^2/2)/Cv with rho and the
    velocity held: dU_E/dT = rho Cv (NodeUnpack.h:9-29 inverted)."""
    U0 = np.asarray(U0)
    rho = U0[wall_nodes, 0]
    return dJdU0[wall_nodes, -1] * rho * Cv


def _torch_function():
    import torch

    class FDVStep(torch.autograd.Function):
        """U1 = step(U0, M) at a fixed dt on the GPU (float64 host tensors in
        and out); backward through the discrete adjoint."""

        @staticmethod
        def forward(tctx, U0, mach, ctx: FDVContext, dt: float, mode: str = "fp64",
                    rtol: float = 1e-12):
            th = dataclasses.replace(ctx.thermo, mach=float(mach))
            ctx.set_thermo(th)
            U0n = U0.detach().cpu().numpy().astype(np.float64)
            U1, du = step_fixed_dt(ctx, U0n, dt, mode=mode, rtol=rtol)
            tctx.fdv = (ctx, dt, mode, rtol, U0n, du, float(mach))
            return torch.from_numpy(U1.reshape(U0n.shape).copy())

        @staticmethod
        def backward(tctx, gU1):
            ctx, dt, mode, rtol, U0n, du, mach = tctx.fdv
            ctx.set_thermo(dataclasses.replace(ctx.thermo, mach=mach))
            ctx.set_state(U0n)
            ctx.build(dt)
            g = gU1.detach().cpu().numpy().astype(np.float64)
            lam, _ = ctx.adjoint_solve(g, mode=mode, rtol=rtol)
            gU, gM = ctx.build_vjp(lam, du)
            gU0 = torch.from_numpy(g.reshape(gU.shape) + gU)
            return gU0, torch.tensor(gM, dtype=torch.float64), None, None, None, None

    return FDVStep


_FDVSTEP = None


def fdv_step(U0, mach, ctx: FDVContext, dt: float, mode="fp64", rtol=1e-12):
    """torch entry: U1 = FDV step of U0 at Mach ``mach`` and fixed ``dt``,
    differentiable with respect to U0 and mach."""
    global _FDVSTEP
    if _FDVSTEP is None:
        _FDVSTEP = _torch_function()
    return _FDVSTEP.apply(U0, mach, ctx, dt, mode, rtol)
