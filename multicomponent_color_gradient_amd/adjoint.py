"""Adjoint of the colour-gradient step: FP64 CPU reference, and a
``torch.autograd.Function`` over the GPU adjoint of the C ABI.

The state is the populations in the reference layout, shape (27, NZ, NY, NX)
per colour (f_i(x) after streaming, ghost slots unused); the parameters are
A (the perturbation amplitude) and beta (the recolouring parameter).  The
domain is periodic (``periodic_bubble_case``), the relaxation rate single,
and the storage FP32.  Gradients with respect to A and beta are accumulated
in FP64.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._abi import McgCase, as_ptr, check, load_library
from .solver import Solver


def cpu_adjoint(case: McgCase, fr0: np.ndarray, fb0: np.ndarray, nsteps: int, lam_fr: np.ndarray | None = None,
                lam_fb: np.ndarray | None = None, fields0: dict | None = None) -> dict:
    """FP64 forward of nsteps steps (y periodic; also with x/z walls, per-node
    viscosity and the body force) and, if lam_* is given, the adjoint: returns
    {"fr", "fb"} (final state) and {"lam_fr", "lam_fb", "dA", "dbeta", "dgravity"}.
    fields0 (the initial fields dict, ghosts included) is needed with walls."""
    lib = load_library()
    shape = (27, case.NZ, case.NY, case.NX)
    d = ctypes.c_double
    fr0 = np.ascontiguousarray(fr0, np.float64).reshape(shape)
    fb0 = np.ascontiguousarray(fb0, np.float64).reshape(shape)
    f0 = None
    if fields0 is not None:
        f0 = np.ascontiguousarray(np.stack([fields0[k] for k in ("rhor", "rhob", "ux", "uy", "uz")]), np.float64)
    out = {"fr": np.empty(shape), "fb": np.empty(shape)}
    dp = np.zeros(3)
    if lam_fr is not None:
        lam_fr = np.ascontiguousarray(lam_fr, np.float64).reshape(shape)
        lam_fb = np.ascontiguousarray(lam_fb, np.float64).reshape(shape)
        out["lam_fr"], out["lam_fb"] = np.empty(shape), np.empty(shape)
        args = (as_ptr(lam_fr, d), as_ptr(lam_fb, d))
        outs = (as_ptr(out["lam_fr"], d), as_ptr(out["lam_fb"], d), as_ptr(dp, d))
    else:
        args = (None, None)
        outs = (None, None, None)
    c = case.copy()
    check(lib.mcg_cpu_adjoint(ctypes.byref(c), int(nsteps), as_ptr(fr0, d), as_ptr(fb0, d),
                              None if f0 is None else as_ptr(f0, d), *args, as_ptr(out["fr"], d), as_ptr(out["fb"], d),
                              *outs))
    if lam_fr is not None:
        out["dA"], out["dbeta"], out["dgravity"] = float(dp[0]), float(dp[1]), float(dp[2])
    return out


def cpu_adjoint_jet(case: McgCase, fr0: np.ndarray, fb0: np.ndarray, fields0: dict, nsteps: int,
                    lam_fr: np.ndarray | None = None, lam_fb: np.ndarray | None = None, w_uy: np.ndarray | None = None,
                    uy_steps: tuple[int, int] = (0, -1), inlet_du: np.ndarray | None = None) -> dict:
    """FP64 forward and adjoint of a jet case (inlet and convective outlet
    differentiated).  fields0: the initial fields dict (rhor, rhob, ux, uy, uz
    over the full grid, ghosts included).  inlet_du: optional inlet velocity
    offset per column, shape (3, NZ, NX).  Returns the final state and, if a
    seed is given, {"lam_fr", "lam_fb", "dA", "dbeta", "dU", "damp", "dinlet"}
    (dinlet: dJ/de per jet column summed over steps, shape (3, NZ, NX)).  J is
    <lam, f_final> plus sum over the moments passes of steps uy_steps (numbered
    from 0) of sum_x w_uy(x) u_y(x)."""
    lib = load_library()
    shape = (27, case.NZ, case.NY, case.NX)
    d = ctypes.c_double
    fr0 = np.ascontiguousarray(fr0, np.float64).reshape(shape)
    fb0 = np.ascontiguousarray(fb0, np.float64).reshape(shape)
    f0 = np.ascontiguousarray(np.stack([fields0[k] for k in ("rhor", "rhob", "ux", "uy", "uz")]), np.float64)
    out = {"fr": np.empty(shape), "fb": np.empty(shape)}
    seeded = lam_fr is not None or w_uy is not None
    if lam_fr is None:
        lam_fr = np.zeros(shape)
        lam_fb = np.zeros(shape)
    lam_fr = np.ascontiguousarray(lam_fr, np.float64).reshape(shape)
    lam_fb = np.ascontiguousarray(lam_fb, np.float64).reshape(shape)
    w = None if w_uy is None else np.ascontiguousarray(w_uy, np.float64).reshape(shape[1:])
    dp = np.zeros(4)
    dinlet = np.zeros((3, case.NZ, case.NX))
    out["lam_fr"], out["lam_fb"] = np.empty(shape), np.empty(shape)
    c = case.copy()
    du = None if inlet_du is None else np.ascontiguousarray(inlet_du, np.float64).reshape((3, case.NZ, case.NX))
    check(lib.mcg_cpu_adjoint_jet(ctypes.byref(c), int(nsteps), as_ptr(fr0, d), as_ptr(fb0, d), as_ptr(f0, d),
                                  None if du is None else as_ptr(du, d), as_ptr(lam_fr, d) if seeded else None, as_ptr(lam_fb, d) if seeded else None,
                                  None if w is None else as_ptr(w, d), int(uy_steps[0]), int(uy_steps[1]),
                                  as_ptr(out["fr"], d), as_ptr(out["fb"], d), as_ptr(out["lam_fr"], d),
                                  as_ptr(out["lam_fb"], d), as_ptr(dp, d), as_ptr(dinlet, d)))
    if seeded:
        out.update(dA=float(dp[0]), dbeta=float(dp[1]), dU=float(dp[2]), damp=float(dp[3]), dinlet=dinlet)
    else:
        del out["lam_fr"], out["lam_fb"]
    return out


def _torch():
    import torch

    return torch


def make_steps_function():
    """``LBMSteps.apply(fr0, fb0, A, beta, case, nsteps, checkpoint_every[, max_states])`` -> (fr, fb):
    nsteps GPU steps from (fr0, fb0) with the case's A_override / beta_recolor
    replaced by the tensors A and beta; backward runs the GPU adjoint with
    checkpointed recomputation (max_states > 0: at most that many stored
    states, binomial recomputation, the same gradients)."""
    torch = _torch()

    class LBMSteps(torch.autograd.Function):
        @staticmethod
        def forward(ctx, fr0, fb0, A, beta, case, nsteps, checkpoint_every, *limit):
            c = case.copy()
            c.A_override = float(A)
            c.beta_recolor = float(beta)
            s = Solver(c)
            s.init_bubble_tanh()
            s.set_populations(fr0.detach().cpu().numpy(), fb0.detach().cpu().numpy())
            s.adjoint_set_checkpoint_limit(int(limit[0]) if limit else 0)
            ctx.n_limit = len(limit)
            s.adjoint_record(nsteps, checkpoint_every)
            fr, fb = s.populations()
            ctx.solver = s
            ctx.device, ctx.dtype = fr0.device, fr0.dtype
            return torch.from_numpy(fr).to(fr0.device), torch.from_numpy(fb).to(fr0.device)

        @staticmethod
        def backward(ctx, g_fr, g_fb):
            s = ctx.solver
            if g_fr is None:
                g_fr = torch.zeros((27,) + s.shape)
            if g_fb is None:
                g_fb = torch.zeros((27,) + s.shape)
            lr, lb, (dA, dbeta) = s.adjoint_backward(g_fr.detach().cpu().numpy(), g_fb.detach().cpu().numpy())
            s.close()
            ctx.solver = None
            t = lambda a: torch.from_numpy(a).to(device=ctx.device, dtype=ctx.dtype)  # noqa: E731
            return (t(lr), t(lb), torch.tensor(dA, dtype=torch.float64), torch.tensor(dbeta, dtype=torch.float64),
                    None, None, None) + (None,) * ctx.n_limit

    return LBMSteps


def make_rising_steps_function():
    """``LBMRisingSteps.apply(fr0, fb0, A, beta, gravity, case, nsteps, checkpoint_every[, max_states])`` -> (fr, fb):
    nsteps GPU steps of a periodic-y case (the rising bubble: x/z walls, body
    force on red, per-node viscosity) from the tanh-bubble initialisation with
    populations (fr0, fb0) and A_override, beta_recolor, gravity replaced by the
    tensors A, beta, gravity; backward runs the GPU adjoint and returns
    gradients for fr0, fb0, A, beta and gravity."""
    torch = _torch()

    class LBMRisingSteps(torch.autograd.Function):
        @staticmethod
        def forward(ctx, fr0, fb0, A, beta, gravity, case, nsteps, checkpoint_every, *limit):
            c = case.copy()
            c.A_override = float(A)
            c.beta_recolor = float(beta)
            c.gravity = float(gravity)
            s = Solver(c)
            s.init_bubble_tanh()
            s.set_populations(fr0.detach().cpu().numpy(), fb0.detach().cpu().numpy())
            s.adjoint_set_checkpoint_limit(int(limit[0]) if limit else 0)
            ctx.n_limit = len(limit)
            s.adjoint_record(nsteps, checkpoint_every)
            fr, fb = s.populations()
            ctx.solver = s
            ctx.device, ctx.dtype = fr0.device, fr0.dtype
            return torch.from_numpy(fr).to(fr0.device), torch.from_numpy(fb).to(fr0.device)

        @staticmethod
        def backward(ctx, g_fr, g_fb):
            s = ctx.solver
            if g_fr is None:
                g_fr = torch.zeros((27,) + s.shape)
            if g_fb is None:
                g_fb = torch.zeros((27,) + s.shape)
            lr, lb, _ = s.adjoint_backward(g_fr.detach().cpu().numpy(), g_fb.detach().cpu().numpy())
            g = s.adjoint_gradients()
            s.close()
            ctx.solver = None
            t = lambda a: torch.from_numpy(a).to(device=ctx.device, dtype=ctx.dtype)  # noqa: E731
            d64 = lambda v: torch.tensor(v, dtype=torch.float64)  # noqa: E731
            return (t(lr), t(lb), d64(g["dA"]), d64(g["dbeta"]), d64(g["dgravity"]), None, None, None) + (None,) * ctx.n_limit

    return LBMRisingSteps


def make_jet_steps_function():
    """``LBMJetSteps.apply(fr0, fb0, A, beta, U, amp, du, case, nsteps, checkpoint_every, w_uy, window[, max_states])``
    -> (fr, fb, J_uy): nsteps GPU steps of a jet case from the jet initialisation
    with populations (fr0, fb0), A_override, beta_recolor, jet_velocity and
    perturbation_amp replaced by the tensors A, beta, U, amp, and the inlet
    velocity offset du (3, NZ, NX) added at every jet column.  J_uy is the u_y
    functional sum over the moments passes of steps window[0]..window[1] of
    sum_x w_uy(x) u_y(x) (0 when w_uy is None).  backward runs the GPU jet
    adjoint and returns gradients for fr0, fb0, A, beta, U, amp and du."""
    torch = _torch()

    class LBMJetSteps(torch.autograd.Function):
        @staticmethod
        def forward(ctx, fr0, fb0, A, beta, U, amp, du, case, nsteps, checkpoint_every, w_uy, window, *limit):
            c = case.copy()
            c.A_override = float(A)
            c.beta_recolor = float(beta)
            c.jet_velocity = float(U)
            c.perturbation_amp = float(amp)
            s = Solver(c)
            s.init_jet()
            s.set_populations(fr0.detach().cpu().numpy(), fb0.detach().cpu().numpy())
            s.set_inlet_offset(du.detach().cpu().numpy())
            w = None if w_uy is None else np.ascontiguousarray(w_uy.detach().cpu().numpy(), np.float32)
            if w is not None:
                s.adjoint_set_uy_seed(w, int(window[0]), int(window[1]))
            s.adjoint_set_checkpoint_limit(int(limit[0]) if limit else 0)
            ctx.n_limit = len(limit)
            s.adjoint_record(nsteps, checkpoint_every)
            J = s.uy_functional() if w is not None else 0.0
            fr, fb = s.populations()
            ctx.solver, ctx.w, ctx.window = s, w, window
            ctx.device, ctx.dtype = fr0.device, fr0.dtype
            return (torch.from_numpy(fr).to(fr0.device), torch.from_numpy(fb).to(fr0.device),
                    torch.tensor(J, dtype=torch.float64))

        @staticmethod
        def backward(ctx, g_fr, g_fb, g_J):
            s = ctx.solver
            if g_fr is None:
                g_fr = torch.zeros((27,) + s.shape)
            if g_fb is None:
                g_fb = torch.zeros((27,) + s.shape)
            if ctx.w is not None:
                gj = 0.0 if g_J is None else float(g_J)
                s.adjoint_set_uy_seed(ctx.w * np.float32(gj), int(ctx.window[0]), int(ctx.window[1]))
            lr, lb, (dA, dbeta) = s.adjoint_backward(g_fr.detach().cpu().numpy(), g_fb.detach().cpu().numpy())
            (dU, damp), din = s.adjoint_inlet_gradient()
            s.close()
            ctx.solver = None
            t = lambda a: torch.from_numpy(a).to(device=ctx.device, dtype=ctx.dtype)  # noqa: E731
            d64 = lambda v: torch.tensor(v, dtype=torch.float64)  # noqa: E731
            return (t(lr), t(lb), d64(dA), d64(dbeta), d64(dU), d64(damp), torch.from_numpy(din),
                    None, None, None, None, None) + (None,) * ctx.n_limit

    return LBMJetSteps


def bubble_centroid_height(fr: "torch.Tensor", fb: "torch.Tensor") -> "torch.Tensor":
    """Height (y) of the blue component's centroid over the interior nodes: the
    objective of the rising-bubble adjoint (differentiable in torch)."""
    torch = _torch()
    rb = fb[:, 1:-1, 1:-1, 1:-1].sum(0).double()
    ny = rb.shape[1]
    y = torch.arange(1, ny + 1, dtype=torch.float64).view(1, ny, 1)
    return (rb * y).sum() / rb.sum()
