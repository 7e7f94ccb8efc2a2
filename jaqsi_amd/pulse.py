"""Batched two-qubit pulse propagators on the GPU, with autograd.

The reference evolves a parametrised Hamiltonian with the Magnus-2 solver
(``Evolution._build_magnus_evolve_solver``, jaqsi/evolution.py:194-263) and
scores the propagator with the process loss of ``unitary_cost_fn``
(jaqsi/qoc.py:335-341).  Here one launch does that for a whole batch of
parameter sets of the drive Hamiltonian

    H(t) = w0/2 Z0 + w1/2 Z1 + J Z0 Z1
           + E(t) (cos phi X0 + sin phi Y0) + Q(t) (-sin phi X0 + cos phi Y0)

with E the lifted Gaussian of amplitude A and width sigma centred at T/2
(jaqsi/pulses.py:324-352) and Q = -beta dE/dt its DRAG quadrature
(pulses.py:377-390), in FP64, and returns the gradient of the infidelity
with respect to (A, sigma, beta, phi) from the same launch.
"""

from __future__ import annotations

import math
from typing import Optional, Tuple

import numpy as np
import torch

from jaqsi_amd import _native as nat

COUPLING_J = 2 * math.pi * 0.005  # rad/ns
T_TOTAL = 50.0  # ns
N_STEPS = 2000


def _target_host(target) -> np.ndarray:
    v = target.detach().cpu().numpy() if isinstance(target, torch.Tensor) else np.asarray(target)
    v = np.asarray(v, dtype=np.complex128)
    if v.shape != (4, 4):
        raise ValueError(f"target must be a 4x4 matrix, got shape {v.shape}")
    return np.ascontiguousarray(np.stack([v.real, v.imag], axis=-1).reshape(-1))


def _params(params: torch.Tensor, omega: torch.Tensor) -> torch.Tensor:
    if not isinstance(params, torch.Tensor) or not isinstance(omega, torch.Tensor):
        raise TypeError("params and omega must be torch tensors")
    if params.dim() != 2 or params.shape[1] != 4:
        raise ValueError(f"params must be (n_sets, 4) = (A, sigma, beta, phi), got {tuple(params.shape)}")
    if omega.shape != (params.shape[0], 2):
        raise ValueError(f"omega must be (n_sets, 2) = (w0, w1), got {tuple(omega.shape)}")
    if params.device.type != "cuda" or omega.device != params.device:
        raise ValueError("params and omega must be on the same CUDA device")
    return torch.cat([params.detach(), omega.detach()], dim=1).to(torch.float64).contiguous()


def propagate(params: torch.Tensor, omega: torch.Tensor, target, *, n_steps: int = N_STEPS,
              t_total: float = T_TOTAL, coupling: float = COUPLING_J, want_unitaries: bool = True,
              want_grad: bool = True) -> Tuple[Optional[torch.Tensor], torch.Tensor, Optional[torch.Tensor]]:
    """Propagators U (n_sets, 4, 4) complex128, infidelities 1 - |Tr(V^dag U)|^2 / 16
    (n_sets,) and d infidelity / d(A, sigma, beta, phi) (n_sets, 4), on the
    current stream of params' device.  Wire 0 is the most significant bit."""
    P = _params(params, omega)
    n = P.shape[0]
    if n_steps < 1:
        raise ValueError("n_steps must be positive")
    Vh = _target_host(target)
    dev = P.device
    U = torch.empty((n, 4, 4), dtype=torch.complex128, device=dev) if want_unitaries else None
    inf = torch.empty(n, dtype=torch.float64, device=dev)
    g = torch.empty((n, 4), dtype=torch.float64, device=dev) if want_grad else None
    if n == 0:
        return U, inf, g
    d = nat.PulseDesc(n, int(n_steps), float(t_total), float(coupling), nat.JQ_ENV_DRAG_GAUSSIAN, 0)
    with torch.cuda.device(dev):
        stream = torch.cuda.current_stream(dev).cuda_stream
        nat.check(nat.lib().jq_pulse_propagate(d, nat.ptr(P), nat.ptr(Vh), nat.ptr(U), nat.ptr(inf), nat.ptr(g),
                                               stream))
    return U, inf, g


class _Infidelity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, omega, target, n_steps, t_total, coupling):
        _, inf, g = propagate(params, omega, target, n_steps=n_steps, t_total=t_total, coupling=coupling,
                              want_unitaries=False, want_grad=ctx.needs_input_grad[0])
        ctx.save_for_backward(g if g is not None else inf.new_empty(0))
        ctx.dtype = params.dtype
        return inf

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        if g.numel() == 0:
            return None, None, None, None, None, None
        return (grad_out[:, None] * g).to(ctx.dtype), None, None, None, None, None


def infidelity(params: torch.Tensor, omega: torch.Tensor, target, *, n_steps: int = N_STEPS,
               t_total: float = T_TOTAL, coupling: float = COUPLING_J) -> torch.Tensor:
    """Process infidelity per parameter set, differentiable w.r.t. params
    (the gradient comes from the same launch; omega is held fixed)."""
    return _Infidelity.apply(params, omega, target, n_steps, t_total, coupling)
