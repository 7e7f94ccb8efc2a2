"""Quantum-information helpers on torch tensors (mirrors jaqsi/math.py:59-206).

``fidelity`` is the objective of pulse-level training (the reference's
tests/test_training.py and ``unitary_cost_fn``); it is differentiable with
torch's autograd, so gradients reach the pulse parameters through the
state-output adjoint of ``Script.execute(type="state")``.
"""

from __future__ import annotations

import torch


def _as(x) -> torch.Tensor:
    return x if isinstance(x, torch.Tensor) else torch.as_tensor(x)


def _is_vector(s: torch.Tensor) -> bool:
    return s.dim() == 1 or (s.dim() == 2 and s.shape[-2] != s.shape[-1])


def _sqrt_psd(rho: torch.Tensor) -> torch.Tensor:
    evs, vecs = torch.linalg.eigh(rho)
    evs = torch.clamp(evs.real, min=0.0).to(rho.dtype)
    return (vecs * torch.sqrt(evs).unsqueeze(-2)) @ vecs.conj().transpose(-1, -2)


def fidelity(state0, state1) -> torch.Tensor:
    """|<psi|phi>|^2 of normalised state vectors, or (Tr sqrt(sqrt(rho) sigma
    sqrt(rho)))^2 of density matrices (math.py:59-143); batched along a
    leading axis."""
    s0, s1 = _as(state0), _as(state1)
    if s0.shape[-1] != s1.shape[-1]:
        raise ValueError("The two states must have the same number of wires.")
    v0, v1 = _is_vector(s0), _is_vector(s1)
    if v0 != v1:
        raise ValueError("Both states must be of the same kind (both state vectors or both density matrices).")
    dt = torch.promote_types(s0.dtype, s1.dtype)
    if not dt.is_complex:
        dt = torch.complex128
    s0, s1 = s0.to(dt), s1.to(s0.device if s0.device.type != "cpu" else s1.device, dt)
    s0 = s0.to(s1.device)
    if v0:
        n0 = torch.linalg.norm(s0, dim=-1, keepdim=True)
        n1 = torch.linalg.norm(s1, dim=-1, keepdim=True)
        s0 = s0 / torch.where(n0 > 0, n0, torch.ones_like(n0))
        s1 = s1 / torch.where(n1 > 0, n1, torch.ones_like(n1))
        return torch.abs((s0.conj() * s1).sum(-1)) ** 2
    r = _sqrt_psd(s0)
    evs = torch.linalg.eigvalsh(r @ s1 @ r)
    return torch.sqrt(torch.clamp(evs.real, min=0.0)).sum(-1) ** 2


def trace_distance(state0, state1) -> torch.Tensor:
    """(1/2) sum |eig(rho - sigma)| (math.py:146-169)."""
    s0, s1 = _as(state0), _as(state1)
    if s0.shape[-1] != s1.shape[-1]:
        raise ValueError("The two states must have the same number of wires.")
    return torch.abs(torch.linalg.eigvalsh(s0 - s1.to(s0.device, s0.dtype))).sum(-1) / 2


def phase_difference(state0, state1) -> torch.Tensor:
    """arg <psi|phi> of two state vectors (math.py:172-206)."""
    s0, s1 = _as(state0), _as(state1)
    if s0.shape[-1] != s1.shape[-1]:
        raise ValueError("The two states must have the same number of wires.")
    return torch.angle((s0.conj() * s1.to(s0.device, s0.dtype)).sum(-1))
