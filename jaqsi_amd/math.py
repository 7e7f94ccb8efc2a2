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


# ---------------------------------------------------------------------------
# quantum Fisher information (math.py:209-448)
# ---------------------------------------------------------------------------
def _fubini_study(jac: torch.Tensor, state: torch.Tensor) -> torch.Tensor:
    """Re[<d_i psi|d_j psi> - <d_i psi|psi><psi|d_j psi>] from the
    Jacobian (d, P) of a normalised state (d,)."""
    A = jac.conj().T @ jac
    v = jac.conj().T @ state
    return (A - torch.outer(v, v.conj())).real


def _qfi_density(jac: torch.Tensor, rho: torch.Tensor, eps: float = 1e-12) -> torch.Tensor:
    """Symmetric-logarithmic-derivative QFI of rho (d, d) from d rho / d theta
    (d, d, P): 2 sum_{k,l: p_k + p_l > eps} Re(<k|d_i rho|l><l|d_j rho|k>) /
    (p_k + p_l) over the eigenpairs (p_k, |k>) of rho (negative eigenvalues
    from rounding are clamped to zero)."""
    p, V = torch.linalg.eigh(rho)
    p = p.real.clamp_min(0.0)
    M = V.conj().T @ jac.permute(2, 0, 1) @ V          # (P, d, d): <k|d_i rho|l>
    s = p[:, None] + p[None, :]
    w = torch.where(s > eps, 2.0 / torch.where(s > eps, s, torch.ones_like(s)), torch.zeros_like(s))
    return torch.einsum("ikl,jkl->ij", M * w, M.conj()).real


def _captured_jacobian(state_fn, params):
    """Run state_fn once under capture; return (state, J w.r.t. the
    recorded angles, dangles/dparams (P_angles, P))."""
    from jaqsi_amd import simulation

    p = params if isinstance(params, torch.Tensor) else torch.as_tensor(params, dtype=torch.float64)
    p = p.detach().to(torch.float64).requires_grad_(True)
    with simulation.capture_runs() as runs:
        out = state_fn(p)
    if len(runs) != 1:
        raise ValueError(f"state_fn must run exactly one circuit, it ran {len(runs)}")
    run = runs[0]
    if run.mats is not None and run.mats.requires_grad:
        raise NotImplementedError("the QFI is taken through angle parameters; explicit matrices that depend on the "
                                  "parameters are not supported")
    if run.psi0 is not None and run.psi0.requires_grad:
        raise NotImplementedError("the QFI through a parameter-dependent initial state is not supported")
    state, J = simulation.angle_jacobian(run)
    ang = run.angles.reshape(-1)
    P = p.numel()
    if ang.numel() and ang.requires_grad:
        eye = torch.eye(ang.numel(), dtype=ang.dtype, device=ang.device)
        (A,) = torch.autograd.grad(ang, p, grad_outputs=eye, is_grads_batched=True, allow_unused=True)
        A = torch.zeros((ang.numel(), P), dtype=torch.float64) if A is None else A.reshape(ang.numel(), P)
    else:
        A = torch.zeros((ang.numel(), P), dtype=torch.float64)
    return out, state, J, A.to(J.device, torch.float64)


def quantum_fisher_information(state_fn, params) -> torch.Tensor:
    """QFI (P, P) of the state ``state_fn(params)`` (math.py:331-383): the
    Fubini-Study form 4 Re[<d_i psi|d_j psi> - <d_i psi|psi><psi|d_j psi>]
    for a state vector, the symmetric-logarithmic-derivative form for a
    density matrix.

    ``state_fn`` maps the parameters to the output of one
    ``Script.execute(type="state" | "density")``.  The Jacobian with respect
    to the circuit's recorded angles comes from one batched adjoint sweep on
    the GPU (``simulation.angle_jacobian``); the chain to ``params`` is the
    Jacobian of the recorded angles, by torch autograd: F = A^T F_angles A."""
    out, state, J, A = _captured_jacobian(state_fn, params)
    if out.dim() == 1:
        F = 4.0 * _fubini_study(J, state)
    elif out.dim() == 2 and out.shape[0] == out.shape[1]:
        d = out.shape[0]
        F = _qfi_density(J.reshape(d, d, -1), state.reshape(d, d))
    else:
        raise ValueError("state_fn must return a state vector of shape (d,) or a density matrix of shape (d, d), "
                         f"got shape {tuple(out.shape)}.")
    return A.T @ F @ A


def fubini_study_metric(state_fn, params) -> torch.Tensor:
    """The Fubini-Study metric (P, P) of a pure state (math.py:386-419):
    a quarter of its QFI."""
    out, state, J, A = _captured_jacobian(state_fn, params)
    if out.dim() != 1:
        raise ValueError("The Fubini-Study metric is only defined for pure states; state_fn must return a state "
                         f"vector of shape (d,), got shape {tuple(out.shape)}.")
    return A.T @ _fubini_study(J, state) @ A


# ---------------------------------------------------------------------------
# reductions (math.py:422-519)
# ---------------------------------------------------------------------------
def partial_trace(rho, n_qubits: int, keep) -> torch.Tensor:
    """Reduced density matrix on the wires ``keep`` (in increasing order) of
    rho (2^n, 2^n) or a batch (B, 2^n, 2^n)."""
    r = _as(rho)
    batched = r.dim() == 3
    B = r.shape[0] if batched else 1
    t = r.reshape((B,) + (2,) * (2 * n_qubits))
    keep = sorted(keep)
    # a traced wire shares one index between its ket and bra axes
    letters = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ"
    ket = [letters[w] for w in range(n_qubits)]
    bra = [letters[n_qubits + w] if w in keep else letters[w] for w in range(n_qubits)]
    res = "".join(ket[w] for w in keep) + "".join(bra[w] for w in keep)
    t = torch.einsum("z" + "".join(ket) + "".join(bra) + "->z" + res, t)
    dim = 2 ** len(keep)
    t = t.reshape(B, dim, dim)
    return t if batched else t[0]


def marginalize_probs(probs, n_qubits: int, keep) -> torch.Tensor:
    """Marginal distribution (B, 2^k) of the wires ``keep`` (in increasing
    order) from probabilities (2^n,) or (B, 2^n)."""
    p = _as(probs).reshape(-1, *(2,) * n_qubits)
    drop = tuple(1 + w for w in range(n_qubits) if w not in keep)
    if drop:
        p = p.sum(dim=drop)
    return p.reshape(p.shape[0], -1)
