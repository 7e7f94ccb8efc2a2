"""Host side of the simulation seam.

``simulate_and_measure`` keeps the signature and behaviour of
jaqsi/simulation.py:658-732: statevector or density-matrix simulation by
whether the tape holds a channel, the adjoint-method VJP for exact
expectation values, and the measurement types "expval", "probs", "state"
and "density".  Everything numeric runs in libjaqsi_amd.so on the GPU; this
module lowers the tape, keeps the native plans cached per circuit
structure and moves tensors across the C-ABI.  There is no CPU path: if no
GPU or no built library is present, the calls raise.
"""

from __future__ import annotations

import threading
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from jaqsi_amd import _native as nat
from jaqsi_amd import lowering
from jaqsi_amd.gateset import Barrier
from jaqsi_amd.noise import KrausChannel
from jaqsi_amd.operations import Operation, x64_enabled

_plans: dict = {}
_plans_lock = threading.Lock()
_PLAN_CACHE_MAX = 256
_jit_mode: Optional[int] = None  # None: the library's default per plan


def specialise(mode: Optional[int]) -> None:
    """Choose the kernels of plans created from now on: nat.JQ_JIT_OFF (the
    program interpreter), JQ_JIT_ASYNC (plan-specialised kernels once their
    background compile ends), JQ_JIT_SYNC (compile when the plan is made) or
    None (the library default: specialise plans of JQ_JIT_MIN_WIRES or more
    wires in the background).  Cached plans are dropped."""
    global _jit_mode
    with _plans_lock:
        _jit_mode = mode
        _plans.clear()


def cached_plans() -> list:
    """The plans currently cached (for inspection)."""
    with _plans_lock:
        return list(_plans.values())


def infer_n_qubits(ops: List[Operation], obs: List[Operation]) -> int:
    """max wire + 1 over the tape and observables (simulation.py:32-47)."""
    wires = set()
    for op in list(ops) + list(obs):
        wires.update(op.wires)
    return max(wires) + 1 if wires else 1


def has_noise(tape: List[Operation]) -> bool:
    """A channel on the tape requires density-matrix simulation (simulation.py:50-62)."""
    return any(isinstance(op, KrausChannel) for op in tape)


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("jaqsi_amd simulates on the GPU only; no GPU is visible")
    return torch.device("cuda", torch.cuda.current_device())


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _dtype_code() -> int:
    return 1 if x64_enabled() else 0


def _cdtype_torch(code: int) -> torch.dtype:
    return torch.complex128 if code else torch.complex64


def get_plan(low: lowering.Lowered, dtype: int, tile_bits: int = 0, checkpoints: int = 1) -> nat.Plan:
    key = (low.key, dtype, tile_bits, checkpoints)
    with _plans_lock:
        p = _plans.get(key)
        if p is None:
            p = nat.Plan(low.n_wires, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, dtype,
                         tile_bits, checkpoints)
            if _jit_mode is not None:
                p.jit(_jit_mode)
            if len(_plans) >= _PLAN_CACHE_MAX:
                _plans.pop(next(iter(_plans)))
            _plans[key] = p
        return p


# ---------------------------------------------------------------------------
# observables
# ---------------------------------------------------------------------------
def diag_observables(obs: Sequence[Operation]) -> Optional[List[Tuple[int, float, float]]]:
    """(wire, d0, d1) for single-qubit diagonal observables, else None.

    Same test as the reference fast path (simulation.py:772-801): the class
    matrix is used, so instances with a custom matrix take the general path.
    """
    out = []
    for ob in obs:
        m = ob.__class__._matrix
        if m is None or len(ob.wires) != 1 or getattr(ob, "_explicit", False):
            return None
        m = np.asarray(m)
        if m.shape != (2, 2) or not np.allclose(m - np.diag(np.diag(m)), 0):
            return None
        d = np.real(np.diag(m))
        out.append((int(ob.wires[0]), float(d[0]), float(d[1])))
    return out


# ---------------------------------------------------------------------------
# general observables: Pauli decomposition + basis-rotated parity groups
# ---------------------------------------------------------------------------
_PAULI_1Q = {
    "I": np.eye(2, dtype=complex),
    "X": np.array([[0, 1], [1, 0]], dtype=complex),
    "Y": np.array([[0, -1j], [1j, 0]], dtype=complex),
    "Z": np.array([[1, 0], [0, -1]], dtype=complex),
}


def pauli_terms(m: np.ndarray, wires: Sequence[int], tol: float = 1e-12) -> List[Tuple[dict, float]]:
    """Decompose a Hermitian matrix on `wires` (first wire = most significant
    index bit) into real-weighted Pauli words: [({wire: letter}, coeff)]."""
    import itertools

    k = len(wires)
    m = np.asarray(m, dtype=complex)
    if m.shape != (2**k, 2**k):
        raise ValueError("observable matrix does not match its wires")
    if not np.allclose(m, m.conj().T, atol=1e-10):
        raise NotImplementedError("gradients of non-Hermitian observables are not supported")
    out = []
    for letters in itertools.product("IXYZ", repeat=k):
        P = np.array([[1.0 + 0j]])
        for c in letters:
            P = np.kron(P, _PAULI_1Q[c])
        c = np.real(np.trace(P @ m)) / 2**k
        if abs(c) > tol:
            out.append(({w: l for w, l in zip(wires, letters) if l != "I"}, float(c)))
    return out


def pauli_groups(obs: Sequence[Operation]):
    """Group the Pauli terms of all observables qubit-wise by measurement
    basis.  Returns (groups, n_obs) with groups = [(basis {wire: X|Y|Z},
    [(obs index, coeff, wire mask)])]; a term with no wires (identity) gets
    mask 0 and contributes coeff * 1."""
    groups: List[Tuple[dict, list]] = []
    for oi, ob in enumerate(obs):
        if len(ob.wires) > 6:
            raise NotImplementedError("observables on more than 6 wires are not supported by the gradient path")
        for word, c in pauli_terms(_obs_matrix_f64(ob), ob.wires):
            mask = 0
            for w in word:
                mask |= 1 << w
            for basis, terms in groups:
                if all(basis.get(w, l) == l for w, l in word.items()):
                    basis.update(word)
                    terms.append((oi, c, mask))
                    break
            else:
                groups.append((dict(word), [(oi, c, mask)]))
    return groups


def _basis_rotations(basis: dict) -> List[Operation]:
    """Gates mapping each wire's measurement basis onto Z: X -> H, Y -> H S^dagger."""
    from jaqsi_amd import gateset as G

    ops: List[Operation] = []
    for w, l in sorted(basis.items()):
        if l == "X":
            ops.append(G.H(wires=w, record=False))
        elif l == "Y":
            ops.append(G.RZ(-np.pi / 2, wires=w, record=False))  # S^dagger up to a global phase
            ops.append(G.H(wires=w, record=False))
    return ops


class _ParityRunner:
    """One adjoint runner per basis group; observables combine the groups'
    Pauli-word expectation values with their coefficients."""

    def __init__(self, tape, n_qubits, groups, n_obs, batch, psi0_batched, dtype, base_low, mats):
        self.groups = []
        P0, E0 = len(base_low.angle_values), base_low.n_matrix_elems
        for basis, terms in groups:
            low = lowering.lower_pure(list(tape) + _basis_rotations(basis), n_qubits)
            plan = get_plan(low, dtype)
            # the rotations only append: angles [0, P0) and matrix elements
            # [0, E0) are the circuit's own, the rest are the rotation constants
            if len(low.angle_values) < P0 or low.n_matrix_elems < E0:
                raise RuntimeError("basis rotations changed the circuit's own parameter layout")
            extra_angles = torch.tensor([float(a) for a in low.angle_values[P0:]], dtype=torch.float64)
            extra_m = lowering.matrix_tensor(low.matrix_values, low.n_matrix_elems, None)
            extra_m = extra_m.reshape(-1)[E0:] if extra_m is not None else None
            if mats is not None:
                gm = torch.cat([mats.reshape(batch, -1).detach().to(torch.complex128).cpu(),
                                extra_m.reshape(1, -1).expand(batch, -1)], dim=1) if extra_m is not None and extra_m.numel() else mats.reshape(batch, -1)
            else:
                gm = extra_m.reshape(1, -1).expand(batch, -1) if extra_m is not None and extra_m.numel() else None
            words = np.zeros((len(terms), 2), dtype=np.int32)
            for t, (_, _, mask) in enumerate(terms):
                words[t, 0] = np.int32(np.uint32(mask & 0xFFFFFFFF).astype(np.int32))
                words[t, 1] = np.int32(np.uint32(mask >> 32).astype(np.int32))
            d = np.tile(np.array([1.0, -1.0]), len(terms))
            W = np.zeros((n_obs, len(terms)))  # expval_o = sum_t W[o, t] <P_t>
            for t, (oi, c, _) in enumerate(terms):
                W[oi, t] = c
            self.groups.append((plan, gm, words, d, torch.tensor(W, dtype=torch.float64), len(terms), extra_angles))
        self.P0 = P0
        self.batch = batch
        self.psi0_batched = psi0_batched
        self.dtype = dtype
        self.n_qubits = n_qubits
        self.n_obs = n_obs

    def _mats(self, mats, dev):
        if mats is None:
            return None
        return mats.reshape(self.batch, -1).to(dev, torch.complex128).contiguous()

    def _angles(self, angles, extra):
        if angles.shape[1] != self.P0:
            raise ValueError("angle vector does not match the circuit")
        if extra.numel() == 0:
            return angles
        return torch.cat([angles, extra.to(angles.device).reshape(1, -1).expand(self.batch, -1)], dim=1).contiguous()

    def forward(self, angles, psi0):
        dev = angles.device
        out = torch.zeros((self.batch, self.n_obs), dtype=torch.float64, device=dev)
        for plan, mats, words, d, W, nt, extra in self.groups:
            angles_g = self._angles(angles, extra)
            state = torch.empty((self.batch, 1 << self.n_qubits), dtype=_cdtype_torch(self.dtype), device=dev)
            ev = torch.empty((self.batch, nt), dtype=torch.float64, device=dev)
            m_g = self._mats(mats, dev)
            nat.check(nat.lib().jq_sv_forward(
                plan.handle, self.batch, nat.ptr(angles_g), nat.ptr(m_g), nat.ptr(psi0),
                int(self.psi0_batched), nat.ptr(state), nat.JQ_OBS_SV_PARITY, nt, nat.ptr(words), nat.ptr(d),
                nat.ptr(ev), _stream()))
            out += ev @ W.to(dev).T
        return out

    def adjoint(self, angles, psi0, ct):
        dev = angles.device
        grad = torch.zeros_like(angles)
        for plan, mats, words, d, W, nt, extra in self.groups:
            angles_g = self._angles(angles, extra)
            m_g = self._mats(mats, dev)
            ctp = (ct @ W.to(dev)).contiguous()  # cotangent of each Pauli word
            state = torch.empty((self.batch, 1 << self.n_qubits), dtype=_cdtype_torch(self.dtype), device=dev)
            ga = torch.empty_like(angles_g)
            nat.check(nat.lib().jq_sv_adjoint(
                plan.handle, self.batch, nat.ptr(angles_g), nat.ptr(m_g), nat.ptr(psi0),
                int(self.psi0_batched), nat.ptr(state), nat.JQ_OBS_SV_PARITY, nt, nat.ptr(words), nat.ptr(d),
                nat.ptr(ctp), None, nat.ptr(ga), None, None, _stream()))
            grad += ga[:, : self.P0]
        return grad


class _ParityExpval(torch.autograd.Function):
    """Expectation values of general observables, adjoint backward per basis group."""

    @staticmethod
    def forward(ctx, angles, psi0, runner):
        ctx.runner = runner
        ctx.save_for_backward(angles, psi0 if psi0 is not None else torch.empty(0))
        ctx.has_psi0 = psi0 is not None
        return runner.forward(angles, psi0)

    @staticmethod
    def backward(ctx, g):
        angles, psi0 = ctx.saved_tensors
        psi0 = psi0 if ctx.has_psi0 else None
        ga = ctx.runner.adjoint(angles, psi0, g.to(torch.float64).contiguous())
        return ga, None, None


def _obs_arrays(diags):
    wires = np.array([w for w, _, _ in diags], dtype=np.int32)
    d = np.array([[a, b] for _, a, b in diags], dtype=np.float64).reshape(-1)
    return wires, d


def _obs_matrix_f64(ob: Operation) -> np.ndarray:
    """The observable's matrix in complex128, rebuilt from float64 sources
    (class matrices, Hermitian's stored array, product factors) so that its
    Pauli coefficients do not inherit complex64 rounding."""
    factors = getattr(ob, "_factors", None)
    if factors:
        wires = list(ob.wires)
        m = np.eye(2 ** len(wires), dtype=complex)
        for f in factors:
            fw = list(f.wires)
            rest = [w for w in wires if w not in fw]
            k = len(wires)
            big = np.kron(_obs_matrix_f64(f), np.eye(2 ** len(rest), dtype=complex))  # wire order fw + rest
            order = fw + rest
            perm = [order.index(w) for w in wires]
            t = big.reshape((2,) * (2 * k)).transpose(perm + [k + q for q in perm])
            m = m @ t.reshape(2**k, 2**k)
        return m
    npm = getattr(ob, "_np_matrix", None)
    if npm is not None:
        return np.asarray(npm, dtype=complex)
    cm = getattr(ob.__class__, "_matrix", None)
    if cm is not None and not getattr(ob, "_explicit", False) and not getattr(ob, "parameters", None):
        return np.asarray(cm, dtype=complex)
    return _obs_matrix(ob)


def _obs_matrix(ob: Operation) -> np.ndarray:
    m = ob.matrix
    m = m.detach().cpu().numpy() if isinstance(m, torch.Tensor) else np.asarray(m)
    return np.ascontiguousarray(m.astype(np.complex128))


# ---------------------------------------------------------------------------
# the autograd bridge (the custom_vjp of simulation.py:586-654)
# ---------------------------------------------------------------------------
class _Runner:
    def __init__(self, plan, batch, density, diags, n_qubits, psi0_batched, dtype):
        self.plan = plan
        self.batch = batch
        self.density = density
        self.diags = diags
        self.n_qubits = n_qubits
        self.psi0_batched = psi0_batched
        self.dtype = dtype
        self.obs_kind = nat.JQ_OBS_DM_DIAG if density else nat.JQ_OBS_SV_DIAG
        self.wires, self.d = _obs_arrays(diags)

    def vec_len(self):
        return 1 << self.plan.n_wires

    def forward(self, angles, mats, psi0):
        dev = angles.device
        state = torch.empty((self.batch, self.vec_len()), dtype=_cdtype_torch(self.dtype), device=dev)
        ev = torch.empty((self.batch, len(self.diags)), dtype=torch.float64, device=dev)
        nat.check(nat.lib().jq_sv_forward(
            self.plan.handle, self.batch, nat.ptr(angles), nat.ptr(mats), nat.ptr(psi0), int(self.psi0_batched),
            nat.ptr(state), self.obs_kind, len(self.diags), nat.ptr(self.wires), nat.ptr(self.d), nat.ptr(ev),
            _stream()))
        return ev, state

    def adjoint(self, angles, mats, psi0, ct, want_psi0):
        dev = angles.device
        state = torch.empty((self.batch, self.vec_len()), dtype=_cdtype_torch(self.dtype), device=dev)
        ev = torch.empty((self.batch, len(self.diags)), dtype=torch.float64, device=dev)
        ga = torch.empty_like(angles)
        gm = torch.empty_like(mats) if mats is not None else None
        gp = torch.empty((self.batch, self.vec_len()), dtype=_cdtype_torch(self.dtype), device=dev) if want_psi0 else None
        nat.check(nat.lib().jq_sv_adjoint(
            self.plan.handle, self.batch, nat.ptr(angles), nat.ptr(mats), nat.ptr(psi0), int(self.psi0_batched),
            nat.ptr(state), self.obs_kind, len(self.diags), nat.ptr(self.wires), nat.ptr(self.d), nat.ptr(ct),
            nat.ptr(ev), nat.ptr(ga), nat.ptr(gm), nat.ptr(gp), _stream()))
        return ev, ga, gm, gp


class _AdjointExpval(torch.autograd.Function):
    """Expectation values whose backward is the native adjoint sweep."""

    @staticmethod
    def forward(ctx, angles, mats, psi0, runner):
        ev, _ = runner.forward(angles, mats, psi0)
        ctx.runner = runner
        ctx.save_for_backward(angles, mats if mats is not None else torch.empty(0), psi0 if psi0 is not None else torch.empty(0))
        ctx.has_mats = mats is not None
        ctx.has_psi0 = psi0 is not None
        return ev

    @staticmethod
    def backward(ctx, g):
        angles, mats, psi0 = ctx.saved_tensors
        mats = mats if ctx.has_mats else None
        psi0 = psi0 if ctx.has_psi0 else None
        r = ctx.runner
        ct = g.to(torch.float64).contiguous()
        want_psi0 = psi0 is not None and ctx.needs_input_grad[2]
        _, ga, gm, gp = r.adjoint(angles, mats, psi0, ct, want_psi0)
        # JAX-convention cotangents -> PyTorch's convention (conjugate)
        gm = gm.conj() if gm is not None else None
        if gp is not None:
            gp = gp.conj()
            if not r.psi0_batched:
                gp = gp.sum(0)
            gp = gp.reshape(psi0.shape)
        return ga, gm, gp, None


# ---------------------------------------------------------------------------
# the seam
# ---------------------------------------------------------------------------
def _prepare(tape, n_qubits, use_density, batch, initial_state, dtype, values=None):
    low = lowering.lower_mixed(tape, n_qubits) if use_density else lowering.lower_pure(tape, n_qubits)
    dev = device()
    if values is None:
        angles = lowering.angle_tensor(low.angle_values, batch)
        mats = lowering.matrix_tensor(low.matrix_values, low.n_matrix_elems, batch)
    else:
        angles, mats = values
    if batch is None:
        angles = angles.reshape(1, -1)
        mats = mats.reshape(1, -1) if mats is not None else None
    angles = angles.to(dev, torch.float64).contiguous()
    mats = mats.to(dev, torch.complex128).contiguous() if mats is not None else None
    psi0 = None
    psi0_batched = False
    if initial_state is not None:
        psi0 = torch.as_tensor(initial_state)
        psi0 = psi0.to(dev, _cdtype_torch(dtype))
        if use_density:
            # rho0 = |psi><psi| (simulation.py:467-469)
            p = psi0.reshape(-1, 1 << n_qubits)
            psi0 = (p[:, :, None] * p.conj()[:, None, :]).reshape(p.shape[0], -1)
        if psi0.dim() == 2 and psi0.shape[0] > 1:
            psi0_batched = True
        psi0 = psi0.reshape(psi0.shape[0] if psi0_batched else 1, -1).contiguous()
    plan = get_plan(low, dtype)
    return low, plan, angles, mats, psi0, psi0_batched


def simulate_and_measure(
    tape: List[Operation],
    n_qubits: int,
    type: str,
    obs: List[Operation],
    use_density: bool,
    shots: Optional[int] = None,
    key=None,
    initial_state=None,
    adjoint: bool = True,
    batch: Optional[int] = None,
    values=None,
):
    """Run simulation and measurement in one dispatch (simulation.py:658-732).

    ``batch``/``values`` carry a batched recording (Script with in_axes):
    ``values`` = (angles (B, P), matrices (B, E)) and results gain a leading
    batch axis.
    """
    if type not in ("expval", "probs", "state", "density"):
        raise ValueError(f"Unknown measurement type: {type!r}")
    if use_density and type == "state":
        raise ValueError(
            "Measurement type 'state' is not defined for mixed (noisy) circuits. Use 'density' instead."
        )
    tape = [op for op in tape if not isinstance(op, Barrier)]
    if any(getattr(op, "_pending", False) for op in tape):
        from jaqsi_amd.evolution import resolve_pending

        resolve_pending(tape)
    dtype = _dtype_code()
    low, plan, angles, mats, psi0, psi0_batched = _prepare(tape, n_qubits, use_density, batch, initial_state, dtype, values)
    B = angles.shape[0]
    rdt = torch.float64 if dtype else torch.float32

    def unbatch(x):
        return x if batch is not None else x[0]

    if type == "expval" and shots is None:
        diags = diag_observables(obs)
        if diags is not None and len(diags) > 0:
            runner = _Runner(plan, B, use_density, diags, n_qubits, psi0_batched, dtype)
            ev = _AdjointExpval.apply(angles, mats, psi0, runner)
            return unbatch(ev.to(rdt))
        # general observables: gradients through Pauli-word parity groups
        if any(t is not None and t.requires_grad for t in (angles, mats, psi0)) and torch.is_grad_enabled():
            if use_density:
                raise NotImplementedError(
                    "gradients of non-diagonal observables of density matrices are not supported yet"
                )
            if (mats is not None and mats.requires_grad) or (psi0 is not None and psi0.requires_grad):
                raise NotImplementedError(
                    "gradients of non-diagonal observables w.r.t. explicit matrices or initial states are not supported yet"
                )
            groups = pauli_groups(obs)
            runner = _ParityRunner(tape, n_qubits, groups, len(obs), B, psi0_batched, dtype, low, mats)
            ev = _ParityExpval.apply(angles, psi0, runner)
            return unbatch(ev.to(rdt))
        state = _forward_state(plan, B, angles, mats, psi0, psi0_batched, dtype)
        out = torch.zeros((B, len(obs)), dtype=torch.float64, device=angles.device)
        for i, ob in enumerate(obs):
            m = torch.as_tensor(_obs_matrix(ob)).to(angles.device)
            col = torch.zeros(B, dtype=torch.float64, device=angles.device)
            w = np.array(ob.wires, dtype=np.int32)
            nat.check(nat.lib().jq_expval_matrix(n_qubits, B, dtype, int(use_density), nat.ptr(state), len(ob.wires),
                                                 nat.ptr(w), nat.ptr(m), nat.ptr(col), _stream()))
            out[:, i] = col
        return unbatch(out.to(rdt))

    if any(t is not None and t.requires_grad for t in (angles, mats, psi0)) and torch.is_grad_enabled():
        raise NotImplementedError(
            f"gradients of {type!r} outputs are not supported by the GPU path; use type='expval'"
        )
    state = _forward_state(plan, B, angles, mats, psi0, psi0_batched, dtype)
    dim = 1 << n_qubits
    if shots is not None and type in ("probs", "expval"):
        probs = _probs(state, n_qubits, use_density, B, dtype)
        return unbatch(sample_shots(probs, n_qubits, type, obs, shots, key))
    if type == "state":
        return unbatch(state)
    if type == "probs":
        return unbatch(_probs(state, n_qubits, use_density, B, dtype))
    # density
    if use_density:
        return unbatch(state.reshape(B, dim, dim))
    rho = torch.empty((B, dim, dim), dtype=state.dtype, device=state.device)
    nat.check(nat.lib().jq_sv_outer(n_qubits, B, dtype, nat.ptr(state), nat.ptr(rho), _stream()))
    return unbatch(rho)


def _forward_state(plan, B, angles, mats, psi0, psi0_batched, dtype):
    state = torch.empty((B, 1 << plan.n_wires), dtype=_cdtype_torch(dtype), device=angles.device)
    nat.check(nat.lib().jq_sv_forward(plan.handle, B, nat.ptr(angles), nat.ptr(mats), nat.ptr(psi0), int(psi0_batched),
                                      nat.ptr(state), nat.JQ_OBS_NONE, 0, None, None, None, _stream()))
    return state


def _probs(state, n_qubits, density, B, dtype):
    dim = 1 << n_qubits
    out = torch.empty((B, dim), dtype=torch.float64 if dtype else torch.float32, device=state.device)
    if density:
        nat.check(nat.lib().jq_dm_diag(n_qubits, B, dtype, nat.ptr(state), nat.ptr(out), _stream()))
    else:
        nat.check(nat.lib().jq_sv_probs(n_qubits, B, dtype, nat.ptr(state), nat.ptr(out), _stream()))
    return out


def sample_shots(probs, n_qubits, type, obs, shots, key=None):
    """Shot sampling from exact probabilities (simulation.py:860-917).

    The reference draws with jax.random.choice; here torch.multinomial with a
    generator seeded from *key*, so samples differ (statistical parity only).
    """
    g = torch.Generator(device=probs.device)
    g.manual_seed(int(key) if key is not None else 0)
    p = probs.to(torch.float64)
    p = p / p.sum(-1, keepdim=True)
    samples = torch.multinomial(p, shots, replacement=True, generator=g)
    counts = torch.zeros_like(p)
    counts.scatter_add_(1, samples, torch.ones_like(samples, dtype=p.dtype))
    est = counts / shots
    if type == "probs":
        return est.to(probs.dtype)
    if type == "expval":
        cols = []
        for ob in obs:
            diag = torch.as_tensor(np.real(_lifted_diag(_obs_matrix(ob), ob.wires, n_qubits))).to(p.device)
            cols.append(est @ diag)
        return torch.stack(cols, dim=-1).to(probs.dtype)
    raise ValueError(f"Shot simulation is only supported for 'probs' and 'expval', got {type!r}.")


def _lifted_diag(m, wires, n):
    """Diagonal of a small observable lifted to n qubits."""
    idx = np.arange(1 << n)
    k = len(wires)
    r = np.zeros_like(idx)
    for j, w in enumerate(wires):
        r |= ((idx >> (n - 1 - w)) & 1) << (k - 1 - j)
    return np.diag(m)[r]


# ---------------------------------------------------------------------------
# thin entry points for a custom_vjp bridge on the reference side: the
# lowered structure is built once, angles and cotangents cross as tensors.
# ---------------------------------------------------------------------------
def lower_for_jax(tape: List[Operation], n_qubits: int, use_density: bool):
    """Lower a recorded tape; returns (lowered structure, float64 angle vector)."""
    tape = [op for op in tape if not isinstance(op, Barrier)]
    low = lowering.lower_mixed(tape, n_qubits) if use_density else lowering.lower_pure(tape, n_qubits)
    angles = np.array([float(a) for a in low.angle_values], dtype=np.float64)
    return (low, n_qubits, use_density), angles


def _runner_for(lowered, obs, batch=1):
    low, n_qubits, use_density = lowered
    dtype = _dtype_code()
    plan = get_plan(low, dtype)
    diags = diag_observables(obs)
    if not diags:
        raise NotImplementedError("the lowered entry points take diagonal single-wire observables")
    runner = _Runner(plan, batch, use_density, diags, n_qubits, False, dtype)
    mats = lowering.matrix_tensor(low.matrix_values, low.n_matrix_elems, None)
    if mats is not None:
        mats = mats.reshape(1, -1).to(device(), torch.complex128).contiguous()
    return runner, mats


def run_lowered(lowered, angles: torch.Tensor, obs) -> torch.Tensor:
    """Expectation values (n_obs,) of the lowered circuit at `angles` (float64)."""
    runner, mats = _runner_for(lowered, obs)
    a = angles.to(device(), torch.float64).reshape(1, -1).contiguous()
    ev, _ = runner.forward(a, mats, None)
    return ev[0]


def adjoint_lowered(lowered, angles: torch.Tensor, obs, ct: torch.Tensor) -> torch.Tensor:
    """Gradient of sum_o ct_o <O_o> w.r.t. the angle vector (float64)."""
    runner, mats = _runner_for(lowered, obs)
    a = angles.to(device(), torch.float64).reshape(1, -1).contiguous()
    c = ct.to(device(), torch.float64).reshape(1, -1).contiguous()
    _, ga, _, _ = runner.adjoint(a, mats, None, c, False)
    return ga[0]
