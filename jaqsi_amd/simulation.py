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

import ctypes
import threading
from contextlib import contextmanager
from typing import List, NamedTuple, Optional, Sequence, Tuple

import numpy as np
import torch

from jaqsi_amd import _native as nat
from jaqsi_amd import lowering
from jaqsi_amd.gateset import Barrier
from jaqsi_amd.noise import KrausChannel
from jaqsi_amd.operations import Operation, _as_tensor, x64_enabled

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


def get_plan(low: lowering.Lowered, dtype: int, tile_bits: int = 0, checkpoints: int = 1, batch: int = 1,
             density: bool = False, knobs: Optional[dict] = None, fold: bool = False) -> nat.Plan:
    """The cached plan of a lowered structure; its tuning knobs come from the
    autotuner's cache for (structure, dtype, batch) unless given.  ``fold``:
    the plan is only measured through diagonal statevector observables, so
    its trailing permutation / diagonal gates are measured through them
    instead of simulated (``fold_tail`` of the C-ABI)."""
    if knobs is None:
        from jaqsi_amd import autotune

        knobs = autotune.config_for(low, dtype, batch, density, fold=fold)
    lb, vf = int(knobs.get("low_bits", 0)), int(knobs.get("variant_flags", 0))
    tile_bits = int(knobs.get("tile_bits", tile_bits))
    fold = bool(fold) and not density
    key = (low.key, dtype, tile_bits, checkpoints, lb, vf, fold)
    with _plans_lock:
        p = _plans.get(key)
        if p is None:
            p = nat.Plan(low.n_wires, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, dtype,
                         tile_bits, checkpoints, lb, vf, fold_tail=fold)
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
        if len(ob.wires) != 1 or ob._explicit:
            return None
        d = _class_diag(ob.__class__)
        if d is None:
            return None
        out.append((int(ob.wires[0]), d[0], d[1]))
    return out


_class_diags: dict = {}


def _class_diag(cls) -> Optional[Tuple[float, float]]:
    """(d0, d1) of a class whose matrix is a diagonal 2x2, else None
    (decided once per class)."""
    try:
        return _class_diags[cls]
    except KeyError:
        pass
    m = cls._matrix
    d = None
    if m is not None:
        m = np.asarray(m)
        if m.shape == (2, 2) and np.allclose(m - np.diag(np.diag(m)), 0):
            dd = np.real(np.diag(m))
            d = (float(dd[0]), float(dd[1]))
    _class_diags[cls] = d
    return d


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

    def forward_ckpt(self, angles, mats, psi0):
        """Forward keeping every pass output in the plan: (expvals or None,
        final state, token for backward())."""
        dev = angles.device
        state = torch.empty((self.batch, self.vec_len()), dtype=_cdtype_torch(self.dtype), device=dev)
        nobs = len(self.diags)
        ev = torch.empty((self.batch, nobs), dtype=torch.float64, device=dev) if nobs else None
        tok = ctypes.c_int64(-1)
        nat.check(nat.lib().jq_sv_forward_ckpt(
            self.plan.handle, self.batch, nat.ptr(angles), nat.ptr(mats), nat.ptr(psi0), int(self.psi0_batched),
            nat.ptr(state), self.obs_kind if nobs else nat.JQ_OBS_NONE, nobs, nat.ptr(self.wires), nat.ptr(self.d),
            nat.ptr(ev), ctypes.byref(tok), _stream()))
        return ev, state, tok.value

    def backward(self, token, angles, mats, psi0, state, ct=None, lam=None, want_psi0=False):
        """Adjoint sweep over the checkpoints of forward_ckpt (the library
        re-runs the forward if the plan has been used since).  The cotangent
        is ``ct`` (per observable) or ``lam``, the JAX-convention cotangent of
        the final vector (overwritten)."""
        dev = angles.device
        ga = torch.empty_like(angles)
        gm = torch.empty_like(mats) if mats is not None else None
        gp = torch.empty((self.batch, self.vec_len()), dtype=_cdtype_torch(self.dtype), device=dev) if want_psi0 else None
        nobs = len(self.diags) if lam is None else 0
        nat.check(nat.lib().jq_sv_backward(
            self.plan.handle, token, self.batch, nat.ptr(angles), nat.ptr(mats), nat.ptr(psi0),
            int(self.psi0_batched), nat.ptr(state), self.obs_kind if nobs else nat.JQ_OBS_NONE, nobs,
            nat.ptr(self.wires), nat.ptr(self.d), nat.ptr(ct), nat.ptr(lam), nat.ptr(ga), nat.ptr(gm), nat.ptr(gp),
            _stream()))
        return ga, gm, gp

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


def _save(ctx, angles, mats, psi0, extra=()):
    ctx.save_for_backward(angles, mats if mats is not None else torch.empty(0),
                          psi0 if psi0 is not None else torch.empty(0), *extra)
    ctx.has_mats = mats is not None
    ctx.has_psi0 = psi0 is not None


def _saved(ctx):
    t = ctx.saved_tensors
    angles, mats, psi0 = t[0], t[1], t[2]
    return angles, (mats if ctx.has_mats else None), (psi0 if ctx.has_psi0 else None), t[3:]


def _to_torch_grads(r, psi0, gm, gp):
    """JAX-convention cotangents -> PyTorch's convention (conjugate)."""
    gm = gm.conj() if gm is not None else None
    if gp is not None:
        gp = gp.conj()
        if not r.psi0_batched:
            gp = gp.sum(0)
        gp = gp.reshape(psi0.shape)
    return gm, gp


class _AdjointExpval(torch.autograd.Function):
    """Expectation values whose backward is the native adjoint sweep over the
    checkpoints the forward left in the plan (the fwd/bwd residual pair of
    _adjoint_expval's custom_vjp, simulation.py:586-654)."""

    @staticmethod
    def forward(ctx, angles, mats, psi0, runner):
        ev, state, token = runner.forward_ckpt(angles, mats, psi0)
        ctx.runner = runner
        ctx.token = token
        _save(ctx, angles, mats, psi0, (state,))
        return ev

    @staticmethod
    def backward(ctx, g):
        angles, mats, psi0, (state,) = _saved(ctx)
        r = ctx.runner
        ct = g.to(torch.float64).contiguous()
        want_psi0 = psi0 is not None and ctx.needs_input_grad[2]
        ga, gm, gp = r.backward(ctx.token, angles, mats, psi0, state, ct=ct, want_psi0=want_psi0)
        gm, gp = _to_torch_grads(r, psi0, gm, gp)
        return ga, gm, gp, None


class _FinalState(torch.autograd.Function):
    """The final vector (state, or rho as a 2n-wire vector), differentiable:
    the backward runs the native adjoint seeded with the cotangent of the
    vector -- the reverse-mode autodiff the reference applies to probs /
    state / density outputs (simulation.py:658-732)."""

    @staticmethod
    def forward(ctx, angles, mats, psi0, runner):
        _, state, token = runner.forward_ckpt(angles, mats, psi0)
        ctx.runner = runner
        ctx.token = token
        _save(ctx, angles, mats, psi0, (state,))
        return state

    @staticmethod
    def backward(ctx, g):
        angles, mats, psi0, (state,) = _saved(ctx)
        r = ctx.runner
        # PyTorch's gradient of a real loss w.r.t. a complex vector is the
        # conjugate of JAX's cotangent
        # (a physical conjugate: g.conj() is a lazy view over unconjugated memory)
        lam = torch.conj_physical(g).to(state.dtype).contiguous()
        want_psi0 = psi0 is not None and ctx.needs_input_grad[2]
        ga, gm, gp = r.backward(ctx.token, angles, mats, psi0, state, lam=lam, want_psi0=want_psi0)
        gm, gp = _to_torch_grads(r, psi0, gm, gp)
        return ga, gm, gp, None


class _SvProbs(torch.autograd.Function):
    """|psi_i|^2 (native), backward 2 g_i psi_i (simulation.py:748-750)."""

    @staticmethod
    def forward(ctx, state, n_qubits, dtype):
        ctx.save_for_backward(state)
        return _probs(state, n_qubits, False, state.shape[0], dtype)

    @staticmethod
    def backward(ctx, g):
        (state,) = ctx.saved_tensors
        return 2 * g.to(state.real.dtype) * state, None, None


class _DmProbs(torch.autograd.Function):
    """Re diag(rho) (native), backward into the diagonal (simulation.py:838-840)."""

    @staticmethod
    def forward(ctx, rho_vec, n_qubits, dtype):
        ctx.dim = 1 << n_qubits
        return _probs(rho_vec, n_qubits, True, rho_vec.shape[0], dtype)

    @staticmethod
    def backward(ctx, g):
        dim = ctx.dim
        out = torch.zeros((g.shape[0], dim * dim), dtype=torch.complex128 if g.dtype == torch.float64 else torch.complex64,
                          device=g.device)
        out[:, torch.arange(dim, device=g.device) * (dim + 1)] = g.to(out.dtype)
        return out, None, None


class _SvOuter(torch.autograd.Function):
    """rho = psi psi^dagger (native), backward (G + G^dagger) psi (simulation.py:752-754)."""

    @staticmethod
    def forward(ctx, state, n_qubits, dtype):
        ctx.save_for_backward(state)
        dim = 1 << n_qubits
        B = state.shape[0]
        rho = torch.empty((B, dim, dim), dtype=state.dtype, device=state.device)
        nat.check(nat.lib().jq_sv_outer(n_qubits, B, dtype, nat.ptr(state), nat.ptr(rho), _stream()))
        return rho

    @staticmethod
    def backward(ctx, G):
        (state,) = ctx.saved_tensors
        G = G.to(state.dtype)
        H = G + G.conj().transpose(-1, -2)
        return torch.matmul(H, state.unsqueeze(-1)).squeeze(-1), None, None


def _lift_apply(m, wires, vec, n):
    """O (on `wires`) applied to each row of vec (B, 2^n)."""
    from jaqsi_amd.operations import _contract

    B = vec.shape[0]
    t = vec.reshape((B,) + (2,) * n)
    o = _as_tensor(m, vec.dtype).to(vec.device)
    return _contract(t, o, [w + 1 for w in wires]).reshape(B, -1)


class _SvExpvalMatrix(torch.autograd.Function):
    """Re <psi|O|psi> for one k-wire observable (native jq_expval_matrix),
    backward 2 g O psi (simulation.py:806-809)."""

    @staticmethod
    def forward(ctx, state, n_qubits, dtype, wires, m):
        B = state.shape[0]
        col = torch.zeros(B, dtype=torch.float64, device=state.device)
        w = np.array(wires, dtype=np.int32)
        mt = _as_tensor(m).to(state.device)
        nat.check(nat.lib().jq_expval_matrix(n_qubits, B, dtype, 0, nat.ptr(state), len(wires), nat.ptr(w),
                                             nat.ptr(mt), nat.ptr(col), _stream()))
        ctx.save_for_backward(state)
        ctx.n, ctx.wires, ctx.m = n_qubits, list(wires), m
        return col

    @staticmethod
    def backward(ctx, g):
        (state,) = ctx.saved_tensors
        # d Re<psi|O|psi> / d psi (PyTorch's convention) = (O + O^dagger) psi
        m = np.asarray(ctx.m, dtype=complex)
        h = m + m.conj().T
        return g.to(state.real.dtype)[:, None] * _lift_apply(h, ctx.wires, state, ctx.n), None, None, None, None


class _DmExpvalMatrix(torch.autograd.Function):
    """Re Tr(O rho) for one k-wire observable (native jq_expval_matrix),
    backward g * O lifted to every wire (simulation.py:850-852)."""

    @staticmethod
    def forward(ctx, rho_vec, n_qubits, dtype, wires, m):
        B = rho_vec.shape[0]
        col = torch.zeros(B, dtype=torch.float64, device=rho_vec.device)
        w = np.array(wires, dtype=np.int32)
        mt = _as_tensor(m).to(rho_vec.device)
        nat.check(nat.lib().jq_expval_matrix(n_qubits, B, dtype, 1, nat.ptr(rho_vec), len(wires), nat.ptr(w),
                                             nat.ptr(mt), nat.ptr(col), _stream()))
        ctx.n = n_qubits
        ctx.wires = list(wires)
        ctx.m = m
        ctx.cdt = rho_vec.dtype
        return col

    @staticmethod
    def backward(ctx, g):
        from jaqsi_amd.operations import _contract

        n, dim = ctx.n, 1 << ctx.n
        eye = torch.eye(dim, dtype=torch.complex128, device=g.device).reshape((2,) * (2 * n))
        # O lifted to all wires: O applied to the ket axes of the identity
        lifted = _contract(eye, _as_tensor(ctx.m, torch.complex128).to(g.device), ctx.wires)
        # d Re Tr(O rho) / d rho (PyTorch's convention): conj(O^T) = O^dagger-transpose;
        # for a Hermitian O this is O itself
        grad = lifted.reshape(dim, dim).transpose(0, 1).conj()
        out = g.to(torch.complex128)[:, None] * grad.reshape(1, -1)
        return out.to(ctx.cdt), None, None, None, None


# ---------------------------------------------------------------------------
# the seam
# ---------------------------------------------------------------------------
def _prepare(tape, n_qubits, use_density, batch, initial_state, dtype, values=None, fold=False):
    low = lowering.lower_mixed(tape, n_qubits) if use_density else lowering.lower_pure(tape, n_qubits)
    dev = device()
    if values is None:
        angles = lowering.angle_tensor(low.angle_values, batch)
        mats = lowering.constant_pool(low.matrix_values, dev) if batch is None else None
        if mats is None:
            mats = lowering.matrix_tensor(low.matrix_values, low.n_matrix_elems, batch)
    else:
        angles, mats = values
    if batch is None:
        angles = angles.reshape(1, -1)
        mats = mats.reshape(1, -1) if mats is not None else None
    angles = angles.to(dev, torch.float64).contiguous()
    # (resolve_conj: native code reads memory, not torch's lazy conjugate views)
    mats = mats.to(dev, torch.complex128).resolve_conj().contiguous() if mats is not None else None
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
        psi0 = psi0.reshape(psi0.shape[0] if psi0_batched else 1, -1).resolve_conj().contiguous()
    plan = get_plan(low, dtype, batch=angles.shape[0], density=use_density, fold=fold)
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
    from jaqsi_amd.evolution import PendingEvolution, resolve_pending

    if any(isinstance(op, PendingEvolution) for op in tape):
        resolve_pending(tape)
    dtype = _dtype_code()
    runs = getattr(_capture, "runs", None)
    # expectation values of diagonal single-wire observables measure the
    # circuit's trailing permutation / diagonal gates through the
    # observables (a captured run may be replayed for its final vector: no fold)
    diags = diag_observables(obs) if type == "expval" and shots is None else None
    fold = bool(diags) and not use_density and runs is None
    low, plan, angles, mats, psi0, psi0_batched = _prepare(tape, n_qubits, use_density, batch, initial_state, dtype,
                                                           values, fold)
    if runs is not None:
        runs.append(CapturedRun(plan, n_qubits, use_density, batch, dtype, angles, mats, psi0, psi0_batched))
    B = angles.shape[0]
    rdt = torch.float64 if dtype else torch.float32

    def unbatch(x):
        return x if batch is not None else x[0]

    needs_grad = torch.is_grad_enabled() and any(t is not None and t.requires_grad for t in (angles, mats, psi0))
    if type == "expval" and shots is None:
        if diags is not None and len(diags) > 0:
            runner = _Runner(plan, B, use_density, diags, n_qubits, psi0_batched, dtype)
            ev = _AdjointExpval.apply(angles, mats, psi0, runner)
            return unbatch(ev.to(rdt))
        if needs_grad:
            # general observables: one forward, then one adjoint sweep seeded
            # with the cotangent of the final vector
            runner = _Runner(plan, B, use_density, [], n_qubits, psi0_batched, dtype)
            vec = _FinalState.apply(angles, mats, psi0, runner)
            fn = _DmExpvalMatrix if use_density else _SvExpvalMatrix
            cols = [fn.apply(vec, n_qubits, dtype, list(ob.wires), _obs_matrix_f64(ob)) for ob in obs]
            out = torch.stack(cols, dim=-1) if cols else torch.zeros((B, 0), dtype=torch.float64, device=angles.device)
            return unbatch(out.to(rdt))
        state = _forward_state(plan, B, angles, mats, psi0, psi0_batched, dtype)
        out = torch.zeros((B, len(obs)), dtype=torch.float64, device=angles.device)
        for i, ob in enumerate(obs):
            m = _as_tensor(_obs_matrix(ob)).to(angles.device)
            col = torch.zeros(B, dtype=torch.float64, device=angles.device)
            w = np.array(ob.wires, dtype=np.int32)
            nat.check(nat.lib().jq_expval_matrix(n_qubits, B, dtype, int(use_density), nat.ptr(state), len(ob.wires),
                                                 nat.ptr(w), nat.ptr(m), nat.ptr(col), _stream()))
            out[:, i] = col
        return unbatch(out.to(rdt))

    if needs_grad and shots is None:
        runner = _Runner(plan, B, use_density, [], n_qubits, psi0_batched, dtype)
        vec = _FinalState.apply(angles, mats, psi0, runner)
        dim = 1 << n_qubits
        if type == "state":
            return unbatch(vec)
        if type == "probs":
            fn = _DmProbs if use_density else _SvProbs
            return unbatch(fn.apply(vec, n_qubits, dtype))
        if use_density:
            return unbatch(vec.reshape(B, dim, dim))
        return unbatch(_SvOuter.apply(vec, n_qubits, dtype))
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


_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def _mix64(z: np.ndarray) -> np.ndarray:
    """The splitmix64 finaliser (Steele, Lea & Flood, OOPSLA 2014): a
    bijective 64-bit mix, used here as a counter-based generator."""
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def shot_seed(key) -> np.uint64:
    """64-bit seed from a shot key: None (the reference's default key 0), an
    integer, or an array of key words (a JAX PRNG key array arrives as its
    uint32 words through np.asarray) folded in order."""
    if key is None:
        words = np.zeros(1, dtype=np.uint64)
    elif isinstance(key, torch.Tensor):
        words = key.detach().cpu().numpy().astype(np.int64).view(np.uint64).ravel()
    else:
        words = np.asarray(key).astype(np.int64).view(np.uint64).ravel()
    s = np.uint64(0)
    with np.errstate(over="ignore"):
        for w in words:
            s = _mix64(np.asarray(s ^ w + _GOLDEN))[()]
    return np.uint64(s)


def shot_uniforms(key, batch: int, shots: int) -> np.ndarray:
    """(batch, shots) uniforms in [0, 1): entry (b, s) is the mix of a
    per-row seed and the shot counter, so row b depends only on the key and
    b (the reference splits its key per batch element, script.py:628-629)."""
    seed = shot_seed(key)
    with np.errstate(over="ignore"):
        rows = _mix64(seed ^ _mix64(np.arange(batch, dtype=np.uint64) * _GOLDEN + _GOLDEN))
        ctr = (np.arange(shots, dtype=np.uint64) + np.uint64(1)) * _GOLDEN
        z = _mix64(rows[:, None] + ctr[None, :])
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def sample_shots(probs, n_qubits, type, obs, shots, key=None):
    """Shot sampling from exact probabilities (simulation.py:860-917).

    Each row of *probs* ((dim,) or (B, dim)) is sampled *shots* times by
    inverse-CDF lookup of counter-based uniforms (``shot_uniforms``); the
    reference draws with jax.random.choice, so individual samples differ and
    parity is statistical.  Returns estimated probabilities, or expectation
    values of the observables' diagonals under them (exact for diagonal
    observables, as in the reference)."""
    if type not in ("probs", "expval"):
        raise ValueError(f"Shot simulation is only supported for 'probs' and 'expval', got {type!r}.")
    if int(shots) <= 0:
        raise ValueError(f"shots must be a positive integer, got {shots!r}")
    p = probs.detach().to(torch.float64)
    squeeze = p.dim() == 1
    p = p.reshape(-1, p.shape[-1])
    B, dim = p.shape
    cdf = torch.cumsum(p.clamp_min(0.0), dim=-1)
    cdf = cdf / cdf[:, -1:]
    u = torch.as_tensor(shot_uniforms(key, B, int(shots)), device=p.device)
    samples = torch.searchsorted(cdf, u, right=True).clamp_(max=dim - 1)
    counts = torch.zeros_like(p)
    counts.scatter_add_(1, samples, torch.ones_like(u))
    est = counts / shots
    if type == "probs":
        out = est
    else:
        cols = []
        for ob in obs:
            diag = torch.as_tensor(np.real(_lifted_diag(_obs_matrix(ob), ob.wires, n_qubits))).to(p.device)
            cols.append(est @ diag)
        out = torch.stack(cols, dim=-1)
    out = out[0] if squeeze else out
    return out.to(probs.dtype if probs.dtype.is_floating_point else torch.float64)


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
# ---------------------------------------------------------------------------
# Jacobians of the final vector w.r.t. the recorded angles (quantum Fisher
# information, math.py:331-383)
# ---------------------------------------------------------------------------
class CapturedRun(NamedTuple):
    plan: object
    n_qubits: int
    use_density: bool
    batch: Optional[int]
    dtype: int
    angles: torch.Tensor
    mats: Optional[torch.Tensor]
    psi0: Optional[torch.Tensor]
    psi0_batched: bool


_capture = threading.local()


@contextmanager
def capture_runs():
    """Collect the lowered inputs (plan, angle vector with its autograd
    history, matrices, initial state) of every simulation run inside the
    block."""
    prev = getattr(_capture, "runs", None)
    _capture.runs = []
    try:
        yield _capture.runs
    finally:
        _capture.runs = prev


def angle_jacobian(run: CapturedRun) -> Tuple[torch.Tensor, torch.Tensor]:
    """(final vector (L,), J (L, P)): J[k, j] = d vec_k / d angle_j of one
    captured circuit, complex128.

    One batched forward of 2L copies of the circuit, then one adjoint sweep
    whose copy k is seeded with the cotangent of Re vec_k (copies 0..L-1) or
    Im vec_k (copies L..2L-1): the per-copy angle gradients are the real and
    imaginary parts of J's rows."""
    if run.batch is not None:
        raise ValueError("the Jacobian is taken of one circuit, not a batch")
    L = 1 << run.plan.n_wires
    B = 2 * L
    dev = run.angles.device
    a = run.angles.detach().reshape(1, -1).expand(B, -1).contiguous()
    m = run.mats.detach().reshape(1, -1).expand(B, -1).contiguous() if run.mats is not None else None
    p0 = run.psi0.detach() if run.psi0 is not None else None
    if p0 is not None and run.psi0_batched:
        raise ValueError("the Jacobian is taken of one circuit, not a batch")
    runner = _Runner(run.plan, B, run.use_density, [], run.n_qubits, False, run.dtype)
    _, state, tok = runner.forward_ckpt(a, m, p0)
    final = state[0].to(torch.complex128, copy=True)  # the sweep below runs the state buffer back to the input
    # native cotangents are conj(torch's): e_k seeds Re vec_k, -i e_k seeds Im vec_k
    lam = torch.zeros((B, L), dtype=_cdtype_torch(run.dtype), device=dev)
    idx = torch.arange(L, device=dev)
    lam[idx, idx] = 1.0
    lam[idx + L, idx] = -1j
    ga, _, _ = runner.backward(tok, a, m, p0, state, lam=lam)
    J = torch.complex(ga[:L], ga[L:])
    return final, J


def lower_for_jax(tape: List[Operation], n_qubits: int, use_density: bool):
    """(lowered structure, angle parameters in plan-slot order, unconverted);
    see ``jaqsi_amd.bridge``."""
    from jaqsi_amd.bridge import lower_for_jax as _lower

    return _lower(tape, n_qubits, use_density)


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
