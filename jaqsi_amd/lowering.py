"""Tape -> compact native op list.

The recorded tape (list of ``Operation``) is lowered to the ``jq_op`` /
``jq_member`` arrays of the C-ABI plus two value vectors: the angles of
every parametrised factor and a complex128 pool of explicit matrices.  The
arrays depend only on the circuit *structure*, so a native plan is cached
per structure key and reused with new values (the role of the reference's
compile cache, jaqsi/script.py:595-705).

Density-matrix lowering follows _compile_mixed (jaqsi/simulation.py:348-381):
the vector has 2n wires (ket 0..n-1, bra n..2n-1); a single-qubit gate
becomes its superoperator U (x) conj(U) on (w, w+n) so that it fuses with
the channels recorded on that wire, a wider or permutation gate stays
two-sided (U on the ket wires, conj(U) on the bra wires), and a channel
becomes sum_k K (x) conj(K).
"""

from __future__ import annotations

from typing import List, NamedTuple, Optional, Sequence, Tuple

import numpy as np
import torch

from jaqsi_amd import _native as nat
from jaqsi_amd.gateset import Barrier
from jaqsi_amd.noise import KrausChannel
from jaqsi_amd.operations import Member, Operation

_CODES = {"I": 0, "X": 1, "Y": 2, "Z": 3}


def word_code(word: str) -> int:
    w = 0
    for c in word:
        w = (w << 2) | _CODES[c]
    return w


class Lowered(NamedTuple):
    key: tuple                # structure key (for plan caching)
    n_wires: int
    ops: list                 # nat.Op
    members: list             # nat.Member
    angle_values: list        # one entry per angle slot: float / tensor (0-d or (B,))
    matrix_values: list       # (offset, value) explicit matrices (numpy / tensor)
    n_matrix_elems: int
    grad_matrices: list       # indices into matrix_values that require grad


def _structure_flags(op: Operation) -> Tuple[int, int, tuple]:
    s = op._structure
    flags = 0
    if s.diagonal:
        flags |= nat.JQ_S_DIAGONAL
    perm = tuple(s.permutation)
    if perm:
        flags |= nat.JQ_S_PERMUTATION
    return flags, int(s.controls), perm


def _requires_grad(v) -> bool:
    return isinstance(v, torch.Tensor) and v.requires_grad


class _Builder:
    def __init__(self, n_wires: int):
        self.n_wires = n_wires
        self.ops: List[nat.Op] = []
        self.members: List[nat.Member] = []
        self.angles: list = []
        self.mats: list = []
        self.n_mat = 0
        self.grad_mats: list = []
        self.key: list = []

    def add_member(self, m: Member, extra_flags: int = 0) -> None:
        flags = extra_flags
        param, mat = -1, -1
        if m.kind in ("rot", "phase"):
            param = len(self.angles)
            self.angles.append(m.param)
            kind = nat.JQ_M_ROT if m.kind == "rot" else nat.JQ_M_PHASE
            word = word_code(m.word) if m.kind == "rot" else 0
            self.key.append((kind, word, m.nctrl, flags))
        else:
            kind, word = nat.JQ_M_MATRIX, 0
            val = m.matrix
            shape = tuple(val.shape[-2:])
            mat = self.n_mat
            self.n_mat += shape[0] * shape[1]
            if _requires_grad(val):
                flags |= nat.JQ_MF_GRAD
                self.grad_mats.append(len(self.mats))
            self.mats.append((mat, val))
            self.key.append((kind, shape, m.nctrl, flags))
        self.members.append(nat.Member(kind, flags, m.nctrl, word, param, mat, 0.0))

    def add_op(self, wires: Sequence[int], members: Sequence[Member], structure: int, controls: int, perm: tuple,
               extra_flags: int = 0, superop_matrix: Optional[np.ndarray] = None, cond: float = 1.0) -> None:
        if len(wires) > nat.MAX_WIRES:
            raise NotImplementedError(f"gates on more than {nat.MAX_WIRES} wires are not supported")
        op = nat.Op()
        op.cond = cond
        op.nwires = len(wires)
        for j, w in enumerate(wires):
            op.wires[j] = int(w)
        op.structure = structure
        op.controls = controls
        for j, p in enumerate(perm):
            op.perm[j] = int(p)
        op.member_begin = len(self.members)
        if superop_matrix is not None:
            self.add_member(Member("matrix", matrix=superop_matrix))
        else:
            for m in members:
                self.add_member(m, extra_flags)
        op.member_end = len(self.members)
        self.ops.append(op)
        # the planner routes ill-conditioned blocks (product of their channels'
        # condition numbers above JQ_COND_MAX) to checkpointed passes, so a
        # channel's condition number is part of the structure
        kc = float(cond) if structure & nat.JQ_S_NONUNITARY else 1.0
        self.key.append(("op", tuple(wires), structure, controls, perm, kc))

    def result(self) -> Lowered:
        return Lowered(tuple(self.key), self.n_wires, self.ops, self.members, self.angles, self.mats, self.n_mat,
                       self.grad_mats)


def lower_pure(tape: Sequence[Operation], n_qubits: int) -> Lowered:
    """Statevector lowering: one jq_op per recorded gate."""
    b = _Builder(n_qubits)
    b.key.append(("sv", n_qubits))
    for op in tape:
        if isinstance(op, Barrier):
            continue
        flags, controls, perm = _structure_flags(op)
        b.add_op(op.wires, op.lower(), flags, controls, perm)
    return b.result()


def lower_mixed(tape: Sequence[Operation], n_qubits: int) -> Lowered:
    """Density-matrix lowering on 2n wires (simulation.py:348-381)."""
    n = n_qubits
    b = _Builder(2 * n)
    b.key.append(("dm", n))
    for op in tape:
        if isinstance(op, Barrier):
            continue
        ket = list(op.wires)
        bra = [w + n for w in ket]
        if isinstance(op, KrausChannel):
            sup = op.superoperator()
            b.add_op(ket + bra, [], nat.JQ_S_NONUNITARY, 0, (), superop_matrix=sup, cond=channel_cond(sup))
            continue
        flags, controls, perm = _structure_flags(op)
        if len(ket) == 1 and not perm:
            sflags = nat.JQ_S_DIAGONAL if (flags & nat.JQ_S_DIAGONAL) else 0
            b.add_op(ket + bra, op.lower(), sflags, 0, (), extra_flags=nat.JQ_MF_SUPEROP)
        else:
            b.add_op(ket, op.lower(), flags, controls, perm)
            b.add_op(bra, op.lower(), flags, controls, perm, extra_flags=nat.JQ_MF_CONJ)
    return b.result()


def channel_cond(sup) -> float:
    """2-norm condition number of a channel superoperator (inf if singular
    or if it is a tensor that requires grad: its value may change between
    calls, so it is never inverted)."""
    if isinstance(sup, torch.Tensor):
        if sup.requires_grad:
            return float("inf")
        sup = sup.detach().cpu().numpy()
    s = np.linalg.svd(np.asarray(sup, dtype=np.complex128), compute_uv=False)
    return float(s[0] / s[-1]) if s[-1] > 0 else float("inf")


def angle_tensor(values: list, batch: Optional[int]) -> torch.Tensor:
    """Stack the angle values into (P,) or (B, P) float64, keeping autograd.

    Constant angles (Python numbers, host tensors) join the device of the
    first device tensor among the values, so a tape may mix both."""
    dev = next((v.device for v in values if isinstance(v, torch.Tensor) and v.device.type != "cpu"), None)
    cols = []
    for v in values:
        if isinstance(v, torch.Tensor):
            t = v.to(torch.float64)
        else:
            t = torch.as_tensor(np.asarray(v, dtype=np.float64))
        if dev is not None and t.device != dev:
            t = t.to(dev)
        if batch is not None:
            t = t.expand(batch) if t.dim() == 0 else t.reshape(batch)
        cols.append(t)
    if not cols:
        return torch.zeros((batch, 0) if batch is not None else (0,), dtype=torch.float64)
    return torch.stack(cols, dim=-1)


def matrix_tensor(mats: list, n_elems: int, batch: Optional[int]) -> Optional[torch.Tensor]:
    """Pack explicit matrices into (E,) or (B, E) complex128, keeping autograd."""
    if not mats:
        return None
    parts = []
    for off, v in mats:
        t = v if isinstance(v, torch.Tensor) else torch.as_tensor(np.asarray(v))
        t = t.to(torch.complex128)
        if batch is not None:
            t = t.expand(batch, *t.shape[-2:]) if t.dim() == 2 else t
            parts.append(t.reshape(batch, -1))
        else:
            parts.append(t.reshape(-1))
    return torch.cat(parts, dim=-1)
