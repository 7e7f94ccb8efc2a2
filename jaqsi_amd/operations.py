"""Operation IR: the base class of every gate, observable and channel.

Mirrors jaqsi/operations.py:176-719.  An ``Operation`` validates its wires,
appends itself to the active tape (operations.py:229-273) and carries the
static ``GateStructure`` tags the lowering reads.  Matrices are torch
tensors, so gate parameters may be tensors that require grad; the device
path never evaluates ``.matrix`` for the built-in gate kinds but lowers them
to (kind, parameter) pairs instead (``lower()``).

Bit order: wire 0 is the most significant bit of the flat index.
"""

from __future__ import annotations

import os
import warnings
from typing import Callable, List, NamedTuple, Optional, Tuple, Union

import numpy as np
import torch

from jaqsi_amd.tape import _local as _tape_local
from jaqsi_amd.tape import active_tape, recording  # noqa: F401  (re-export)

_X64 = os.environ.get("JAQSI_AMD_X64", "0") not in ("0", "", "false", "False")


def enable_x64(flag: bool = True) -> None:
    """Select complex128 (True) or complex64 (False) simulation.

    The reference switches with ``jax_enable_x64`` (operations.py:26-43).
    """
    global _X64
    _X64 = bool(flag)


def x64_enabled() -> bool:
    return _X64


def cdtype() -> torch.dtype:
    """Active complex dtype: complex128 with x64, else complex64."""
    return torch.complex128 if _X64 else torch.complex64


def rdtype() -> torch.dtype:
    return torch.float64 if _X64 else torch.float32


class GateStructure(NamedTuple):
    """Static matrix guarantees (operations.py:176-185): ``permutation`` maps
    output rows to input columns, ``diagonal`` and the number of leading
    ``controls`` (active when all are one)."""

    permutation: Tuple[int, ...] = ()
    diagonal: bool = False
    controls: int = 0


class Member(NamedTuple):
    """One primitive factor of a lowered gate.

    kind:  "rot"    -> controlled(cos(t/2) I - i sin(t/2) P_word, nctrl)
           "phase"  -> controlled(diag(1, e^{i t}), nctrl)
           "matrix" -> an explicit matrix (value may require grad)
    """

    kind: str
    param: object = None  # float / 0-d tensor / (B,) tensor
    word: str = ""
    nctrl: int = 0
    matrix: object = None  # numpy or torch (d, d) / (B, d, d)


def _as_tensor(x, dtype=None):
    if isinstance(x, torch.Tensor):
        return x if dtype is None else x.to(dtype)
    a = np.asarray(x)
    if not a.flags.writeable:  # shared read-only constants (class gate matrices)
        a = a.copy()
    return torch.as_tensor(a, dtype=dtype)


def _embed_matrix(mat: torch.Tensor, op_wires: list, all_wires: list) -> torch.Tensor:
    """Tensor *mat* with identities on the missing wires, then reorder the
    qubits to *all_wires* (operations.py:107-148)."""
    n = len(all_wires)
    if list(op_wires) == list(all_wires):
        return mat
    missing = [w for w in all_wires if w not in op_wires]
    full = mat
    for _ in missing:
        full = torch.kron(full, torch.eye(2, dtype=mat.dtype))
    order = list(op_wires) + missing
    perm = [order.index(w) for w in all_wires]
    t = full.reshape([2] * (2 * n)).permute(perm + [p + n for p in perm])
    return t.reshape(2**n, 2**n)


def _contract(tensor: torch.Tensor, gate: torch.Tensor, wires: List[int]) -> torch.Tensor:
    """Contract a (2^k, 2^k) *gate* into the *wires* axes of a rank-N tensor
    (operations.py:46-104)."""
    k = len(wires)
    gt = gate.reshape((2,) * (2 * k))
    out = torch.tensordot(gt, tensor, dims=(list(range(k, 2 * k)), list(wires)))
    return torch.movedim(out, list(range(k)), list(wires))


class Operation:
    """Base class for any quantum operation or observable
    (operations.py:188-644)."""

    is_controlled = False
    _structure = GateStructure()
    is_clifford = False
    is_unitary = False
    _matrix = None
    _explicit = False  # True on instances built from a given matrix
    _num_wires: Optional[int] = None
    _param_names: Tuple[str, ...] = ()

    def __init__(
        self,
        wires: Union[int, List[int]] = 0,
        matrix=None,
        record: bool = True,
        name: Optional[str] = None,
    ) -> None:
        self.name = name or self.__class__.__name__
        ws = list(wires) if isinstance(wires, (list, tuple)) else [wires]
        self._wires = ws
        if self._num_wires is not None and len(ws) != self._num_wires:
            raise ValueError(f"{self.name} expects {self._num_wires} wire(s), got {len(ws)}: {ws}")
        if len(ws) > 1 and len(ws) != len(set(ws)):
            raise ValueError(f"{self.name} received duplicate wires: {ws}")
        if matrix is not None:
            self._matrix = matrix
            self._structure = GateStructure()
            self._explicit = True
        if record:
            stack = getattr(_tape_local, "stack", None)
            if stack:
                stack[-1].append(self)

    # -- introspection -----------------------------------------------------
    @property
    def parameters(self) -> list:
        return [getattr(self, n) for n in self._param_names]

    def __repr__(self) -> str:
        params = self.parameters
        if params:
            ps = ", ".join(f"{float(v):.4f}" if np.ndim(_detach(v)) == 0 else str(v) for v in params)
            return f"{self.name}({ps}, wires={self.wires})"
        return f"{self.name}(wires={self.wires})"

    @property
    def wires(self) -> List[int]:
        return self._wires

    @wires.setter
    def wires(self, wires) -> None:
        self._wires = list(wires) if isinstance(wires, (list, tuple)) else [wires]

    @property
    def matrix(self) -> torch.Tensor:
        """The gate matrix in the active complex dtype (operations.py:307-326)."""
        if self._matrix is None:
            raise NotImplementedError(f"{self.__class__.__name__} does not define a matrix.")
        return _as_tensor(self._matrix).to(cdtype())

    def lower(self) -> List[Member]:
        """Primitive factors of this gate in application order.

        The default is the explicit matrix; built-in parametrised gates
        override this with (kind, parameter) members so the device computes
        their matrices and derivatives itself.
        """
        m = self._matrix
        if m is None:
            raise NotImplementedError(f"{self.__class__.__name__} does not define a matrix.")
        return [Member("matrix", matrix=m)]

    def decompose(self) -> List["Operation"]:
        raise NotImplementedError(f"{self.__class__.__name__} does not define a decomposition.")

    # -- algebra (operations.py:387-532) ----------------------------------
    def _update_tape_operation(self, op: "Operation") -> None:
        tape = active_tape()
        if tape is not None:
            if tape and tape[-1] is self:
                tape[-1] = op
            else:
                tape.append(op)

    def dagger(self) -> "Operation":
        mat = self.matrix.conj().transpose(-1, -2)
        op = Operation(wires=self.wires, matrix=mat, record=False)
        op.is_unitary = self.is_unitary
        perm = self._structure.permutation
        if perm:
            perm = tuple(perm.index(i) for i in range(len(perm)))
        op._structure = self._structure._replace(permutation=perm)
        self._update_tape_operation(op)
        return op

    def power(self, power: int) -> "Operation":
        mat = torch.linalg.matrix_power(self.matrix, power)
        op = Operation(wires=self.wires, matrix=mat, record=False)
        op.is_unitary = self.is_unitary
        self._update_tape_operation(op)
        return op

    def __mul__(self, other):
        if isinstance(other, Operation):
            return self.__matmul__(other)
        op = Operation(wires=self.wires, matrix=other * self.matrix, record=False)
        self._update_tape_operation(op)
        return op

    __rmul__ = __mul__

    def __add__(self, other: "Operation") -> "Operation":
        if sorted(self.wires) != sorted(other.wires):
            raise ValueError(
                f"Can only add operations acting on the same set of wires, got {self.wires} and {other.wires}"
            )
        return Operation(wires=self.wires, matrix=self.matrix + other.matrix, record=False)

    def prod(self, *ops: "Operation") -> "Operation":
        if not ops:
            return self
        all_ops = (self,) + ops
        all_wires: List[int] = []
        for o in all_ops:
            for w in o.wires:
                if w not in all_wires:
                    all_wires.append(w)
        mat = _embed_matrix(all_ops[0].matrix, all_ops[0].wires, all_wires)
        for o in all_ops[1:]:
            mat = mat @ _embed_matrix(o.matrix, o.wires, all_wires)
        names = "*".join(o.name for o in all_ops)
        op = Operation(wires=all_wires, matrix=mat, name=f"Prod({names})", record=False)
        op.is_unitary = all(o.is_unitary for o in all_ops)
        op._factors = all_ops  # kept for float64 re-evaluation of observables
        return op

    def __matmul__(self, other):
        if not isinstance(other, Operation):
            return NotImplemented
        return self.prod(other)

    # -- dense helpers (operations.py:534-644); off the device path ---------
    def lifted_matrix(self, n_qubits: int) -> torch.Tensor:
        dim = 2**n_qubits
        eye = torch.eye(dim, dtype=cdtype())
        cols = [self.apply_to_state(eye[j], n_qubits) for j in range(dim)]
        return torch.stack(cols, dim=1)

    def apply_to_state(self, state: torch.Tensor, n_qubits: int) -> torch.Tensor:
        psi = _as_tensor(state).to(cdtype()).reshape((2,) * n_qubits)
        return _contract(psi, self.matrix, self.wires).reshape(2**n_qubits)

    def apply_to_density(self, rho: torch.Tensor, n_qubits: int) -> torch.Tensor:
        u = self.matrix
        t = _as_tensor(rho).to(cdtype()).reshape((2,) * (2 * n_qubits))
        t = _contract(t, u, self.wires)
        t = _contract(t, u.conj(), [w + n_qubits for w in self.wires])
        return t.reshape(2**n_qubits, 2**n_qubits)


def _detach(v):
    return v.detach().cpu().numpy() if isinstance(v, torch.Tensor) else v


class Hermitian(Operation):
    """A Hermitian observable/gate from an arbitrary matrix
    (operations.py:647-719)."""

    def __init__(self, matrix, wires: Union[int, List[int]] = 0, record: bool = True) -> None:
        super().__init__(wires=wires, matrix=_as_tensor(matrix).to(torch.complex128), record=record)
        self._np_matrix = None if isinstance(matrix, torch.Tensor) and matrix.requires_grad else np.asarray(_detach(matrix), dtype=complex)

    def __rmul__(self, coeff_fn: Callable):
        if not callable(coeff_fn):
            raise TypeError(f"Left operand of `* Hermitian` must be callable, got {type(coeff_fn)}")
        from jaqsi_amd.evolution import ParametrizedHamiltonian

        H = self._np_matrix if self._np_matrix is not None else self.matrix
        return ParametrizedHamiltonian(terms=[(coeff_fn, H, self.wires)])

    def evolve(self, name: Optional[str] = None, **odeint_kwargs) -> Callable:
        """Gate factory ``(t, wires=0) -> Operation`` of exp(-i t H)
        (operations.py:704-719)."""
        from jaqsi_amd.evolution import Evolution

        return Evolution.evolve(self, name=name, **odeint_kwargs)


def __getattr__(name):  # ParametrizedHamiltonian lives with its solver
    if name == "ParametrizedHamiltonian":
        from jaqsi_amd.evolution import ParametrizedHamiltonian

        return ParametrizedHamiltonian
    raise AttributeError(name)
