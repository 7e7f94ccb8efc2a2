"""NumPy complex128 restatement of the reference statevector engine.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates, in plain NumPy:

* gate matrices and structure tags          -- jaqsi/gateset.py:32-890
* rank-n tensor gate application            -- jaqsi/simulation.py:70-112,
                                               jaqsi/operations.py:46-77
* same-wire fusion (``_fuse``)              -- jaqsi/simulation.py:214-240
* the reversed-plan adjoint VJP             -- jaqsi/simulation.py:546-655
* diagonal / general expectation values     -- jaqsi/simulation.py:735-811

Bit order: wire 0 is the most significant bit of the flat amplitude index
(jaqsi/simulation.py:794-796).

How it is pinned (tests/test_oracle_sv.py): every gate matrix against its
textbook closed form and unitarity; application against brute-force
kron-lifted dense matrices; expectation values against closed forms
(RX|0> -> cos, Bell/GHZ); adjoint gradients against central finite
differences and the exact parameter-shift rule; and the full value+gradient
path against the golden Adam trajectories of tests/test_training.py:54-106.
"""

from __future__ import annotations

import math
from functools import reduce
from typing import List, NamedTuple, Optional, Sequence, Tuple

import numpy as np

C = np.complex128
I2 = np.eye(2, dtype=C)
PX = np.array([[0, 1], [1, 0]], dtype=C)
PY = np.array([[0, -1j], [1j, 0]], dtype=C)
PZ = np.array([[1, 0], [0, -1]], dtype=C)
HAD = np.array([[1, 1], [1, -1]], dtype=C) / math.sqrt(2)
PAULI = {"I": I2, "X": PX, "Y": PY, "Z": PZ}
P0 = np.array([[1, 0], [0, 0]], dtype=C)
P1 = np.array([[0, 0], [0, 1]], dtype=C)


def pauli_tensor(word: str) -> np.ndarray:
    """Kron of single-qubit Paulis, first letter = first wire (gateset.py:392-394)."""
    return reduce(np.kron, [PAULI[c] for c in word])


def rot(theta: float, P: np.ndarray) -> np.ndarray:
    """cos(t/2) I - i sin(t/2) P  (gateset.py:399-405)."""
    d = P.shape[0]
    return math.cos(theta / 2) * np.eye(d, dtype=C) - 1j * math.sin(theta / 2) * P


def drot(theta: float, P: np.ndarray) -> np.ndarray:
    """d/dtheta of :func:`rot`."""
    d = P.shape[0]
    return -0.5 * math.sin(theta / 2) * np.eye(d, dtype=C) - 0.5j * math.cos(theta / 2) * P


def controlled(block: np.ndarray, n_controls: int) -> np.ndarray:
    """Identity with *block* in the all-controls-one corner (gateset.py:801-806)."""
    dt = block.shape[0]
    dim = (2**n_controls) * dt
    m = np.eye(dim, dtype=C)
    s = dim - dt
    m[s:, s:] = block
    return m


def dcontrolled(dblock: np.ndarray, n_controls: int) -> np.ndarray:
    dt = dblock.shape[0]
    dim = (2**n_controls) * dt
    m = np.zeros((dim, dim), dtype=C)
    s = dim - dt
    m[s:, s:] = dblock
    return m


class Op(NamedTuple):
    """One recorded gate: ``name``, ``wires``, real ``params`` and ``extra``.

    ``extra`` holds the Pauli word (PauliRot / ControlledPauliRot), the
    number of controls (ControlledPauliRot), the diagonal
    (DiagonalQubitUnitary) or an explicit matrix (``Matrix``).
    """

    name: str
    wires: Tuple[int, ...]
    params: Tuple[float, ...] = ()
    extra: object = None


class Structure(NamedTuple):
    """Static guarantees (operations.py:176-185): permutation rows->cols,
    diagonal flag, number of leading control wires."""

    permutation: Tuple[int, ...] = ()
    diagonal: bool = False
    controls: int = 0


_FIXED = {
    "Id": (I2, Structure(diagonal=True)),
    "PauliX": (PX, Structure(permutation=(1, 0))),
    "PauliY": (PY, Structure()),
    "PauliZ": (PZ, Structure(diagonal=True)),
    "H": (HAD, Structure()),
    "S": (np.diag([1, 1j]).astype(C), Structure(diagonal=True)),
    "SWAP": (
        np.eye(4, dtype=C)[[0, 2, 1, 3]],
        Structure(permutation=(0, 2, 1, 3)),
    ),
    "CX": (np.kron(P0, I2) + np.kron(P1, PX), Structure((0, 1, 3, 2), False, 1)),
    "CY": (np.kron(P0, I2) + np.kron(P1, PY), Structure((), False, 1)),
    "CZ": (np.kron(P0, I2) + np.kron(P1, PZ), Structure((), True, 1)),
    "CCX": (
        np.eye(8, dtype=C)[[0, 1, 2, 3, 4, 5, 7, 6]],
        Structure((0, 1, 2, 3, 4, 5, 7, 6), False, 2),
    ),
    "CSWAP": (
        np.eye(8, dtype=C)[[0, 1, 2, 3, 4, 6, 5, 7]],
        Structure((0, 1, 2, 3, 4, 6, 5, 7), False, 1),
    ),
}

_ROT_AXIS = {"RX": "X", "RY": "Y", "RZ": "Z"}
_WORD_GATES = {"RXX": "XX", "RYY": "YY", "RZZ": "ZZ", "RZX": "ZX"}
_CROT = {"CRX": "X", "CRY": "Y", "CRZ": "Z"}


def op_matrix(op: Op) -> np.ndarray:
    """The gate matrix, row = output, MSB = first wire (gateset.py)."""
    name, p = op.name, op.params
    if name in _FIXED:
        m = _FIXED[name][0]
        if name == "Id" and len(op.wires) > 1:
            return np.eye(2 ** len(op.wires), dtype=C)
        return m.copy()
    if name in _ROT_AXIS:  # gateset.py:408-451
        return rot(p[0], PAULI[_ROT_AXIS[name]])
    if name in _WORD_GATES:  # gateset.py:736-739
        return rot(p[0], pauli_tensor(_WORD_GATES[name]))
    if name == "PauliRot":  # gateset.py:670-717
        return rot(p[0], pauli_tensor(op.extra))
    if name in _CROT:  # gateset.py:864-890
        return controlled(rot(p[0], PAULI[_CROT[name]]), 1)
    if name == "ControlledPauliRot":  # gateset.py:764-818
        word, nc = op.extra
        return controlled(rot(p[0], pauli_tensor(word)), nc)
    if name in ("CPhase", "ControlledPhaseShift"):  # gateset.py:587-619
        return np.diag([1, 1, 1, np.exp(1j * p[0])]).astype(C)
    if name == "Rot":  # gateset.py:622-667: RZ(omega) RY(theta) RZ(phi)
        phi, theta, omega = p
        return rot(omega, PZ) @ rot(theta, PY) @ rot(phi, PZ)
    if name == "DiagonalQubitUnitary":
        return np.diag(np.asarray(op.extra, dtype=C))
    if name == "Matrix":
        return np.asarray(op.extra, dtype=C)
    raise ValueError(f"unknown gate {name}")


def op_dmatrices(op: Op) -> List[np.ndarray]:
    """d(matrix)/d(param_j) for each real parameter, in ``op.params`` order."""
    name, p = op.name, op.params
    if name in _ROT_AXIS:
        return [drot(p[0], PAULI[_ROT_AXIS[name]])]
    if name in _WORD_GATES:
        return [drot(p[0], pauli_tensor(_WORD_GATES[name]))]
    if name == "PauliRot":
        return [drot(p[0], pauli_tensor(op.extra))]
    if name in _CROT:
        return [dcontrolled(drot(p[0], PAULI[_CROT[name]]), 1)]
    if name == "ControlledPauliRot":
        word, nc = op.extra
        return [dcontrolled(drot(p[0], pauli_tensor(word)), nc)]
    if name in ("CPhase", "ControlledPhaseShift"):
        return [np.diag([0, 0, 0, 1j * np.exp(1j * p[0])]).astype(C)]
    if name == "Rot":
        phi, theta, omega = p
        a, b, c = rot(omega, PZ), rot(theta, PY), rot(phi, PZ)
        return [a @ b @ drot(phi, PZ), a @ drot(theta, PY) @ c, drot(omega, PZ) @ b @ c]
    return []


def op_structure(op: Op) -> Structure:
    name = op.name
    if name in _FIXED:
        return _FIXED[name][1]
    if name == "RZ":
        return Structure(diagonal=True)
    if name in _WORD_GATES:
        return Structure(diagonal=set(_WORD_GATES[name]) <= {"I", "Z"})
    if name == "PauliRot":
        return Structure(diagonal=set(op.extra) <= {"I", "Z"})
    if name in _CROT:
        return Structure(diagonal=name == "CRZ", controls=1)
    if name == "ControlledPauliRot":
        word, nc = op.extra
        return Structure(diagonal=set(word) <= {"I", "Z"}, controls=nc)
    if name in ("CPhase", "ControlledPhaseShift"):
        return Structure(diagonal=True, controls=1)
    if name == "DiagonalQubitUnitary":
        return Structure(diagonal=True)
    return Structure()


# ---------------------------------------------------------------------------
# application
# ---------------------------------------------------------------------------


def apply(psi: np.ndarray, m: np.ndarray, wires: Sequence[int]) -> np.ndarray:
    """Contract a k-qubit matrix into the *wires* axes of a rank-n tensor.

    Same contraction as operations.py:46-77 (`_einsum_subscript`) and the
    slice form of simulation.py:96-112.
    """
    k = len(wires)
    mt = m.reshape((2,) * (2 * k))
    out = np.tensordot(mt, psi, axes=(list(range(k, 2 * k)), list(wires)))
    return np.moveaxis(out, list(range(k)), list(wires))


def initial_state(n: int, psi0: Optional[np.ndarray] = None) -> np.ndarray:
    """|0...0> or the given state (simulation.py:384-388)."""
    if psi0 is None:
        s = np.zeros(2**n, dtype=C)
        s[0] = 1
        return s
    return np.asarray(psi0, dtype=C).reshape(2**n).copy()


def simulate(ops: Sequence[Op], n: int, psi0: Optional[np.ndarray] = None) -> np.ndarray:
    """Gate-by-gate statevector simulation (simulation.py:412-440)."""
    psi = initial_state(n, psi0).reshape((2,) * n)
    for op in ops:
        if op.name == "Barrier":
            continue
        psi = apply(psi, op_matrix(op), op.wires)
    return psi.reshape(-1)


def lifted(m: np.ndarray, wires: Sequence[int], n: int) -> np.ndarray:
    """Dense 2^n x 2^n embedding by applying to every basis vector
    (operations.py:534-551)."""
    dim = 2**n
    cols = [apply(np.eye(dim, dtype=C)[j].reshape((2,) * n), m, wires).reshape(-1) for j in range(dim)]
    return np.stack(cols, axis=1)


def lifted_kron(m: np.ndarray, wires: Sequence[int], n: int) -> np.ndarray:
    """Independent embedding: kron with identities, then a wire permutation.

    Used only to pin :func:`apply`/:func:`lifted` (operations.py:107-173).
    """
    k = len(wires)
    full = m
    missing = [w for w in range(n) if w not in wires]
    for _ in missing:
        full = np.kron(full, I2)
    order = list(wires) + missing  # current axis order
    perm = [order.index(w) for w in range(n)]
    t = full.reshape((2,) * (2 * n))
    t = np.transpose(t, perm + [p + n for p in perm])
    return t.reshape(2**n, 2**n)


# ---------------------------------------------------------------------------
# measurement
# ---------------------------------------------------------------------------


def z_signs(n: int) -> np.ndarray:
    """(n, 2^n) array of the PauliZ eigenvalue of wire q at each index."""
    idx = np.arange(2**n)
    return np.stack([1 - 2 * ((idx >> (n - 1 - q)) & 1) for q in range(n)]).astype(np.float64)


def expval_diag(psi: np.ndarray, n: int, diags: Sequence[Tuple[int, float, float]]) -> np.ndarray:
    """Single-qubit diagonal observables: sum |psi_i|^2 d[bit_q(i)]
    (simulation.py:772-801).  *diags* holds (wire, d0, d1)."""
    p = np.abs(psi.reshape(-1)) ** 2
    idx = np.arange(p.size)
    out = []
    for q, d0, d1 in diags:
        bit = (idx >> (n - q - 1)) & 1
        out.append(float(np.dot(p, np.where(bit, d1, d0))))
    return np.array(out)


def expval_op(psi: np.ndarray, n: int, m: np.ndarray, wires: Sequence[int]) -> float:
    """<psi|O|psi> for an observable on *wires* (simulation.py:806-809)."""
    t = psi.reshape((2,) * n)
    return float(np.real(np.vdot(t.reshape(-1), apply(t, m, wires).reshape(-1))))


# ---------------------------------------------------------------------------
# plan: same-wire fusion and the adjoint
# ---------------------------------------------------------------------------


class Block(NamedTuple):
    """A fused plan entry: the product of ``members`` (tape order)."""

    matrix: np.ndarray
    wires: Tuple[int, ...]
    members: Tuple[int, ...]  # indices into the op list
    structure: Structure


def fused_structure(left: Structure, right: Structure) -> Structure:
    """Guarantees kept by right @ left (simulation.py:202-211)."""
    perm = ()
    if left.permutation and right.permutation:
        perm = tuple(left.permutation[i] for i in right.permutation)
    return Structure(perm, left.diagonal and right.diagonal, min(left.controls, right.controls))


def fuse(ops: Sequence[Op]) -> List[Block]:
    """Merge runs on identical wire tuples (simulation.py:214-240).

    A gate multiplies into the previous block on the same wire tuple when no
    block in between touched any of those wires; the product is new @ old.
    """
    fused: List[list] = []
    last = {}
    for j, op in enumerate(ops):
        if op.name == "Barrier":
            continue
        m, w, s = op_matrix(op), tuple(op.wires), op_structure(op)
        i = last.get(w[0])
        if i is not None and fused[i][1] == w and all(last.get(x) == i for x in w):
            fused[i][0] = m @ fused[i][0]
            fused[i][2] = fused[i][2] + (j,)
            fused[i][3] = fused_structure(fused[i][3], s)
            continue
        fused.append([m, w, (j,), s])
        for x in w:
            last[x] = len(fused) - 1
    return [Block(m, w, mem, s) for m, w, mem, s in fused]


def observable_apply(psi: np.ndarray, n: int, obs: Sequence[Tuple[np.ndarray, Tuple[int, ...]]], ct: np.ndarray) -> np.ndarray:
    """sum_o ct_o O_o psi (flat)."""
    t = psi.reshape((2,) * n)
    acc = np.zeros_like(t)
    for (m, w), c in zip(obs, ct):
        acc = acc + c * apply(t, m, w)
    return acc.reshape(-1)


def adjoint(
    ops: Sequence[Op],
    n: int,
    obs: Sequence[Tuple[np.ndarray, Tuple[int, ...]]],
    ct: Optional[np.ndarray] = None,
    psi0: Optional[np.ndarray] = None,
):
    """Expectation values and their adjoint-method VJP (simulation.py:546-655).

    Forward: run the fused plan, measure.  Backward: lambda = VJP of the
    measurement for cotangent *ct* (JAX convention: lambda = 2 conj(sum ct O psi)),
    then for each plan entry in reverse: psi <- U^dagger psi, bar[o,i] =
    sum_rest lambda[o,rest] psi[i,rest] (no conjugation), lambda <- U^T lambda.
    Entry cotangents are chained to the real gate parameters with
    d f / d theta = Re sum bar * dU/dtheta.

    Returns (expvals, param_grads, bars, lambda0) where param_grads[j] lists
    the gradient of each real parameter of ``ops[j]`` and lambda0 is the
    initial-state cotangent in JAX convention.
    """
    plan = fuse(ops)
    if ct is None:
        ct = np.ones(len(obs))
    psi = initial_state(n, psi0).reshape((2,) * n)
    for b in plan:
        psi = apply(psi, b.matrix, b.wires)
    flat = psi.reshape(-1)
    vals = np.array([expval_op(flat, n, m, w) for m, w in obs])
    lam = (2 * np.conj(observable_apply(flat, n, obs, ct))).reshape((2,) * n)
    bars: List[np.ndarray] = [None] * len(plan)
    for bi in range(len(plan) - 1, -1, -1):
        b = plan[bi]
        k = len(b.wires)
        psi = apply(psi, np.conj(b.matrix).T, b.wires)
        rest = [a for a in range(n) if a not in b.wires]
        lt = np.moveaxis(lam, list(b.wires), list(range(k))).reshape(2**k, -1)
        pt = np.moveaxis(psi, list(b.wires), list(range(k))).reshape(2**k, -1)
        bars[bi] = lt @ pt.T
        lam = apply(lam, b.matrix.T, b.wires)
        del rest
    grads: List[List[float]] = [[] for _ in ops]
    for b, bar in zip(plan, bars):
        mats = [op_matrix(ops[j]) for j in b.members]
        for pos, j in enumerate(b.members):
            left = reduce(lambda a, c: c @ a, mats[pos + 1 :], np.eye(mats[0].shape[0], dtype=C))
            right = reduce(lambda a, c: c @ a, mats[:pos], np.eye(mats[0].shape[0], dtype=C))
            grads[j] = [float(np.real(np.sum(bar * (left @ dm @ right)))) for dm in op_dmatrices(ops[j])]
    return vals, grads, bars, lam.reshape(-1)


def flat_grads(ops: Sequence[Op], grads: List[List[float]]) -> np.ndarray:
    return np.array([g for gl in grads for g in gl])


def expval_fn(ops_fn, n: int, obs) -> callable:
    """theta -> expvals, for finite differences in the tests."""

    def f(theta):
        psi = simulate(ops_fn(theta), n)
        return np.array([expval_op(psi, n, m, w) for m, w in obs])

    return f


# ---------------------------------------------------------------------------
# the ansatz of the benchmark workloads
# ---------------------------------------------------------------------------


def hea_ops(theta: np.ndarray, n: int, features: Optional[np.ndarray] = None) -> List[Op]:
    """Hardware-efficient ansatz: optional RX(x_q) encoding layer, then per
    layer RX, RY, RZ on every wire and the CX ring q -> q+1 mod n.
    ``theta`` has shape (layers, n, 3)."""
    ops: List[Op] = []
    if features is not None:
        for q in range(n):
            ops.append(Op("RX", (q,), (float(features[q]),)))
    for layer in theta:
        for q in range(n):
            ops.append(Op("RX", (q,), (float(layer[q, 0]),)))
            ops.append(Op("RY", (q,), (float(layer[q, 1]),)))
            ops.append(Op("RZ", (q,), (float(layer[q, 2]),)))
        for q in range(n):
            ops.append(Op("CX", (q, (q + 1) % n)))
    return ops


def z_obs(n: int):
    return [(PZ, (q,)) for q in range(n)]
