"""Trailing permutation / diagonal gates measured through the observables
(``fold_tail`` of jq_plan_desc).

The affine index map the planner reports is checked against brute force
over every basis state (the oracle applies the folded gates to each basis
vector).  The folded program, run on the CPU by tests/plan_interp.py, with
the observables carried through that map, reproduces the oracle's
expectation values and adjoint gradients of the full circuit."""

import numpy as np
import pytest

from oracle import sv


def _lower(fn, n):
    from jaqsi_amd import lowering
    from jaqsi_amd.tape import recording

    with recording() as tape:
        fn()
    return lowering.lower_pure(tape, n)


def _plan(low, n, fold, dtype=1):
    from jaqsi_amd import _native as nat

    return nat.Plan(n, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, dtype, fold_tail=fold)


def _bits(x, n):
    """Bits of basis index x by wire (wire 0 is the most significant bit)."""
    return np.array([(x >> (n - 1 - w)) & 1 for w in range(n)])


def _basis_map(ops, n):
    """Final basis index of every basis state under `ops` (each must map
    basis states to basis states; phases are dropped)."""
    out = np.zeros(1 << n, dtype=np.int64)
    for x in range(1 << n):
        psi = np.zeros(1 << n, dtype=complex)
        psi[x] = 1
        y = np.flatnonzero(np.abs(sv.simulate(ops, n, psi0=psi)) > 0.5)
        assert len(y) == 1
        out[x] = y[0]
    return out


def _assert_map(plan, tail_ops, n):
    nf, rows, flip = plan.fold()
    assert nf == len(tail_ops)
    perm = _basis_map(tail_ops, n)
    for x in range(1 << n):
        xb, yb = _bits(x, n), _bits(int(perm[x]), n)
        for w in range(n):
            sel = [(int(rows[w]) >> v) & 1 for v in range(n)]
            assert (int(np.dot(sel, xb)) & 1) ^ ((flip >> w) & 1) == yb[w], (x, w)


def _word_signs(plan, n, words):
    """(2^n, len(words)) signs (+1 / -1) of Pauli-Z words on the final vector,
    expressed on the vector the folded plan computes."""
    _, rows, flip = plan.fold()
    xs = np.array([_bits(x, n) for x in range(1 << n)])
    out = np.zeros((1 << n, len(words)))
    for k, word in enumerate(words):
        m, f = 0, 0
        for w in word:
            m ^= int(rows[w])
            f ^= (flip >> w) & 1
        sel = np.array([(m >> v) & 1 for v in range(n)])
        out[:, k] = 1 - 2 * (((xs @ sel) & 1) ^ f)
    return out


def _hea_ring(theta, n):
    import jaqsi_amd as J

    for layer in range(theta.shape[0]):
        for q in range(n):
            J.RX(theta[layer, q, 0], q)
            J.RY(theta[layer, q, 1], q)
            J.RZ(theta[layer, q, 2], q)
        for q in range(n):
            J.CX([q, (q + 1) % n])


def test_cx_ring_folds_into_parity_rows():
    n = 7
    theta = np.random.default_rng(70).uniform(0, 2 * np.pi, (2, n, 3))
    low = _lower(lambda: _hea_ring(theta, n), n)
    plain, folded = _plan(low, n, 0), _plan(low, n, 1)
    assert plain.fold()[0] == 0
    _assert_map(folded, [sv.Op("CX", (q, (q + 1) % n)) for q in range(n)], n)
    # the last ring is gone from the passes, the rotations are all there
    assert folded.info()["n_blocks"] == plain.info()["n_blocks"]


def _mixed(theta, n):
    """Rotations, then a tail in which RZ on wire 5 stops the fold on wires 5
    and (through CX(2, 5)) 2; everything after it or on other wires folds."""
    import jaqsi_amd as J

    for q in range(n):
        J.RX(theta[q, 0], q)
        J.RY(theta[q, 1], q)
    J.PauliX(2)          # blocked (CX(2, 5) on wire 2 stays)
    J.SWAP([0, 4])       # folds
    J.CZ([1, 3])         # folds (diagonal)
    J.S(5)               # blocked
    J.CX([2, 5])         # stays: RZ(5) follows
    J.RZ(theta[0, 2], 5)  # a parameter: stays
    J.CX([4, 1])         # folds
    J.SWAP([3, 2])       # folds (after every kept gate on wires 2, 3)
    J.CX([3, 2])         # folds (fused with the SWAP)
    J.S(0)               # folds (diagonal)


MIXED_TAIL = [sv.Op("SWAP", (0, 4)), sv.Op("CZ", (1, 3)), sv.Op("CX", (4, 1)), sv.Op("SWAP", (3, 2)),
              sv.Op("CX", (3, 2)), sv.Op("S", (0,))]


def _mixed_ops(theta, n):
    ops = []
    for q in range(n):
        ops += [sv.Op("RX", (q,), (theta[q, 0],)), sv.Op("RY", (q,), (theta[q, 1],))]
    return ops + [sv.Op("PauliX", (2,)), sv.Op("SWAP", (0, 4)), sv.Op("CZ", (1, 3)), sv.Op("S", (5,)),
                  sv.Op("CX", (2, 5)), sv.Op("RZ", (5,), (theta[0, 2],)), sv.Op("CX", (4, 1)),
                  sv.Op("SWAP", (3, 2)), sv.Op("CX", (3, 2)), sv.Op("S", (0,))]


def test_mixed_tail_folds_only_what_commutes_to_the_end():
    n = 6
    theta = np.random.default_rng(61).uniform(0, 2 * np.pi, (n, 3))
    low = _lower(lambda: _mixed(theta, n), n)
    _assert_map(_plan(low, n, 1), MIXED_TAIL, n)


@pytest.mark.parametrize("case", ["ring", "mixed"])
def test_folded_program_matches_full_circuit(case):
    """Expectation values of single-wire projectors and Pauli-Z words, and
    the adjoint gradient of their weighted sum, from the folded program
    against the oracle's full circuit."""
    from plan_interp import Program

    rng = np.random.default_rng(5)
    if case == "ring":
        n = 6
        theta = rng.uniform(0, 2 * np.pi, (2, n, 3))
        low = _lower(lambda: _hea_ring(theta, n), n)
        ops = sv.hea_ops(theta, n)
    else:
        n = 6
        theta = rng.uniform(0, 2 * np.pi, (n, 3))
        low = _lower(lambda: _mixed(theta, n), n)
        ops = _mixed_ops(theta, n)
    from jaqsi_amd import lowering

    plan = _plan(low, n, 1)
    assert plan.fold()[0] > 0
    angles = np.array([float(a) for a in low.angle_values])
    xm = lowering.matrix_tensor(low.matrix_values, low.n_matrix_elems, None)
    xm = xm.numpy() if xm is not None else None
    # observables: Z on every wire, |1><1| on wire 2 (d = (0, 1)), Z0 Z3 Z5
    words = [(w,) for w in range(n)] + [(2,), (0, 3, 5)]
    d = np.array([[1.0, -1.0]] * n + [[0.0, 1.0], [1.0, -1.0]])
    ct = rng.normal(size=len(words))
    signs = _word_signs(plan, n, words)
    vals_of = lambda s: np.where(s > 0, d[:, 0], d[:, 1])  # d0 on even parity, d1 on odd
    weights = np.array([vals_of(s) for s in signs])  # (2^n, n_obs)
    prog = Program(plan)
    pre = prog.forward(angles, xm)
    ev = (np.abs(pre) ** 2) @ weights
    full = sv.simulate(ops, n)
    yb = np.array([_bits(y, n) for y in range(1 << n)])
    ref = np.zeros(len(words))
    for k, word in enumerate(words):
        par = yb[:, list(word)].sum(1) & 1
        ref[k] = (np.abs(full) ** 2) @ np.where(par == 0, d[k, 0], d[k, 1])
    np.testing.assert_allclose(ev, ref, atol=1e-12)
    # gradient of sum_k ct_k <O_k>: seed 2 w(x) conj(psi) on the pre-tail vector
    w = weights @ ct
    _, _, grads, _ = prog.adjoint(angles, lambda f: 2 * w * np.conj(f), xm)
    # the oracle's adjoint of the full circuit, each observable as its
    # diagonal matrix on the word's wires
    mats = []
    for k, word in enumerate(words):
        sub = np.array([d[k, 0] if bin(i).count("1") % 2 == 0 else d[k, 1] for i in range(1 << len(word))])
        mats.append((np.diag(sub).astype(complex), tuple(word)))
    _, og, _, _ = sv.adjoint(ops, n, mats, ct=ct)
    np.testing.assert_allclose(grads, sv.flat_grads(ops, og), atol=1e-10)


def test_nothing_folds_without_a_foldable_tail():
    import jaqsi_amd as J

    n = 4
    th = np.random.default_rng(4).uniform(0, 2 * np.pi, 4)
    # a parameterised gate last on every wire
    low = _lower(lambda: ([J.CX([q, (q + 1) % n]) for q in range(n)], [J.RY(th[q], q) for q in range(n)]), n)
    assert _plan(low, n, 1).fold()[0] == 0
    # only permutations: at least one op stays, so none folds
    low = _lower(lambda: [J.CX([q, (q + 1) % n]) for q in range(n)], n)
    assert _plan(low, n, 1).fold()[0] == 0
    # a Toffoli permutes basis states non-affinely
    low = _lower(lambda: ([J.RY(th[q], q) for q in range(n)], J.CCX([0, 1, 2])), n)
    assert _plan(low, n, 1).fold()[0] == 0


def test_final_ring_fold_saves_passes_at_24_qubits():
    n, layers = 24, 8
    theta = np.zeros((layers, n, 3))
    low = _lower(lambda: _hea_ring(theta, n), n)
    a, b = _plan(low, n, 0, dtype=0), _plan(low, n, 1, dtype=0)
    assert b.fold()[0] == n
    assert b.info()["n_passes"] < a.info()["n_passes"]
