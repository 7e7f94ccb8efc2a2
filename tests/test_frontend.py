"""Host-side API behaviour (no GPU): tape recording, validation, lowering."""

import numpy as np
import pytest
import torch


def test_recording_and_wire_validation():
    import jaqsi_amd as J

    with J.recording() as tape:
        J.Gates.RX(0.1, wires=0)
        J.Gates.CX(wires=[0, 1])
        J.Gates.Barrier(wires=[0, 1])
    assert [op.name for op in tape] == ["RX", "CX", "Barrier"]
    with pytest.raises(ValueError):
        J.CX(wires=[0, 0])
    with pytest.raises(ValueError):
        J.RX(0.1, wires=[0, 1])
    with pytest.raises(TypeError):
        J.Gates.RX(0.1, wires=0, pulse="yes")


def test_pulse_gates_record_pending_evolutions_and_validate():
    """Gates.<G>(pulse=True) records pending evolutions of the pulse
    decomposition (gates.py:88-176, pulses.py:1370-1665); pulse_params are
    checked for type and length; the pulse manager scales the calibrated
    parameters; the batched propagator API checks its arguments."""
    import jaqsi_amd as J
    from jaqsi_amd import pulse
    from jaqsi_amd.evolution import PendingEvolution
    from jaqsi_amd.pulses import PulseInformation
    from jaqsi_amd.tape import recording

    with recording() as tape:
        J.Gates.RX(0.1, wires=0, pulse=True)
        J.Gates.CX(wires=[0, 1], pulse=True)
    assert all(isinstance(op, PendingEvolution) for op in tape)
    # RX; CX = H(target) CZ H(target), H = RZ(pi) RY(pi/2) + phase correction
    assert [op.name for op in tape] == ["RX", "RZ", "RY", "H", "CZ", "RZ", "RY", "H"]
    assert PulseInformation.H.size == 4 and PulseInformation.CX.size == 9
    assert [len(x) for x in PulseInformation.CX.split_params(np.arange(9.0))] == [4, 1, 4]
    with pytest.raises(ValueError, match="expects 3 pulse parameters"):
        J.Gates.RX(0.1, wires=0, pulse=True, pulse_params=[1.0, 2.0])
    with pytest.raises(TypeError):
        J.Gates.RX(0.1, wires=0, pulse=True, pulse_params="abc")
    with pytest.raises(TypeError, match="'pulse' must be a bool"):
        J.Gates.RX(0.1, wires=0, pulse=1)
    with recording() as tape:
        with J.Gates.pulse_manager_context(np.array([2.0, 1.0, 1.0, 0.5])):
            J.Gates.RX(0.3, wires=0, pulse=True)
            J.Gates.RZ(0.3, wires=0, pulse=True)
    base = np.asarray(PulseInformation.RX.params)
    np.testing.assert_allclose(np.asarray(tape[0]._params[0])[:3], base * [2.0, 1.0, 1.0])
    np.testing.assert_allclose(np.asarray(tape[1]._params[0])[0], 0.5 * 0.5)
    with PulseInformation.preserve_state():
        PulseInformation.set_envelope("drag", rwa=True)
        with recording() as tape:
            J.Gates.RX(0.3, wires=0, pulse=True)
        assert len(tape[0]._fns) == 2  # DRAG drives the second quadrature
    assert PulseInformation.get_envelope() == "gaussian"
    p, w = torch.zeros(3, 4, dtype=torch.float64), torch.zeros(3, 2, dtype=torch.float64)
    with pytest.raises(ValueError, match="CUDA"):
        pulse.propagate(p, w, np.eye(4))
    with pytest.raises(ValueError, match="omega"):
        pulse.propagate(p, w[:2], np.eye(4))
    with pytest.raises(ValueError, match="params"):
        pulse.propagate(p[:, :3], w, np.eye(4))
    with pytest.raises(ValueError, match="4x4"):
        pulse._target_host(np.eye(2))


def test_parametrized_hamiltonian_algebra():
    """coeff * Hermitian builds one term, + concatenates, - negates; terms
    must share wires and shapes (operations.py:722-863)."""
    import jaqsi_amd as J

    X = np.array([[0, 1], [1, 0]])
    Zm = np.array([[1, 0], [0, -1]])
    f = lambda p, t: p * t  # noqa: E731
    g = lambda p, t: p  # noqa: E731
    H = f * J.Hermitian(X, wires=0, record=False) + g * J.Hermitian(Zm, wires=0, record=False)
    assert H.n_terms == 2 and H.wires == [0]
    Hn = -H
    assert Hn.coeff_fns[0](2.0, 3.0) == -6.0
    assert (H - H).n_terms == 4
    with pytest.raises(ValueError, match="same wires"):
        f * J.Hermitian(X, wires=0, record=False) + g * J.Hermitian(X, wires=1, record=False)
    with pytest.raises(ValueError, match="same shape"):
        f * J.Hermitian(X, wires=0, record=False) + g * J.Hermitian(np.eye(4), wires=0, record=False)
    with pytest.raises(TypeError):
        3.0 * J.Hermitian(X, wires=0, record=False)
    with pytest.raises(ValueError, match="parameter set"):
        H.evolve()([1.0], 1.0)
    with pytest.raises(ValueError, match="Unknown solver"):
        H.evolve(solver="rk4")


def test_noise_params_attach_channels():
    import jaqsi_amd as J
    from jaqsi_amd.noise import BitFlip, DepolarizingChannel, QubitChannel

    with J.recording() as tape:
        J.Gates.RY(0.2, wires=1, noise_params={"BitFlip": 0.1, "Depolarizing": 0.01})
        J.Gates.CX(wires=[0, 1], noise_params={"MultiQubitDepolarizing": 0.02})
    kinds = [type(op) for op in tape]
    assert kinds[1] is BitFlip and kinds[2] is DepolarizingChannel and kinds[-1] is QubitChannel


def _rx_rho(theta):
    psi = np.array([np.cos(theta / 2), -1j * np.sin(theta / 2)])
    return np.outer(psi, psi.conj())


@pytest.mark.parametrize("pe,t1,t2,tg", [(0.0, 1e-4, 5e-5, 1e-6), (0.0, 1.0, 0.3, 0.8), (0.25, 1.0, 0.8, 0.6),
                                         (0.4, 1.0, 1.7, 0.9), (0.1, 2.0, 4.0, 3.0), (0.3, 1.0, 1.0, 0.0)])
def test_thermal_relaxation_matches_t1_t2_decay(pe, t1, t2, tg):
    """The reference's tests/test_jaqsi.py:689-711 (pe = 0, t2 < t1) and the
    general case: populations relax toward (1 - pe, pe) with exp(-tg/t1),
    coherences decay with exp(-tg/t2), in both regimes (t2 <= t1 and
    t1 < t2 <= 2 t1); the state after RX(1) is used as input."""
    from jaqsi_amd import noise

    rho = _rx_rho(1.0)
    ch = noise.ThermalRelaxationError(pe, t1, t2, tg, wires=0)
    ks = ch.kraus_matrices()
    assert len(ks) <= 4
    assert np.allclose(sum(k.conj().T @ k for k in ks), np.eye(2), atol=1e-14)
    out = sum(k @ rho @ k.conj().T for k in ks)
    e1, e2 = np.exp(-tg / t1), np.exp(-tg / t2)
    p1 = e1 * rho[1, 1] + (1 - e1) * pe
    ref = np.array([[1 - p1, e2 * rho[0, 1]], [e2 * rho[1, 0], p1]])
    np.testing.assert_allclose(out, ref, atol=1e-14)
    # the channel fixes the thermal state
    th = np.diag([1 - pe, pe]).astype(complex)
    np.testing.assert_allclose(sum(k @ th @ k.conj().T for k in ks), th, atol=1e-14)


def test_thermal_relaxation_rejects_non_physical_times():
    from jaqsi_amd import noise

    for args in [(0.1, 0.0, 1.0, 0.1), (0.1, 1.0, -1.0, 0.1), (0.1, 1.0, 2.5, 0.1), (0.1, 1.0, 1.0, -0.1),
                 (1.2, 1.0, 1.0, 0.1)]:
        with pytest.raises(ValueError):
            noise.ThermalRelaxationError(*args)


def test_kraus_from_superoperator_round_trip():
    """Random CPTP maps (random Stinespring isometries) on 1 and 2 qubits:
    the Kraus set read off the Choi matrix rebuilds the same superoperator
    and is trace preserving."""
    from jaqsi_amd.noise import kraus_from_superoperator

    rng = np.random.default_rng(7)
    for d, r in ((2, 3), (4, 5), (2, 1)):
        A = rng.normal(size=(d * r, d)) + 1j * rng.normal(size=(d * r, d))
        V, _ = np.linalg.qr(A)  # isometry C^d -> C^(d r)
        ks = [V[i * d:(i + 1) * d] for i in range(r)]
        S = sum(np.kron(k, k.conj()) for k in ks)
        ks2 = kraus_from_superoperator(S)
        assert len(ks2) == r
        np.testing.assert_allclose(sum(np.kron(k, k.conj()) for k in ks2), S, atol=1e-13)
        np.testing.assert_allclose(sum(k.conj().T @ k for k in ks2), np.eye(d), atol=1e-13)


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_diagonal_decompose_recomposes_the_diagonal(k):
    """The reference's tests/test_jaqsi.py:540-565: the product of the
    PauliRot factors equals diag(exp(-i x g)) times the dropped global phase
    exp(i x mean(g)) (the identity term), checked with the PauliRot matrices
    themselves; zero-weight strings are omitted."""
    import jaqsi_amd as J
    from jaqsi_amd.gateset import DiagonalQubitUnitary

    J.enable_x64(True)
    try:
        rng = np.random.default_rng(k)
        x = 0.83
        for g in (rng.normal(size=2**k), np.arange(2**k, dtype=float) ** 2):
            d = np.exp(-1j * g * x)
            op = DiagonalQubitUnitary(torch.tensor(d), wires=list(range(k)), generator=g, scale=x, record=False)
            U = np.eye(2**k, dtype=complex)
            for f in op.decompose():
                U = f.matrix.numpy() @ U
            np.testing.assert_allclose(U, np.diag(d) * np.exp(1j * x * g.mean()), atol=1e-12)
        # a generator that is one Z string has exactly one factor
        z = np.array([1.0, -1.0] * 2 ** (k - 1))
        facs = DiagonalQubitUnitary(torch.tensor(np.exp(-1j * z)), wires=list(range(k)), generator=z, scale=1.0,
                                    record=False).decompose()
        assert len(facs) == 1 and facs[0].pauli_word == "I" * (k - 1) + "Z"
    finally:
        J.enable_x64(False)


def test_kraus_channels_are_trace_preserving():
    from jaqsi_amd import noise

    for ch in (noise.BitFlip(0.1), noise.PhaseFlip(0.2), noise.DepolarizingChannel(0.3), noise.AmplitudeDamping(0.4),
               noise.PhaseDamping(0.5), noise.ThermalRelaxationError(0.1, 1.0, 0.8, 0.2),
               noise.ThermalRelaxationError(0.1, 1.0, 1.5, 0.2)):
        ks = ch.kraus_matrices()
        assert np.allclose(sum(k.conj().T @ k for k in ks), np.eye(2)), type(ch).__name__
    with pytest.raises(ValueError):
        noise.BitFlip(1.5)


def test_gate_matrices_match_oracle():
    import jaqsi_amd as J
    from oracle import sv

    t = 0.37
    pairs = [
        (J.RX(t, record=False), sv.Op("RX", (0,), (t,))),
        (J.CRY(t, wires=[0, 1], record=False), sv.Op("CRY", (0, 1), (t,))),
        (J.Rot(0.1, 0.2, 0.3, record=False), sv.Op("Rot", (0,), (0.1, 0.2, 0.3))),
        (J.PauliRot(t, "XZ", wires=[0, 1], record=False), sv.Op("PauliRot", (0, 1), (t,), "XZ")),
        (J.ControlledPhaseShift(t, record=False), sv.Op("CPhase", (0, 1), (t,))),
        (J.CCX(record=False), sv.Op("CCX", (0, 1, 2))),
    ]
    for op, ref in pairs:
        assert np.allclose(op.matrix.numpy(), sv.op_matrix(ref)), op.name


def test_lowering_structure_key_ignores_values():
    import jaqsi_amd as J
    from jaqsi_amd import lowering

    def tape(v):
        with J.recording() as t:
            J.RX(v, wires=0)
            J.CX(wires=[0, 1])
        return t

    a, b = lowering.lower_pure(tape(0.1), 2), lowering.lower_pure(tape(0.9), 2)
    assert a.key == b.key
    assert lowering.angle_tensor(a.angle_values, None).tolist() == [0.1]


def test_operation_algebra_on_host():
    import jaqsi_amd as J

    x = J.PauliX(0, record=False)
    z = J.PauliZ(1, record=False)
    prod = x @ z
    assert prod.wires == [0, 1]
    assert np.allclose(prod.matrix.numpy(), np.kron([[0, 1], [1, 0]], [[1, 0], [0, -1]]))
    d = J.RX(0.3, record=False).dagger()
    assert np.allclose((d.matrix @ J.RX(0.3, record=False).matrix).numpy(), np.eye(2), atol=1e-6)


def test_naive_gpu_comparator_matches_the_oracle():
    """The bench's plain-PyTorch comparator (gate by gate, autograd) computes
    the same objective and gradient as the oracle (run here on the CPU)."""
    import sys
    from pathlib import Path

    sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
    import bench
    from oracle import sv

    n, layers = 6, 2
    theta = np.random.default_rng(5).uniform(0, 2 * np.pi, (layers, n, 3))
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    f = bench.naive_hea_expval(th, n)
    f.backward()
    ops = sv.hea_ops(theta, n)
    vals, grads, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
    np.testing.assert_allclose(f.item(), vals.sum(), atol=1e-6)
    np.testing.assert_allclose(th.grad.numpy().reshape(-1), sv.flat_grads(ops, grads), atol=1e-6)


def test_naive_batched_comparator_matches_the_oracle():
    """The bench's batched plain-PyTorch comparator (one structure group, with
    the RX(x) encoding layer) against the oracle, circuit by circuit."""
    import sys
    from pathlib import Path

    sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
    import bench
    from oracle import sv

    n, layers, B = 5, 2, 3
    rng = np.random.default_rng(6)
    theta = rng.uniform(0, 2 * np.pi, (B, layers, n, 3))
    x = rng.uniform(-np.pi, np.pi, (B, n))
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    f = bench.naive_hea_batched(th, n, torch.tensor(x))
    f.sum().backward()
    for b in range(B):
        ops = sv.hea_ops(theta[b], n, features=x[b])
        vals, grads, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
        np.testing.assert_allclose(f[b].item(), vals.sum(), atol=1e-6)
        np.testing.assert_allclose(th.grad[b].numpy().reshape(-1), sv.flat_grads(ops, grads)[n:], atol=1e-6)


def test_cached_gate_lowering_matches_the_full_lowering():
    """Gates of static classes replay cached records: the plan key, op and
    member records, angle slots and matrices equal a full lowering (gate zoo
    plus a tensor-parameterised ansatz with explicit-matrix gates)."""
    import jaqsi_amd as J
    from circuits import record_zoo, zoo_params
    from jaqsi_amd import lowering
    from jaqsi_amd.tape import recording

    t = torch.randn(2, 4, 3, dtype=torch.float64, requires_grad=True)
    for fn, n in ((lambda: record_zoo(zoo_params()), 5),
                  (lambda: [J.Gates.RX(t[l, q, 0], wires=q) for l in range(2) for q in range(4)]
                   + [J.Operation(wires=[0, 2], matrix=np.eye(4, dtype=complex)), J.Gates.CRY(t[0, 1, 2], wires=[3, 1])], 4)):
        for _ in range(2):  # second pass hits the fragment cache
            with recording() as tape:
                fn()
            fast, slow = lowering.lower_pure(tape, n), lowering.lower_pure(tape, n, fast=False)
            assert fast.key == slow.key
            assert fast._ops_raw == slow._ops_raw and fast._members_raw == slow._members_raw
            assert all(a is b for a, b in zip(fast.angle_values, slow.angle_values))
            assert len(fast.angle_values) == len(slow.angle_values)
            assert [(o, id(v)) for o, v in fast.matrix_values] == [(o, id(v)) for o, v in slow.matrix_values]


def test_repeated_tape_reuses_records_with_new_parameters():
    """A tape of class-and-wires gates seen before is served from the tape
    cache: same key object (so plan lookups hash it once), same records, the
    new tape's parameters; a different structure misses it, and a gate built
    from an explicit matrix is lowered in full."""
    import jaqsi_amd as J
    from jaqsi_amd import lowering
    from jaqsi_amd.gateset import RX
    from jaqsi_amd.tape import recording

    n = 4

    def tape_of(t, swap=False, explicit=False):
        with recording() as tape:
            for q in range(n):
                J.Gates.RY(t[q], wires=q)
            for q in range(n):
                J.Gates.CX(wires=[q, (q + 1) % n] if not swap else [(q + 1) % n, q])
            if explicit:
                RX(0.0, wires=0, matrix=np.eye(2, dtype=complex))
            else:
                RX(0.0, wires=0)
        return tape

    t1, t2 = torch.randn(n, dtype=torch.float64), torch.randn(n, dtype=torch.float64)
    a = lowering.lower_pure(tape_of(t1), n)
    tape2 = tape_of(t2)
    b = lowering.lower_pure(tape2, n)
    assert b.key is a.key and b._ops_raw is a._ops_raw and b._members_raw is a._members_raw
    assert all(x is op.theta for x, op in zip(b.angle_values, tape2[:n] + tape2[-1:]))
    full = lowering.lower_pure(tape_of(t2), n, fast=False)
    assert b.key == full.key and hash(b.key) == hash(tuple(full.key))
    assert b._ops_raw == full._ops_raw and b._members_raw == full._members_raw
    assert [float(v) for v in b.angle_values] == [float(v) for v in full.angle_values]
    c = lowering.lower_pure(tape_of(t1, swap=True), n)
    assert c.key != a.key
    te = tape_of(t1, explicit=True)
    e, ef = lowering.lower_pure(te, n), lowering.lower_pure(te, n, fast=False)
    assert e.key == ef.key and e._ops_raw == ef._ops_raw and e._members_raw == ef._members_raw
