"""The statevector oracle is pinned against closed forms, brute force and the
reference's golden training trajectories (CPU only)."""

import math

import numpy as np
import pytest

from oracle import sv
from oracle.adam import adam_trajectory

RNG = np.random.default_rng(0)


def test_gate_matrices_closed_forms_and_unitarity():
    t = 0.7
    assert np.allclose(sv.op_matrix(sv.Op("RX", (0,), (t,))), [[math.cos(t / 2), -1j * math.sin(t / 2)], [-1j * math.sin(t / 2), math.cos(t / 2)]])
    assert np.allclose(sv.op_matrix(sv.Op("RY", (0,), (t,))), [[math.cos(t / 2), -math.sin(t / 2)], [math.sin(t / 2), math.cos(t / 2)]])
    assert np.allclose(sv.op_matrix(sv.Op("RZ", (0,), (t,))), np.diag([np.exp(-0.5j * t), np.exp(0.5j * t)]))
    cx = sv.op_matrix(sv.Op("CX", (0, 1)))
    assert np.allclose(cx @ np.array([0, 0, 1, 0]), [0, 0, 0, 1])  # |10> -> |11>, control = first wire
    assert np.allclose(sv.op_matrix(sv.Op("CPhase", (0, 1), (t,))), np.diag([1, 1, 1, np.exp(1j * t)]))
    rot = sv.op_matrix(sv.Op("Rot", (0,), (0.1, 0.2, 0.3)))
    assert np.allclose(rot, sv.op_matrix(sv.Op("RZ", (0,), (0.3,))) @ sv.op_matrix(sv.Op("RY", (0,), (0.2,))) @ sv.op_matrix(sv.Op("RZ", (0,), (0.1,))))
    crx = sv.op_matrix(sv.Op("CRX", (0, 1), (t,)))
    assert np.allclose(crx[2:, 2:], sv.op_matrix(sv.Op("RX", (0,), (t,)))) and np.allclose(crx[:2, :2], np.eye(2))
    for name, w, p, e in [("H", (0,), (), None), ("SWAP", (0, 1), (), None), ("CCX", (0, 1, 2), (), None),
                          ("CSWAP", (0, 1, 2), (), None), ("PauliRot", (0, 1), (0.4,), "XY"),
                          ("ControlledPauliRot", (0, 1, 2), (0.4,), ("XY", 1)), ("RZZ", (0, 1), (0.3,), None)]:
        m = sv.op_matrix(sv.Op(name, w, p, e))
        assert np.allclose(m @ m.conj().T, np.eye(m.shape[0])), name


def test_structure_tags_match_matrices():
    for name, w, p, e in [("CX", (0, 1), (), None), ("SWAP", (0, 1), (), None), ("CCX", (0, 1, 2), (), None),
                          ("CSWAP", (0, 1, 2), (), None), ("PauliX", (0,), (), None)]:
        op = sv.Op(name, w, p, e)
        m, s = sv.op_matrix(op), sv.op_structure(op)
        perm = np.zeros_like(m)
        for r, c in enumerate(s.permutation):
            perm[r, c] = 1
        assert np.allclose(m, perm), name
    for name, w, p, e in [("RZ", (0,), (0.3,), None), ("CZ", (0, 1), (), None), ("CPhase", (0, 1), (0.3,), None),
                          ("RZZ", (0, 1), (0.2,), None), ("CRZ", (0, 1), (0.5,), None)]:
        op = sv.Op(name, w, p, e)
        m = sv.op_matrix(op)
        assert sv.op_structure(op).diagonal and np.allclose(m, np.diag(np.diag(m))), name


@pytest.mark.parametrize("wires", [(0,), (2,), (1, 3), (3, 0), (2, 0, 3)])
def test_apply_matches_kron_lift(wires):
    n = 4
    k = len(wires)
    m = RNG.normal(size=(2**k, 2**k)) + 1j * RNG.normal(size=(2**k, 2**k))
    psi = RNG.normal(size=2**n) + 1j * RNG.normal(size=2**n)
    out = sv.apply(psi.reshape((2,) * n), m, wires).reshape(-1)
    assert np.allclose(out, sv.lifted_kron(m, wires, n) @ psi)
    assert np.allclose(sv.lifted(m, wires, n), sv.lifted_kron(m, wires, n))


def test_closed_form_expvals():
    for t in (0.0, 0.3, 2.0):
        psi = sv.simulate([sv.Op("RX", (0,), (t,))], 1)
        assert np.isclose(sv.expval_diag(psi, 1, [(0, 1, -1)])[0], math.cos(t))
    bell = sv.simulate([sv.Op("H", (0,)), sv.Op("CX", (0, 1))], 2)
    assert np.allclose(bell, [1 / math.sqrt(2), 0, 0, 1 / math.sqrt(2)])
    ghz = sv.simulate([sv.Op("H", (0,)), sv.Op("CX", (0, 1)), sv.Op("CX", (1, 2))], 3)
    assert np.isclose(abs(ghz[0]) ** 2, 0.5) and np.isclose(abs(ghz[7]) ** 2, 0.5)
    # wire 0 is the most significant bit
    psi = sv.simulate([sv.Op("PauliX", (0,))], 3)
    assert np.isclose(abs(psi[4]), 1)


def test_fusion_keeps_the_state():
    theta = RNG.uniform(0, 6, (2, 4, 3))
    ops = sv.hea_ops(theta, 4)
    plan = sv.fuse(ops)
    assert len(plan) == 2 * (4 + 4)  # one fused block per wire and layer + the CX ring
    psi = sv.initial_state(4).reshape((2,) * 4)
    for b in plan:
        psi = sv.apply(psi, b.matrix, b.wires)
    assert np.allclose(psi.reshape(-1), sv.simulate(ops, 4))


def _fd_grad(f, x, h=1e-6):
    g = np.zeros_like(x)
    for i in range(x.size):
        e = np.zeros_like(x)
        e[i] = h
        g[i] = (f(x + e) - f(x - e)) / (2 * h)
    return g


def test_adjoint_against_finite_differences_and_parameter_shift():
    n = 3
    theta = RNG.uniform(0, 6, (2, n, 3))
    obs = sv.z_obs(n)

    def f(t):
        return sv.expval_fn(lambda tt: sv.hea_ops(tt.reshape(theta.shape), n), n, obs)(t).sum()

    ops = sv.hea_ops(theta, n)
    _, grads, _, _ = sv.adjoint(ops, n, obs)
    g = sv.flat_grads(ops, grads)
    assert np.allclose(g, _fd_grad(f, theta.reshape(-1)), atol=1e-7)
    # exact parameter-shift rule for Pauli rotations
    for i in (0, 5, 11):
        e = np.zeros(theta.size)
        e[i] = math.pi / 2
        ps = (f(theta.reshape(-1) + e) - f(theta.reshape(-1) - e)) / 2
        assert np.isclose(g[i], ps, atol=1e-12)


def test_adjoint_rich_circuit_all_gate_kinds():
    from circuits import zoo_ops, zoo_params

    p = zoo_params()
    obs = [(sv.PZ, (q,)) for q in range(5)] + [(sv.PX, (1,)), (np.kron(sv.PZ, sv.PZ), (0, 2))]

    def f(pp):
        return sv.expval_fn(lambda x: zoo_ops(x), 5, obs)(pp) @ np.arange(1, len(obs) + 1)

    ops = zoo_ops(p)
    _, grads, _, _ = sv.adjoint(ops, 5, obs, ct=np.arange(1, len(obs) + 1, dtype=float))
    flat = sv.flat_grads(ops, grads)
    g = flat[:12].copy()
    g[0] += 0.5 * flat[12]
    assert np.allclose(g, _fd_grad(f, p), atol=1e-7)


def test_golden_minimal_training_loop():
    """tests/test_training.py:33-66 of the reference (rtol 1e-12)."""

    def vg(p):
        ops = [sv.Op("RY", (0,), (p[0],)), sv.Op("RY", (1,), (p[1],)), sv.Op("CX", (0, 1))]
        vals, grads, _, _ = sv.adjoint(ops, 2, [(sv.PZ, (0,))])
        return vals[0], sv.flat_grads(ops, grads)

    losses, params = adam_trajectory(vg, [0.1, 0.2], 0.05, 100)
    np.testing.assert_allclose(
        losses[24:100:25],
        [0.18954755673957352, -0.9329512143583855, -0.9923251517460319, -0.9999997904383473],
        rtol=1e-12,
    )
    np.testing.assert_allclose(params[0], 3.1389723449742055, rtol=1e-12)
    np.testing.assert_allclose(params[1], 0.2, atol=1e-8)


def test_golden_batched_fit():
    """tests/test_training.py:69-106 of the reference.  The final w0 is the
    result of a cancellation; summation order shifts it by ~7e-15 absolute, so
    it is checked at atol 2e-14 instead of the reference's rtol 1e-12."""
    xs = np.linspace(0.0, np.pi, 16)
    ys = np.cos(xs)

    def vg(w):
        tot, g = 0.0, np.zeros(2)
        for x, y in zip(xs, ys):
            ops = [sv.Op("RX", (0,), (x,)), sv.Op("RY", (0,), (w[0],)), sv.Op("CX", (0, 1)), sv.Op("RY", (1,), (w[1],))]
            vals, _, _, _ = sv.adjoint(ops, 2, [(sv.PZ, (0,))])
            r = vals[0] - y
            _, grads, _, _ = sv.adjoint(ops, 2, [(sv.PZ, (0,))], ct=np.array([2 * r / 16]))
            tot += r * r / 16
            g += sv.flat_grads(ops, grads)[1:3]
        return tot, g

    losses, w = adam_trajectory(vg, [0.1, 0.2], 0.05, 100)
    np.testing.assert_allclose(
        losses[24:100:25],
        [2.007329667037132e-07, 1.3095863100844179e-08, 1.0464218803546601e-11, 7.771503138489971e-14],
        rtol=1e-12,
        atol=1e-15,
    )
    np.testing.assert_allclose(w[0], -0.000856006429212641, atol=2e-14)
    np.testing.assert_allclose(w[1], 0.2, atol=1e-8)


def test_cpu_baseline_matches_oracle():
    from oracle import cpu

    n = 6
    theta = RNG.uniform(0, 2 * np.pi, (3, n, 3))
    x = RNG.uniform(-np.pi, np.pi, n)
    ev, g = cpu.hea_grad(theta, x, threads=2)
    ops = sv.hea_ops(theta, n, features=x)
    vals, grads, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
    assert np.allclose(ev, vals, atol=2e-6)
    assert np.allclose(g, sv.flat_grads(ops, grads)[n:], atol=2e-5)


def test_cpu_baseline_c128_matches_oracle_to_rounding():
    """The complex128 build of the C++ restatement is the full-size reference
    of the 24-qubit headline workload; at 8 qubits it agrees with the NumPy
    oracle to rounding, and a planted mistake (one angle's gradient read from
    its neighbour) is caught at that bar."""
    from oracle import cpu

    n = 8
    theta = RNG.uniform(0, 2 * np.pi, (3, n, 3))
    ev, g = cpu.hea_grad(theta, None, threads=2, c128=True)
    ops = sv.hea_ops(theta, n)
    vals, grads, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
    ref = sv.flat_grads(ops, grads)
    np.testing.assert_allclose(ev, vals, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(g, ref, rtol=1e-12, atol=1e-13)
    assert not np.allclose(np.roll(g, 1), ref, rtol=1e-12, atol=1e-13)


def test_cpu_density_matrix_forward_matches_oracle():
    """The C++ density-matrix port (the reference's fused superoperator plan,
    complex64) against the complex128 density oracle."""
    from oracle import cpu, dm

    n = 3
    theta = RNG.uniform(0, 2 * np.pi, (2, n, 3))
    z = cpu.dm_forward(theta, 0.01, 0.005, threads=2)
    rho = dm.simulate(dm.noisy_hea(theta, n, 0.01, 0.005), n)
    np.testing.assert_allclose(z, dm.expval(rho, sv.PZ, (0,), n), atol=1e-6)


def test_cpu_pulse_gradient_matches_the_oracle():
    """The C++ Magnus-2 pulse port (expm per step, adjoint gradient) against
    the SciPy restatement: infidelities to rounding, gradients to central
    differences of it."""
    from oracle import cpu, pulse

    params, omega = pulse.workload(3, 1005)
    V = pulse.target_rx_half_pi()
    N = 120
    inf, g = cpu.pulse_grad(params, omega, N, 50.0, pulse.J_C5, V, threads=2)
    for s in range(3):
        U = pulse.propagate(params[s], omega[s], pulse.J_C5, 50.0, N)
        np.testing.assert_allclose(inf[s], pulse.infidelity(U, V), rtol=1e-12, atol=1e-14)
        for k in range(4):
            h = 1e-5 * max(1.0, abs(params[s, k]))
            pp, pm = params[s].copy(), params[s].copy()
            pp[k] += h
            pm[k] -= h
            fd = (pulse.infidelity(pulse.propagate(pp, omega[s], pulse.J_C5, 50.0, N), V)
                  - pulse.infidelity(pulse.propagate(pm, omega[s], pulse.J_C5, 50.0, N), V)) / (2 * h)
            np.testing.assert_allclose(g[s, k], fd, rtol=1e-5, atol=1e-9)
