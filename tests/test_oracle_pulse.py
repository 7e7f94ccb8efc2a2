"""The pulse oracle is pinned by closed forms, convergence order and an
independent per-block SU(2) restatement (CPU only)."""

import math

import numpy as np

from oracle import pulse, sv


def test_envelope_stable_form_matches_plain_form():
    t = np.linspace(0.0, 50.0, 101)
    for sigma in (5.0, 12.0, 20.0):
        g = np.exp(-((t - 25.0) ** 2) / (2 * sigma**2))
        g0 = math.exp(-(25.0**2) / (2 * sigma**2))
        np.testing.assert_allclose(pulse.lifted_gaussian(t, sigma, 25.0), (g - g0) / (1 - g0), atol=1e-13)
        E, Q = pulse.envelopes(t, 0.7, sigma, 0.3, 25.0)
        h = 1e-5
        Ep, _ = pulse.envelopes(t + h, 0.7, sigma, 0.3, 25.0)
        Em, _ = pulse.envelopes(t - h, 0.7, sigma, 0.3, 25.0)
        np.testing.assert_allclose(Q, -0.3 * (Ep - Em) / (2 * h), atol=1e-8)
    assert abs(pulse.lifted_gaussian(0.0, 10.0, 25.0)) < 1e-15
    assert abs(pulse.lifted_gaussian(25.0, 10.0, 25.0) - 1.0) < 1e-15


def test_zero_drive_closed_form():
    omega = (30.0, 31.5)
    U = pulse.propagate((0.0, 10.0, 0.5, 0.3), omega, pulse.J_C5, 50.0, 50)
    z = np.array([1, -1])
    e = np.array([omega[0] / 2 * z0 + omega[1] / 2 * z1 + pulse.J_C5 * z0 * z1 for z0 in z for z1 in z])
    np.testing.assert_allclose(U, np.diag(np.exp(-1j * 50.0 * e)), atol=1e-12)


def test_unitary_and_second_order_convergence():
    p, om = (0.5, 8.0, 0.4, 1.1), (29.0, 32.0)
    ref = pulse.propagate(p, om, pulse.J_C5, 50.0, 6400)
    errs = []
    for N in (800, 1600):
        U = pulse.propagate(p, om, pulse.J_C5, 50.0, N)
        np.testing.assert_allclose(U @ U.conj().T, np.eye(4), atol=1e-12)
        errs.append(np.abs(U - ref).max())
    assert 3.5 < errs[0] / errs[1] < 5.0  # Magnus-2 is second order


def test_block_su2_restatement():
    """H commutes with Z1, so each step is exp(-i h e_s) times an SU(2)
    rotation per value of qubit 1: an independent form of the same product."""
    p, om, N = (0.9, 15.0, -0.6, 4.0), (28.5, 33.0), 300
    U = pulse.propagate(p, om, pulse.J_C5, 50.0, N)
    h = 50.0 / N
    blocks = [np.eye(2, dtype=complex), np.eye(2, dtype=complex)]
    for n in range(N):
        t = (n + 0.5) * h
        E, Q = pulse.envelopes(t, p[0], p[1], p[2], 25.0)
        a = E * math.cos(p[3]) - Q * math.sin(p[3])
        b = E * math.sin(p[3]) + Q * math.cos(p[3])
        for s, z1 in enumerate((1, -1)):
            c = om[0] / 2 + pulse.J_C5 * z1
            r = math.sqrt(a * a + b * b + c * c)
            K = (a * sv.PX + b * sv.PY + c * sv.PZ) / r
            step = np.exp(-1j * h * om[1] / 2 * z1) * (math.cos(h * r) * np.eye(2) - 1j * math.sin(h * r) * K)
            blocks[s] = step @ blocks[s]
    full = np.zeros((4, 4), dtype=complex)
    for s in range(2):
        full[np.ix_([s, 2 + s], [s, 2 + s])] = blocks[s]
    np.testing.assert_allclose(U, full, atol=1e-12)


def test_infidelity_of_target_is_zero():
    V = pulse.target_rx_half_pi()
    assert abs(pulse.infidelity(V, V)) < 1e-15
    assert abs(pulse.infidelity(np.exp(0.3j) * V, V)) < 1e-15


def _gauss_legendre(T, n=200):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * T * (x + 1), 0.5 * T * w


def _qubit0_block(U, q1=0):
    """The 2x2 block of a qubit-0 operator at qubit 1 = q1 (wire 0 is the
    most significant index bit)."""
    idx = [q1, 2 + q1]
    return U[np.ix_(idx, idx)]


def test_drive_axis_against_the_pulse_area_closed_form():
    """beta = 0 and w0 = w1 = J = 0: H(t) = E(t) (cos phi X + sin phi Y) commutes
    with itself, so U = exp(-i S (cos phi X + sin phi Y)) (x) I with S the pulse
    area, by Gauss-Legendre (the reference's own check, tests/test_training.py:
    131-157).  The off-diagonal entry -i sin(S) e^{-i phi} fixes the sign and
    the axis of the in-phase drive."""
    T, N = 50.0, 4000
    t, w = _gauss_legendre(T)
    for A, sigma, phi in ((0.7, 10.0, 1.1), (0.3, 6.0, -2.4), (1.0, 18.0, 0.0)):
        E, _ = pulse.envelopes(t, A, sigma, 0.0, T / 2)
        S = float(np.sum(w * E))
        U = pulse.propagate((A, sigma, 0.0, phi), (0.0, 0.0), 0.0, T, N)
        ref = math.cos(S) * np.eye(2) - 1j * math.sin(S) * (math.cos(phi) * sv.PX + math.sin(phi) * sv.PY)
        for q1 in (0, 1):
            # the midpoint rule's area differs from S by O(h^2) (~1e-7 here)
            np.testing.assert_allclose(_qubit0_block(U, q1), ref, atol=1e-6)
        np.testing.assert_allclose(U[0, 1], 0.0, atol=1e-15)  # qubit 1 untouched
        assert abs(U[0, 2] - (-1j * math.sin(S) * np.exp(-1j * phi))) < 1e-6


def test_quadrature_axis_against_the_second_order_magnus_term():
    """phi = 0, w0 = w1 = J = 0, small A: H = E X + Q Y with Q = -beta dE/dt.
    The area of Q vanishes, and the second Magnus term is
    -1/2 int int_{t1>t2} [H(t1), H(t2)] = -i K Z with
    K = int int_{t1>t2} (E1 Q2 - Q1 E2) = -2 beta int E^2 dt (integrating
    Q = -beta E' by parts with E(0) = E(T) = 0).  So the Z component of U is
    -i K sin(r)/r, r = |(S, K)|, up to O(A^3): its sign fixes the axis of the
    quadrature (Y for phi = 0, i.e. phi + pi/2)."""
    T, N, A, sigma = 50.0, 4000, 1e-3, 10.0
    t, w = _gauss_legendre(T)
    for beta in (0.8, -0.5):
        E, _ = pulse.envelopes(t, A, sigma, beta, T / 2)
        S = float(np.sum(w * E))
        K = -2.0 * beta * float(np.sum(w * E * E))
        r = math.hypot(S, K)
        U = pulse.propagate((A, sigma, beta, 0.0), (0.0, 0.0), 0.0, T, N)
        u = _qubit0_block(U)
        zc = np.trace(sv.PZ @ u) / 2
        xc = np.trace(sv.PX @ u) / 2
        yc = np.trace(sv.PY @ u) / 2
        assert abs(zc - (-1j * K * math.sin(r) / r)) < 1e-3 * abs(K)
        assert abs(xc - (-1j * S * math.sin(r) / r)) < 1e-6 * abs(S)
        assert abs(yc) < 1e-3 * abs(K)
        # a quadrature on -Y instead would flip the sign of the Z component
        assert (zc.imag > 0) == (beta > 0)
