"""The pulse-evolution oracle is pinned by closed forms, convergence orders,
frame identities and finite differences (CPU only)."""

import math

import numpy as np
from scipy.linalg import expm
from scipy.special import erf

from oracle import evolve as ev

RNG = np.random.default_rng(2024)


def test_magnus2_of_a_self_commuting_drive_is_the_midpoint_area_rotation():
    p = [0.7, 2.0, 6.0, 1.3]
    f = ev.coeff_fns("gaussian")[0]
    N = 37
    U = ev.magnus2([f], [p], [ev.X], 0.0, 6.0, N)
    h = 6.0 / N
    S = sum(h * f(p, (n + 0.5) * h) for n in range(N))
    np.testing.assert_allclose(U, math.cos(S) * ev.I2 - 1j * math.sin(S) * ev.X, atol=1e-14)
    # and the midpoint area converges to the integral at second order
    e1 = abs(S - ev.area(f, p, 0.0, 6.0))
    h2 = 6.0 / (2 * N)
    S2 = sum(h2 * f(p, (n + 0.5) * h2) for n in range(2 * N))
    assert 3.5 < e1 / abs(S2 - ev.area(f, p, 0.0, 6.0)) < 4.5


def _two_term():
    p = [0.9, 0.6, 1.8, 4.0, 1.7]
    fx, fy = ev.coeff_fns("drag")[:2]
    return [fx, fy], [p, p], [ev.X, ev.Y], 4.0


def test_fixed_step_orders():
    fns, params, Hs, T = _two_term()
    ref = ev.cfm4(fns, params, Hs, 0.0, T, 1024)
    e2 = [np.abs(ev.magnus2(fns, params, Hs, 0.0, T, N) - ref).max() for N in (64, 128)]
    e4 = [np.abs(ev.cfm4(fns, params, Hs, 0.0, T, N) - ref).max() for N in (16, 32)]
    assert 3.5 < e2[0] / e2[1] < 4.5  # Magnus-2: second order
    assert 13.0 < e4[0] / e4[1] < 19.0  # CFM4: fourth order


def test_closed_form_matches_the_erf_area_of_the_gaussian():
    """The reference's own check (tests/test_pulse_state.py:238-259): the RWA
    gaussian RY with the lifted-Gaussian area in closed form."""
    A, sigma, T = 0.449245182920846, 1.8896592598281816, 3.4886409030956145
    w = 60.0
    U = ev.ry_rwa(w, [A, sigma, T])
    g0 = math.exp(-(T**2) / (8 * sigma**2))
    a = sigma * math.sqrt(2 * math.pi) * erf(T / (2 * math.sqrt(2) * sigma)) - T * g0
    F = w * A * a / (2 * (1 - g0))
    np.testing.assert_allclose(U, math.cos(F) * ev.I2 - 1j * math.sin(F) * ev.Y, atol=1e-12)


def test_rwa_is_the_slow_half_of_the_lab_frame():
    """On resonance the lab-frame RX_X coefficient E cos^2(w t) w equals the
    RWA's (E / 2) w plus the fast (E / 2) cos(2 w t) w."""
    p = [0.5, 1.5, 3.0, 0.8]
    t = RNG.uniform(0, 3.0, 50)
    lab = ev.coeff_fns("gaussian", rwa=False, frame="lab")
    rwa = ev.coeff_fns("gaussian", rwa=True)
    E = ev.envelope("gaussian", p, t, 1.5)
    np.testing.assert_allclose(lab[0](p, t) - rwa[0](p, t), 0.5 * E * np.cos(2 * ev.OMEGA * t) * p[-1], atol=1e-13)


def test_drive_frame_equals_lab_frame():
    p = [0.6, 0.3, 1.2, 3.0, 1.1]
    t = RNG.uniform(0, 3.0, 40)
    for wq in (ev.OMEGA, 0.93 * ev.OMEGA):
        lab = ev.coeff_fns("drag", rwa=False, frame="lab", wq=wq)
        drv = ev.coeff_fns("drag", rwa=False, frame="drive", wq=wq)
        for a, b in zip(lab, drv):
            np.testing.assert_allclose(a(p, t), b(p, t), atol=1e-13)


def test_envelopes_against_plain_forms():
    t = np.linspace(0.0, 4.0, 81)
    p = [0.8, 1.3, 4.0]
    g = np.exp(-((t - 2.0) ** 2) / (2 * 1.3**2))
    g0 = math.exp(-(2.0**2) / (2 * 1.3**2))
    np.testing.assert_allclose(ev.envelope("gaussian", p, t, 2.0), 0.8 * (g - g0) / (1 - g0), atol=1e-14)
    q = [0.8, 0.4, 1.3, 4.0]
    h = 1e-6
    dE = (ev.envelope("drag", q, t + h, 2.0) - ev.envelope("drag", q, t - h, 2.0)) / (2 * h)
    np.testing.assert_allclose(ev.quadrature("drag", q, t, 2.0), -0.4 * dE, atol=1e-8)
    np.testing.assert_allclose(ev.envelope("sech", p, t, 2.0), 0.8 / np.cosh((t - 2.0) / 1.3), atol=1e-15)


def test_product_gradient_against_central_differences():
    for d, Hs in ((2, [ev.X, ev.Z]), (4, [np.kron(ev.X, ev.I2), np.kron(ev.Z, ev.Z), np.kron(ev.I2, ev.Y)])):
        coef = RNG.normal(size=(5, len(Hs)))
        G = RNG.normal(size=(d, d)) + 1j * RNG.normal(size=(d, d))
        g = ev.product_grad(coef, Hs, G)
        L = lambda c: float(np.real(np.trace(G.conj().T @ ev.product(c, Hs))))  # noqa: E731
        fd = np.zeros_like(coef)
        for idx in np.ndindex(coef.shape):
            cp, cm = coef.copy(), coef.copy()
            cp[idx] += 1e-6
            cm[idx] -= 1e-6
            fd[idx] = (L(cp) - L(cm)) / 2e-6
        np.testing.assert_allclose(g, fd, atol=1e-8)


def test_pulse_cz_is_cz():
    np.testing.assert_allclose(ev.cz_pulse(), np.diag([1, 1, 1, -1]), atol=1e-8)
    np.testing.assert_allclose(ev.rz_pulse(0.7), np.diag(np.exp([-0.35j, 0.35j])), atol=1e-15)
