"""Quantum Fisher information formulas (jaqsi_amd.math, mirrors
jaqsi/math.py:209-448) pinned by closed forms on the CPU; the GPU Jacobian
that feeds them is tested in test_gpu_qfi.py."""

import numpy as np
import torch

from jaqsi_amd.math import _fubini_study, _qfi_density, marginalize_probs, partial_trace

X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.diag([1.0 + 0j, -1.0])


def _rho(r):
    return 0.5 * (np.eye(2) + r[0] * X + r[1] * Y + r[2] * Z)


def test_mixed_qubit_rotation_qfi_is_the_perpendicular_bloch_length():
    """rho(t) = exp(-i t Z/2) rho0 exp(i t Z/2): the SLD QFI is r_x^2 + r_y^2,
    whatever the purity."""
    for r in ([0.3, -0.4, 0.5], [0.1, 0.2, -0.9], [0.6, 0.0, 0.0]):
        rho = _rho(np.array(r))
        drho = -0.5j * (Z @ rho - rho @ Z)
        F = _qfi_density(torch.tensor(drho[:, :, None]), torch.tensor(rho))
        np.testing.assert_allclose(F.numpy(), [[r[0] ** 2 + r[1] ** 2]], atol=1e-14)


def test_purity_direction_qfi():
    """Shrinking the Bloch vector r(t) = t n (|n| = 1): F = 1 / (1 - t^2)."""
    n = np.array([0.6, 0.0, 0.8])
    for t in (0.2, 0.5, 0.9):
        F = _qfi_density(torch.tensor((0.5 * (n[0] * X + n[2] * Z))[:, :, None]), torch.tensor(_rho(t * n)))
        np.testing.assert_allclose(F.numpy(), [[1 / (1 - t * t)]], rtol=1e-13)


def test_pure_state_forms_agree():
    """For psi(a, b) = RY(b) RZ(a) RY(c)|0>, the Fubini-Study QFI equals the
    SLD QFI of |psi><psi| (zero eigenvalues masked), and the single-rotation
    entry is 1 - <P>^2 of the generator P before that rotation."""
    c, a, b = 0.7, 0.4, 1.1

    def psi(a, b):
        ry = lambda t: np.array([[np.cos(t / 2), -np.sin(t / 2)], [np.sin(t / 2), np.cos(t / 2)]])  # noqa: E731
        rz = lambda t: np.diag([np.exp(-0.5j * t), np.exp(0.5j * t)])  # noqa: E731
        return ry(b) @ rz(a) @ ry(c) @ np.array([1.0, 0.0])

    s = psi(a, b)
    h = 1e-6
    J = np.stack([(psi(a + h, b) - psi(a - h, b)) / (2 * h), (psi(a, b + h) - psi(a, b - h)) / (2 * h)], axis=1)
    Ffs = 4 * _fubini_study(torch.tensor(J), torch.tensor(s)).numpy()
    rho = np.outer(s, s.conj())
    drho = np.stack([np.outer(J[:, i], s.conj()) + np.outer(s, J[:, i].conj()) for i in range(2)], axis=-1)
    Fsld = _qfi_density(torch.tensor(drho), torch.tensor(rho)).numpy()
    np.testing.assert_allclose(Ffs, Fsld, atol=1e-8)
    # d/da: generator Z/2 on RY(c)|0>: 1 - <Z>^2 = sin^2 c
    np.testing.assert_allclose(Ffs[0, 0], np.sin(c) ** 2, atol=1e-8)


def test_partial_trace_and_marginals():
    rng = np.random.default_rng(0)
    A = rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8))
    rho = A @ A.conj().T
    rho /= np.trace(rho)
    t = rho.reshape((2,) * 6)
    ref = np.einsum("abcdbf->acdf", t).reshape(4, 4)  # trace wire 1
    np.testing.assert_allclose(partial_trace(torch.tensor(rho), 3, [0, 2]).numpy(), ref, atol=1e-15)
    batch = torch.tensor(np.stack([rho, rho.conj()]))
    assert partial_trace(batch, 3, [2, 0]).shape == (2, 4, 4)
    p = torch.tensor(np.real(np.diag(rho)))
    np.testing.assert_allclose(marginalize_probs(p, 3, [0, 2]).numpy()[0], np.real(np.diag(ref)), atol=1e-15)
    np.testing.assert_allclose(marginalize_probs(p, 3, [1]).numpy()[0],
                               np.real(np.diag(np.einsum("abcaec->be", t))), atol=1e-15)
