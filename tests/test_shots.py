"""Shot sampling (simulation.py:860-917): statistics of the counter-based
sampler, key handling and batch independence.  Samples differ from the
reference's jax.random.choice draws, so parity is statistical: every check
here is a goodness-of-fit bound at a fixed key."""

import numpy as np
import pytest
import torch
from scipy import stats

from jaqsi_amd.simulation import sample_shots, shot_uniforms


def test_uniforms_are_uniform_and_rows_independent():
    u = shot_uniforms(123, 4, 200_000)
    assert u.shape == (4, 200_000) and u.min() >= 0.0 and u.max() < 1.0
    for row in u:
        assert stats.kstest(row, "uniform").pvalue > 1e-4
    c = np.corrcoef(u)
    assert np.abs(c - np.eye(4)).max() < 5 / np.sqrt(u.shape[1])
    # consecutive draws within a row are uncorrelated
    assert abs(np.corrcoef(u[0, :-1], u[0, 1:])[0, 1]) < 5 / np.sqrt(u.shape[1])


def test_probs_estimate_passes_chi_square():
    rng = np.random.default_rng(3)
    p = rng.dirichlet(np.ones(16))
    shots = 200_000
    est = sample_shots(torch.tensor(p), 4, "probs", [], shots, key=11).numpy()
    assert abs(est.sum() - 1.0) < 1e-12
    counts = np.round(est * shots)
    assert stats.chisquare(counts, p * shots).pvalue > 1e-4


def test_expval_estimate_within_five_sigma():
    import jaqsi_amd as J

    p = torch.tensor([0.5, 0.1, 0.3, 0.1], dtype=torch.float64)
    obs = [J.PauliZ(wires=0, record=False), J.PauliZ(wires=1, record=False)]
    exact = np.array([0.6 - 0.4, 0.8 - 0.2])
    shots = 50_000
    got = sample_shots(p, 2, "expval", obs, shots, key=np.array([0, 42], dtype=np.uint32)).numpy()
    sigma = np.sqrt((1 - exact**2) / shots)
    assert np.all(np.abs(got - exact) < 5 * sigma)


def test_keys_and_batches():
    p = torch.full((5, 8), 1 / 8, dtype=torch.float64)
    a = sample_shots(p, 3, "probs", [], 1000, key=7)
    assert torch.equal(a, sample_shots(p, 3, "probs", [], 1000, key=7))
    assert not torch.equal(a, sample_shots(p, 3, "probs", [], 1000, key=8))
    # a JAX-style key array (two uint32 words), a torch tensor of the same
    # words and a list give the same draws
    k = np.array([0, 7], dtype=np.uint32)
    b = sample_shots(p, 3, "probs", [], 1000, key=k)
    assert torch.equal(b, sample_shots(p, 3, "probs", [], 1000, key=torch.tensor([0, 7])))
    assert torch.equal(b, sample_shots(p, 3, "probs", [], 1000, key=[0, 7]))
    # identical rows get independent draws, and row b depends only on (key, b)
    assert not torch.equal(a[0], a[1])
    assert torch.equal(a[:3], sample_shots(p[:3], 3, "probs", [], 1000, key=7))
    # an unbatched vector is row 0
    assert torch.equal(a[0], sample_shots(p[0], 3, "probs", [], 1000, key=7))


def test_zero_probability_outcomes_are_never_drawn():
    p = torch.tensor([0.0, 0.5, 0.0, 0.5], dtype=torch.float64)
    est = sample_shots(p, 2, "probs", [], 10_000, key=1)
    assert est[0] == 0 and est[2] == 0


def test_errors():
    p = torch.tensor([1.0, 0.0], dtype=torch.float64)
    with pytest.raises(ValueError, match="only supported"):
        sample_shots(p, 1, "state", [], 10, key=0)
    with pytest.raises(ValueError, match="positive"):
        sample_shots(p, 1, "probs", [], 0, key=0)
