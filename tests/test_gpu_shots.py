"""Shot-sampled Script outputs on the GPU: probabilities and expectation
values from shots agree with the exact GPU results within five standard
errors, batched (in_axes) and unbatched, for statevectors and density
matrices."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _circ(t):
    import jaqsi_amd as J

    J.Gates.RX(t, wires=0)
    J.Gates.RY(0.7, wires=1)
    J.Gates.CX(wires=[0, 1])


def _noisy(t):
    import jaqsi_amd as J

    _circ(t)
    J.Gates.RZ(0.3, wires=1, noise_params={"Depolarizing": 0.05})


@pytest.mark.parametrize("fn", [_circ, _noisy])
def test_shots_match_exact_results(fn):
    import jaqsi_amd as J

    print(f"shots {fn.__name__}", flush=True)
    J.enable_x64(True)
    try:
        script = J.Script(fn, n_qubits=2)
        obs = [J.PauliZ(wires=0, record=False), J.PauliZ(wires=1, record=False)]
        shots = 40_000
        ts = torch.tensor([0.3, 1.1, 2.5], dtype=torch.float64, device="cuda")
        exact_p = script.execute(type="probs", args=(ts,), in_axes=(0,)).cpu().numpy()
        exact_e = script.execute(type="expval", obs=obs, args=(ts,), in_axes=(0,)).cpu().numpy()
        got_p = script.execute(type="probs", args=(ts,), in_axes=(0,), shots=shots, key=5).cpu().numpy()
        got_e = script.execute(type="expval", obs=obs, args=(ts,), in_axes=(0,), shots=shots,
                               key=np.array([0, 9], dtype=np.uint32)).cpu().numpy()
        assert got_p.shape == exact_p.shape and got_e.shape == exact_e.shape
        sp = np.sqrt(exact_p * (1 - exact_p) / shots)
        assert np.all(np.abs(got_p - exact_p) <= 5 * sp + 1e-12)
        se = np.sqrt(np.clip(1 - exact_e**2, 0, None) / shots)
        assert np.all(np.abs(got_e - exact_e) <= 5 * se + 1e-12)
        one = script.execute(type="probs", args=(0.3,), shots=shots, key=5).cpu().numpy()
        np.testing.assert_array_equal(one, got_p[0])  # row 0 of the batch, same key
    finally:
        J.enable_x64(False)
