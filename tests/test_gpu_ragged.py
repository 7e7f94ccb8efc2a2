"""GPU parity of ragged batches (circuits of uneven depth in one
jq_sv_adjoint_groups call) against the complex128 oracle, per circuit."""

import numpy as np
import pytest
import torch

from circuits import record_hea
from oracle import sv
from parity import assert_parity

pytestmark = pytest.mark.gpu



@pytest.mark.parametrize("specialised", [False, True])
@pytest.mark.parametrize("x64", [False, True])
@pytest.mark.parametrize("n", [5, 14])
def test_uneven_depth_batch(n, x64, specialised):
    """Grouped interpreter launches, or (specialised) every group through its
    plan-specialised kernels."""
    import jaqsi_amd as J
    from jaqsi_amd import _native as nat
    from jaqsi_amd import simulation
    from jaqsi_amd.ragged import RaggedBatch, record

    print(f"ragged n={n} x64={x64} specialised={specialised}", flush=True)
    J.enable_x64(x64)
    simulation.specialise(None)  # fresh plans for this case
    rng = np.random.default_rng(3000 + n)
    depths = [1, 3, 2]
    per = [3, 2, 4]
    tapes = [record(record_hea, np.zeros((d, n, 3)), n, np.zeros(n)) for d in depths]
    rb = RaggedBatch(tapes, per, n, [(q, 1.0, -1.0) for q in range(n)])
    phases = rb.specialise(nat.JQ_JIT_SYNC if specialised else nat.JQ_JIT_OFF)
    if specialised:
        assert all(ph == nat.JQ_JIT_LOADED for ph in phases)
    thetas, feats, angles, cts = [], [], [], []
    for d, B in zip(depths, per):
        th = rng.uniform(0, 2 * np.pi, (B, d, n, 3))
        x = rng.uniform(-np.pi, np.pi, (B, n))
        thetas.append(th)
        feats.append(x)
        angles.append(torch.tensor(np.concatenate([x, th.reshape(B, -1)], axis=1), device="cuda"))
        cts.append(torch.tensor(rng.uniform(0.5, 1.5, (B, n)), device="cuda"))
    evs, grads = rb.adjoint(angles, cts)
    for g, (d, B) in enumerate(zip(depths, per)):
        for b in range(B):
            ops = sv.hea_ops(thetas[g][b], n, features=feats[g][b])
            ct = cts[g][b].cpu().numpy()
            vals, og, _, _ = sv.adjoint(ops, n, sv.z_obs(n), ct=ct)
            assert_parity(evs[g][b].cpu().numpy(), vals, x64)
            assert_parity(grads[g][b].cpu().numpy(), sv.flat_grads(ops, og), x64)


def test_groups_reject_shared_plan():
    import jaqsi_amd as J
    from jaqsi_amd.ragged import RaggedBatch, record

    J.enable_x64(False)
    n = 4
    t = record(record_hea, np.zeros((1, n, 3)), n, None)
    rb = RaggedBatch([t, t], [1, 1], n, [(0, 1.0, -1.0)])
    a = [torch.zeros((1, rb.n_params[0]), dtype=torch.float64, device="cuda") for _ in range(2)]
    c = [torch.ones((1, 1), dtype=torch.float64, device="cuda") for _ in range(2)]
    with pytest.raises(ValueError):
        rb.adjoint(a, c)
