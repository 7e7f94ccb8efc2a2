"""14-bit tiles (complex64): 2^14 amplitudes per 1024-thread forward /
512-thread backward workgroup, plan-specialised kernels only.  Expectation
values, angle gradients and initial-state cotangents against the complex128
oracle, and the ragged batch against the same batch on 13-bit tiles."""

import numpy as np
import pytest
import torch

from circuits import hea_theta, record_hea
from oracle import sv
from parity import assert_parity

pytestmark = pytest.mark.gpu


@pytest.fixture
def tile14(monkeypatch):
    """Every plan the Script path creates takes 14-bit tiles."""
    from jaqsi_amd import autotune, simulation

    monkeypatch.setattr(autotune, "config_for",
                        lambda *a, **k: {"low_bits": 5, "variant_flags": 0, "tile_bits": 14})
    yield simulation


def _plans_have_tile14(simulation):
    from jaqsi_amd import _native as nat

    plans = [p for p in simulation.cached_plans() if p.info()["tile_bits"] == 14]
    assert plans, "no 14-bit plan was created"
    for p in plans:
        assert p.jit() == nat.JQ_JIT_LOADED


@pytest.mark.parametrize("n,layers", [(14, 3), (14, 7), (15, 2)])
def test_tile14_expval_and_gradient(tile14, n, layers):
    import jaqsi_amd as J

    print(f"tile14 hea n={n} layers={layers}", flush=True)
    J.enable_x64(False)
    theta = hea_theta(n, layers, 2000 + 10 * n + layers)
    x = np.random.default_rng(n + layers).uniform(-np.pi, np.pi, n)
    th = torch.tensor(theta, dtype=torch.float64, device="cuda", requires_grad=True)
    obs = [J.PauliZ(wires=q, record=False) for q in range(n)]
    ev = J.Script(lambda t: record_hea(t, n, features=x), n_qubits=n).execute(type="expval", obs=obs, args=(th,))
    ev.sum().backward()
    _plans_have_tile14(tile14)
    ops = sv.hea_ops(theta, n, features=x)
    vals, grads, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
    assert_parity(ev.detach().cpu().numpy(), vals, False)
    assert_parity(th.grad.cpu().numpy().reshape(-1), sv.flat_grads(ops, grads)[n:], False)


def test_tile14_initial_state_and_its_cotangent(tile14):
    """A given initial state streams into LDS (1024-thread direct load) and
    gets its cotangent back from the 512-thread backward."""
    import jaqsi_amd as J

    n, layers = 14, 4
    J.enable_x64(False)
    rng = np.random.default_rng(77)
    theta = hea_theta(n, layers, 77)
    psi0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi0 /= np.linalg.norm(psi0)
    p0 = torch.tensor(psi0.astype(np.complex64), device="cuda", requires_grad=True)
    th = torch.tensor(theta, dtype=torch.float64, device="cuda", requires_grad=True)
    obs = [J.PauliZ(wires=q, record=False) for q in range(n)]
    ev = J.Script(lambda t: record_hea(t, n), n_qubits=n).execute(type="expval", obs=obs, args=(th,),
                                                                   initial_state=p0)
    ev.sum().backward()
    _plans_have_tile14(tile14)
    ops = sv.hea_ops(theta, n)
    vals, grads, _, lam0 = sv.adjoint(ops, n, sv.z_obs(n), psi0=psi0)
    assert_parity(ev.detach().cpu().numpy(), vals, False)
    assert_parity(th.grad.cpu().numpy().reshape(-1), sv.flat_grads(ops, grads), False)
    # torch's cotangent convention is the conjugate of JAX's
    assert_parity(p0.grad.detach().cpu().numpy().astype(np.complex128), np.conj(lam0), False)


def test_tile14_ragged_batch_matches_13_bit_tiles():
    """Groups of uneven depth on 14-bit tiles give the outputs of the same
    batch on 13-bit tiles, and one circuit per group matches the oracle."""
    from jaqsi_amd import _native as nat
    from jaqsi_amd.ragged import RaggedBatch, record

    n = 14
    depths = (2, 5, 9)
    batch = (5, 3, 4)
    rng = np.random.default_rng(1403)
    thetas = [rng.uniform(0, 2 * np.pi, (B, d, n, 3)) for d, B in zip(depths, batch)]
    feats = [rng.uniform(-np.pi, np.pi, (B, n)) for B in batch]
    tapes = [record(record_hea, np.zeros((d, n, 3)), n, np.zeros(n)) for d in depths]
    angles = [torch.tensor(np.concatenate([f, t.reshape(B, -1)], 1), device="cuda")
              for f, t, B in zip(feats, thetas, batch)]
    cts = [torch.ones((B, n), dtype=torch.float64, device="cuda") for B in batch]
    out = {}
    for tb in (13, 14):
        print(f"ragged tile_bits={tb}", flush=True)
        rb = RaggedBatch(tapes, list(batch), n, [(q, 1.0, -1.0) for q in range(n)], dtype=0,
                         knobs={"low_bits": 5, "variant_flags": 0, "tile_bits": tb})
        assert all(ph == nat.JQ_JIT_LOADED for ph in rb.specialise(nat.JQ_JIT_SYNC))
        assert all(p.info()["tile_bits"] == tb for p in rb.plans)
        evs, grads = rb.adjoint(angles, cts)
        out[tb] = ([e.cpu().numpy() for e in evs], [g.cpu().numpy() for g in grads])
    for g in range(len(depths)):
        assert_parity(out[14][0][g], out[13][0][g], False)
        assert_parity(out[14][1][g], out[13][1][g], False)
        ops = sv.hea_ops(thetas[g][0], n, features=feats[g][0])
        vals, gr, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
        assert_parity(out[14][0][g][0], vals, False)
        assert_parity(out[14][1][g][0], sv.flat_grads(ops, gr), False)
