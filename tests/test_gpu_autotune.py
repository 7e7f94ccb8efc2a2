"""Every autotuner candidate (low_bits x kernel variants) computes the same
expectation values and adjoint gradients (checked against the complex128
oracle), on the plan-specialised kernels and the interpreter; and a tuning
run records a choice that plan creation then uses."""

import numpy as np
import pytest
import torch

from circuits import hea_theta, record_hea
from oracle import sv
from parity import assert_parity

pytestmark = pytest.mark.gpu

_CASES = [(5, 0), (4, 0), (6, 0), (3, 0), (5, 1), (5, 2), (5, 4), (4, 7)]


@pytest.mark.parametrize("n,jit", [(20, True), (14, False), (14, True)])
def test_candidates_agree_with_the_oracle(n, jit):
    from jaqsi_amd import _native as nat, lowering
    from jaqsi_amd.simulation import _Runner
    from jaqsi_amd.tape import recording

    theta = hea_theta(n, 2, 31)
    with recording() as tape:
        record_hea(theta, n)
    low = lowering.lower_pure(tape, n)
    ops = sv.hea_ops(theta, n)
    vals, grads, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
    want_g = sv.flat_grads(ops, grads)
    dev = torch.device("cuda", 0)
    a = torch.tensor([float(x) for x in low.angle_values], dtype=torch.float64, device=dev).reshape(1, -1)
    m = lowering.matrix_tensor(low.matrix_values, low.n_matrix_elems, None)
    m = m.reshape(1, -1).to(dev) if m is not None else None
    ct = torch.ones((1, n), dtype=torch.float64, device=dev)
    for lb, fl in _CASES:
        print(f"autotune candidate n={n} jit={jit} low_bits={lb} flags={fl}", flush=True)
        plan = nat.Plan(n, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, 0, 0, 1, lb, fl)
        plan.jit(nat.JQ_JIT_SYNC if jit else nat.JQ_JIT_OFF)
        r = _Runner(plan, 1, False, [(q, 1.0, -1.0) for q in range(n)], n, False, 0)
        ev, ga, _, _ = r.adjoint(a, m, None, ct, False)
        assert_parity(ev[0].cpu().numpy(), vals, False, err_msg=f"expvals {lb} {fl}")
        assert_parity(ga[0].cpu().numpy(), want_g, False, err_msg=f"gradients {lb} {fl}")


def test_tuning_records_a_choice_plan_creation_uses(tmp_path, monkeypatch):
    import jaqsi_amd as J
    from jaqsi_amd import autotune, simulation

    monkeypatch.setenv("JAQSI_AMD_TUNE_DB", str(tmp_path / "tune.json"))
    monkeypatch.setattr(autotune, "_db", None)
    n = 12
    theta = hea_theta(n, 2, 5)
    script = J.Script(lambda t: record_hea(t, n), n_qubits=n)
    from jaqsi_amd import lowering
    from jaqsi_amd.tape import recording

    with recording() as tape:
        record_hea(theta, n)
    low = lowering.lower_pure(tape, n)
    rec = autotune.tune(low, 0, 1, reps=2, rounds=2)
    assert rec["source"] == "tuned" and len(rec["tried"]) >= 4
    autotune.set_mode("lookup")
    try:
        obs = [J.PauliZ(wires=q, record=False) for q in range(n)]
        t = torch.tensor(theta, dtype=torch.float64, device="cuda", requires_grad=True)
        ev = script.execute(type="expval", obs=obs, args=(t,))
        ev.sum().backward()
        plans = [p for p in simulation.cached_plans() if p.n_wires == n]
        assert any(p.low_bits == rec["low_bits"] and p.variant_flags == rec["variant_flags"] for p in plans)
        ops = sv.hea_ops(theta, n)
        vals, grads, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
        assert_parity(ev.detach().cpu().numpy(), vals, False)
    finally:
        autotune._memo.clear()
        monkeypatch.setattr(autotune, "_db", None)
