"""The native planner's programs, executed on the CPU by tests/plan_interp.py,
reproduce the oracle (statevector forward and adjoint gradients)."""

import numpy as np
import pytest

from circuits import record_zoo, zoo_ops, zoo_params
from oracle import sv


def _plan(tape, n, dtype=0):
    from jaqsi_amd import _native as nat, lowering

    low = lowering.lower_pure(tape, n)
    plan = nat.Plan(n, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, dtype)
    angles = np.array([float(a) for a in low.angle_values])
    xm = lowering.matrix_tensor(low.matrix_values, low.n_matrix_elems, None)
    return plan, angles, (xm.numpy() if xm is not None else None)


def _record(fn):
    from jaqsi_amd.tape import recording

    with recording() as tape:
        fn()
    return tape


@pytest.mark.parametrize("dtype", [0, 1])
@pytest.mark.parametrize("n,layers", [(3, 2), (9, 2), (14, 1), (16, 1)])
def test_hea_program_matches_oracle(n, layers, dtype):
    from circuits import record_hea
    from plan_interp import Program

    theta = np.random.default_rng(n).uniform(0, 2 * np.pi, (layers, n, 3))
    plan, angles, xm = _plan(_record(lambda: record_hea(theta, n)), n, dtype)
    prog = Program(plan)
    ops = sv.hea_ops(theta, n)
    np.testing.assert_allclose(prog.forward(angles, xm), sv.simulate(ops, n), atol=1e-12)
    zs = sv.z_signs(n).sum(0)
    _, _, grads, _ = prog.adjoint(angles, lambda f: 2 * zs * np.conj(f), xm)
    _, og, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
    np.testing.assert_allclose(grads, sv.flat_grads(ops, og), atol=1e-12)


@pytest.mark.parametrize("dtype", [0, 1])
def test_gate_zoo_program_matches_oracle(dtype):
    from plan_interp import Program

    p = zoo_params()
    plan, angles, xm = _plan(_record(lambda: record_zoo(p)), 5, dtype)
    prog = Program(plan)
    ops = zoo_ops(p)
    np.testing.assert_allclose(prog.forward(angles, xm), sv.simulate(ops, 5), atol=1e-12)
    zs = sv.z_signs(5).sum(0)
    _, _, grads, _ = prog.adjoint(angles, lambda f: 2 * zs * np.conj(f), xm)
    _, og, _, _ = sv.adjoint(ops, 5, sv.z_obs(5))
    np.testing.assert_allclose(grads, sv.flat_grads(ops, og), atol=1e-12)


def test_c2_schedule_shape():
    """The 24-qubit, 8-layer workload schedules into 21 passes (the CX ring
    advances about 8 wires per pass), and every tile holds global bits 0..4
    (256-byte contiguous runs of complex64)."""
    from circuits import record_hea

    theta = np.zeros((8, 24, 3))
    plan, _, _ = _plan(_record(lambda: record_hea(theta, 24)), 24)
    info = plan.info()
    assert info["n_passes"] == 21 and info["n_blocks"] == 192 and info["tile_bits"] == 13
    w = plan.program()
    T = int(w[2])
    for q in range(info["n_passes"]):
        rec = int(w[16 + q])
        tbits = [int(x) for x in w[rec + 4: rec + 4 + T]]
        assert tbits[:5] == [0, 1, 2, 3, 4] and tbits == sorted(tbits)


def test_planner_errors_map_to_exceptions():
    import jaqsi_amd as J
    from jaqsi_amd import _native as nat, lowering

    tape = _record(lambda: J.PauliRot(0.3, "XYZX", wires=[0, 1, 2, 3]))
    low = lowering.lower_pure(tape, 4)
    with pytest.raises(NotImplementedError):
        nat.Plan(4, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, 0)
    with pytest.raises(ValueError):
        nat.Plan(2, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, 0)


def test_deep_circuit_passes_stay_within_the_slot_budget():
    """A deep circuit on few wires (one tile holds every wire) is split into
    passes whose cotangent slots fit the LDS next to the two state tiles, and
    the split program still reproduces the oracle."""
    from circuits import record_hea
    from plan_interp import Program

    n, layers = 11, 48  # 528 fused 2x2 blocks x 8 reals > the ~3.8K reals of one pass
    theta = np.random.default_rng(77).uniform(0, 2 * np.pi, (layers, n, 3))
    plan, angles, xm = _plan(_record(lambda: record_hea(theta, n)), n, 0)
    prog = Program(plan)
    budget = (160 * 1024 - 2 * (8 << 13) - 1792) // 8 - 64
    nslots = [ns for _, _, ns, _, _, _ in prog.passes()]
    assert len(nslots) >= 2 and max(nslots) <= budget
    assert sum(nslots) == 2 * 4 * n * layers
    ops = sv.hea_ops(theta, n)
    np.testing.assert_allclose(prog.forward(angles, xm), sv.simulate(ops, n), atol=1e-11)
    zs = sv.z_signs(n).sum(0)
    _, _, grads, _ = prog.adjoint(angles, lambda f: 2 * zs * np.conj(f), xm)
    _, og, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
    np.testing.assert_allclose(grads, sv.flat_grads(ops, og), atol=1e-10)
