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


@pytest.mark.parametrize("n,layers", [(3, 2), (9, 2)])
def test_rotation_blocks_reduced_cotangent_matches_oracle(n, layers):
    """The specialised backward keeps only t_k = Im Tr(bar^T U sigma_k)
    (3 reals) of blocks the planner flags as Pauli-rotation products, and the
    chain turns them into an effective cotangent of U; the angle gradients
    through the chain rule stay exact."""
    from circuits import record_hea
    from plan_interp import Program

    theta = np.random.default_rng(n).uniform(0, 2 * np.pi, (layers, n, 3))
    plan, angles, xm = _plan(_record(lambda: record_hea(theta, n)), n)
    prog = Program(plan, quat=True)
    zs = sv.z_signs(n).sum(0)
    _, bars, grads, _ = prog.adjoint(angles, lambda f: 2 * zs * np.conj(f), xm)
    ops = sv.hea_ops(theta, n)
    _, og, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
    np.testing.assert_allclose(grads, sv.flat_grads(ops, og), atol=1e-12)
    # the same with a generic complex cotangent seed, against the full bars
    seed = np.random.default_rng(n + 1).normal(size=(2, 1 << n)).T @ np.array([1, 1j])
    _, _, g_full, _ = Program(plan).adjoint(angles, lambda f: seed, xm)
    _, _, g_red, _ = prog.adjoint(angles, lambda f: seed, xm)
    np.testing.assert_allclose(g_red, g_full, atol=1e-12)
    # every HEA rotation block is flagged: 3 of its 8 bar reals are kept
    _, full, _, _ = Program(plan).adjoint(angles, lambda f: 2 * zs * np.conj(f), xm)
    assert np.count_nonzero(bars) * 8 == np.count_nonzero(full) * 3


def test_rotation_block_flags_on_the_gate_zoo():
    """The zoo holds H, phase shifts, explicit matrices and controlled
    rotations: with the reduced cotangent of the flagged blocks the
    gradients stay the oracle's."""
    from plan_interp import Program

    p = zoo_params()
    plan, angles, xm = _plan(_record(lambda: record_zoo(p)), 5)
    zs = sv.z_signs(5).sum(0)
    ops = zoo_ops(p)
    _, og, _, _ = sv.adjoint(ops, 5, sv.z_obs(5))
    _, _, grads, _ = Program(plan, quat=True).adjoint(angles, lambda f: 2 * zs * np.conj(f), xm)
    np.testing.assert_allclose(grads, sv.flat_grads(ops, og), atol=1e-12)


def test_rotations_fused_with_fixed_gates_are_not_flagged():
    """S . RY . H fuses into a 2x2 block outside the quaternion form: the
    planner leaves it unflagged (gradients exact), and flagging every 2x2
    block anyway (a planted mistake) breaks the gradients.  (The block acts
    after an entangling gate: on |0> the bar01 / bar11 entries vanish and the
    mistake would not show.)"""
    import jaqsi_amd as J
    from plan_interp import Program

    a, b, c = 0.7, -1.1, 0.4

    def circ():
        J.RY(a, wires=0)
        J.RX(b, wires=1)
        J.CX(wires=[0, 1])
        J.H(1)
        J.RY(c, wires=1)
        J.S(wires=1)
        J.CX(wires=[1, 0])

    ops = [sv.Op("RY", (0,), (a,)), sv.Op("RX", (1,), (b,)), sv.Op("CX", (0, 1)), sv.Op("H", (1,)),
           sv.Op("RY", (1,), (c,)), sv.Op("S", (1,)), sv.Op("CX", (1, 0))]
    plan, angles, xm = _plan(_record(circ), 2)
    zs = sv.z_signs(2).sum(0)
    _, og, _, _ = sv.adjoint(ops, 2, sv.z_obs(2))
    expect = sv.flat_grads(ops, og)
    _, _, grads, _ = Program(plan, quat=True).adjoint(angles, lambda f: 2 * zs * np.conj(f), xm)
    np.testing.assert_allclose(grads, expect, atol=1e-12)
    bad = Program(plan, quat=True)
    orig = bad._grad

    def flag_all(psi_in, lam_out, g, *rest):
        g = np.array(g)
        if int(g[0]) == 1:
            g[15] = 1
        return orig(psi_in, lam_out, g, *rest)

    bad._grad = flag_all
    # with a generic complex cotangent seed every bar entry matters
    seed = np.random.default_rng(7).normal(size=(4, 2)) @ np.array([1, 1j])
    _, _, full, _ = Program(plan).adjoint(angles, lambda f: seed, xm)
    _, _, reduced, _ = Program(plan, quat=True).adjoint(angles, lambda f: seed, xm)
    np.testing.assert_allclose(reduced, full, atol=1e-12)
    _, _, wrong, _ = bad.adjoint(angles, lambda f: seed, xm)
    assert np.abs(wrong - full).max() > 1e-2


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

    # dense blocks run on up to 5 wires (a pass of their own beyond the tile's 3)
    tape = _record(lambda: J.PauliRot(0.3, "XYZXYZ", wires=[0, 1, 2, 3, 4, 5]))
    low = lowering.lower_pure(tape, 6)
    with pytest.raises(NotImplementedError):
        nat.Plan(6, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, 0)
    with pytest.raises(ValueError):
        nat.Plan(2, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, 0)
    tape = _record(lambda: J.PauliRot(0.3, "XYZXYZX", wires=list(range(7))))
    with pytest.raises(NotImplementedError):
        lowering.lower_pure(tape, 7)


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


def _wide_circuit(p):
    """Gates the tile passes cannot hold with a gradient: a dense 3-wire
    rotation, a diagonal 3-wire rotation and a dense 4-wire rotation, between
    single-qubit layers (5 wires)."""
    import jaqsi_amd as J

    for q in range(5):
        J.RY(p[q], wires=q)
    J.PauliRot(p[5], "XYZ", wires=[0, 2, 4])
    J.PauliRot(p[6], "ZZZ", wires=[3, 1, 2])
    J.CX(wires=[1, 3])
    J.PauliRot(p[7], "XXYZ", wires=[4, 0, 1, 3])
    for q in range(5):
        J.RX(p[8 + q], wires=q)


def _wide_ops(p):
    O = sv.Op
    ops = [O("RY", (q,), (p[q],)) for q in range(5)]
    ops += [O("PauliRot", (0, 2, 4), (p[5],), "XYZ"), O("PauliRot", (3, 1, 2), (p[6],), "ZZZ"),
            O("CX", (1, 3)), O("PauliRot", (4, 0, 1, 3), (p[7],), "XXYZ")]
    ops += [O("RX", (q,), (p[8 + q],)) for q in range(5)]
    return ops


@pytest.mark.parametrize("dtype", [0, 1])
def test_wide_blocks_run_in_passes_of_their_own(dtype):
    """Dense blocks of 3+ target wires with a gradient, 4-wire dense blocks and
    3-wire diagonals with a gradient are scheduled as wide-block passes
    (program flag PF_WIDE); their adjoint reads the checkpointed input state.
    Forward and gradients match the oracle."""
    from plan_interp import PF_WIDE, Program

    p = np.random.default_rng(31).uniform(0, 2 * np.pi, 13)
    plan, angles, xm = _plan(_record(lambda: _wide_circuit(p)), 5, dtype)
    prog = Program(plan)
    flags = [f for _, f, _, _, _, _ in prog.passes()]
    assert sum(1 for f in flags if f & PF_WIDE) == 3
    ops = _wide_ops(p)
    np.testing.assert_allclose(prog.forward(angles, xm), sv.simulate(ops, 5), atol=1e-12)
    zs = sv.z_signs(5).sum(0)
    _, _, grads, _ = prog.adjoint(angles, lambda f: 2 * zs * np.conj(f), xm)
    _, og, _, _ = sv.adjoint(ops, 5, sv.z_obs(5))
    np.testing.assert_allclose(grads, sv.flat_grads(ops, og), atol=1e-12)


def _noisy_strong(t):
    """Channels whose superoperators are singular or badly conditioned
    between rotations on 2 qubits."""
    import jaqsi_amd as J
    from jaqsi_amd import noise

    J.RX(t[0], wires=0)
    noise.BitFlip(0.5, wires=0)
    J.RY(t[1], wires=0)
    J.RY(t[2], wires=1)
    noise.DepolarizingChannel(0.7, wires=1)
    J.CX(wires=[0, 1])
    noise.AmplitudeDamping(0.999, wires=0)
    J.RX(t[3], wires=0)
    noise.PhaseFlip(0.5, wires=1)
    J.RY(t[4], wires=1)
    noise.AmplitudeDamping(0.99, wires=1)
    J.RZ(t[5], wires=1)
    J.RX(t[6], wires=1)
    J.CX(wires=[1, 0])
    noise.PhaseDamping(0.3, wires=0)
    J.RY(t[7], wires=0)


def _noisy_strong_items(t):
    from oracle import dm

    O = sv.Op
    return [O("RX", (0,), (t[0],)), ("kraus", dm.kraus("BitFlip", 0.5), (0,)), O("RY", (0,), (t[1],)),
            O("RY", (1,), (t[2],)), ("kraus", dm.kraus("DepolarizingChannel", 0.7), (1,)), O("CX", (0, 1)),
            ("kraus", dm.kraus("AmplitudeDamping", 0.999), (0,)), O("RX", (0,), (t[3],)),
            ("kraus", dm.kraus("PhaseFlip", 0.5), (1,)), O("RY", (1,), (t[4],)),
            ("kraus", dm.kraus("AmplitudeDamping", 0.99), (1,)), O("RZ", (1,), (t[5],)), O("RX", (1,), (t[6],)),
            O("CX", (1, 0)), ("kraus", dm.kraus("PhaseDamping", 0.3), (0,)), O("RY", (0,), (t[7],))]


@pytest.mark.parametrize("dtype", [0, 1])
def test_singular_and_ill_conditioned_channels_are_never_inverted(dtype):
    """BitFlip(0.5) and PhaseFlip(0.5) have singular superoperators,
    DepolarizingChannel(0.7) and AmplitudeDamping(0.99 / 0.999) condition
    numbers of 15 to 1000: the blocks holding them run as wide-block passes
    (no inverse in the pool), PhaseDamping(0.3) (condition 1.2) stays in a
    tile pass and is un-applied by its inverse.  The adjoint gradient of
    <Z_0> + <Z_1> matches the exact parameter-shift rule of the density oracle."""
    from jaqsi_amd import _native as nat, lowering
    from oracle import dm
    from plan_interp import PF_WIDE, Program

    n = 2
    t = np.random.default_rng(5).uniform(0, 2 * np.pi, 8)
    low = lowering.lower_mixed(_record(lambda: _noisy_strong(t)), n)
    conds = sorted(op.cond for op in low.ops if op.structure & nat.JQ_S_NONUNITARY)
    assert conds[0] < 1.5 and np.isinf(conds[-1])
    plan = nat.Plan(2 * n, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, dtype)
    angles = np.array([float(a) for a in low.angle_values])
    xm = lowering.matrix_tensor(low.matrix_values, low.n_matrix_elems, None).numpy()
    prog = Program(plan)
    flags = [f for _, f, _, _, _, _ in prog.passes()]
    assert sum(1 for f in flags if f & PF_WIDE) >= 4
    bt = prog.bt
    wide_blocks = {int(g[1]) for _, f, _, _, _, subs in prog.passes() if f & PF_WIDE for g in subs[0][2]}
    assert all(bt[b][6] < 0 for b in wide_blocks)  # no inverse for a wide block
    assert any(bt[b][6] >= 0 for b in range(len(bt)) if b not in wide_blocks)  # PhaseDamping(0.3)
    rho = prog.forward(angles, xm).reshape(4, 4)
    items = _noisy_strong_items(t)
    np.testing.assert_allclose(rho, dm.simulate(items, n), atol=1e-12)
    dim = 1 << n
    zsum = sv.z_signs(n).sum(0)
    seed = np.zeros(dim * dim, dtype=complex)
    seed[np.arange(dim) * (dim + 1)] = zsum
    _, _, grads, _ = prog.adjoint(angles, lambda f: seed, xm)
    v, g_ref = dm.value_and_shift_grad(items, n, [(sv.PZ, (0,)), (sv.PZ, (1,))], np.array([1.0, 1.0]))
    assert np.all(np.isfinite(grads))
    np.testing.assert_allclose(grads, g_ref, atol=1e-12)


@pytest.mark.parametrize("low_bits,dtype", [(3, 0), (4, 0), (6, 0), (4, 1), (5, 1)])
def test_low_bits_knob_keeps_results(low_bits, dtype):
    """The autotuner's low_bits knob changes the tiles (the lowest low_bits
    global bits in every tile) and the pass count, not the results."""
    from circuits import record_hea
    from jaqsi_amd import _native as nat, lowering
    from plan_interp import Program

    n, layers = 16, 2
    theta = np.random.default_rng(9).uniform(0, 2 * np.pi, (layers, n, 3))
    low = lowering.lower_pure(_record(lambda: record_hea(theta, n)), n)
    plan = nat.Plan(n, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, dtype, low_bits=low_bits)
    w = plan.program()
    T = int(w[2])
    for q in range(plan.info()["n_passes"]):
        rec = int(w[16 + q])
        assert [int(x) for x in w[rec + 4: rec + 4 + low_bits]] == list(range(low_bits))
    prog = Program(plan)
    angles = np.array([float(a) for a in low.angle_values])
    ops = sv.hea_ops(theta, n)
    np.testing.assert_allclose(prog.forward(angles, None), sv.simulate(ops, n), atol=1e-12)
    zs = sv.z_signs(n).sum(0)
    _, _, grads, _ = prog.adjoint(angles, lambda f: 2 * zs * np.conj(f), None)
    _, og, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
    np.testing.assert_allclose(grads, sv.flat_grads(ops, og), atol=1e-12)


def test_tuning_knobs_are_validated():
    from jaqsi_amd import _native as nat, lowering

    low = lowering.lower_pure(_record(lambda: __import__("jaqsi_amd").Gates.RX(0.1, wires=0)), 2)
    for kw in ({"low_bits": 2}, {"low_bits": 7}, {"variant_flags": 16}):
        with pytest.raises(ValueError):
            nat.Plan(2, low.ops, low.members, 1, low.n_matrix_elems, 0, **kw)


@pytest.mark.parametrize("layers", [3, 7])
def test_fourteen_bit_tile_programs_match_oracle(layers):
    """14-bit tiles (complex64): a 14-qubit circuit is one tile held by a
    1024-thread forward / 512-thread backward workgroup; its programs (passes
    capped at a block budget, 32-amplitude backward layouts) reproduce the
    oracle's state and adjoint gradients, in fewer passes than 13-bit tiles."""
    from circuits import record_hea
    from jaqsi_amd import _native as nat, lowering
    from plan_interp import Program

    n = 14
    rng = np.random.default_rng(40 + layers)
    theta = rng.uniform(0, 2 * np.pi, (layers, n, 3))
    x = rng.uniform(-np.pi, np.pi, n)
    low = lowering.lower_pure(_record(lambda: record_hea(theta, n, features=x)), n)
    angles = np.array([float(a) for a in low.angle_values])
    ops = sv.hea_ops(theta, n, features=x)
    zs = sv.z_signs(n).sum(0)
    _, og, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
    passes = {}
    for tb in (13, 14):
        plan = nat.Plan(n, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, 0, tb)
        prog = Program(plan, quat=True)
        assert prog.T == tb
        np.testing.assert_allclose(prog.forward(angles, None), sv.simulate(ops, n), atol=1e-12)
        _, _, grads, _ = prog.adjoint(angles, lambda f: 2 * zs * np.conj(f), None)
        np.testing.assert_allclose(grads, sv.flat_grads(ops, og), atol=1e-10)
        passes[tb] = plan.info()["n_passes"]
    assert passes[14] < passes[13]


def test_fourteen_bit_tiles_refuse_what_they_cannot_run():
    """14-bit tiles need at least 14 wires and no wide-pass or dense 3-wire
    blocks (those keep 13-bit tiles); complex128 keeps 12-bit tiles."""
    import jaqsi_amd as J
    from jaqsi_amd import _native as nat, lowering

    def plan(fn, n, tb, dtype=0):
        low = lowering.lower_pure(_record(fn), n)
        return nat.Plan(n, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, dtype, tb)

    with pytest.raises(ValueError, match="at least 14 wires"):
        plan(lambda: J.Gates.RX(0.1, wires=0), 12, 14)
    with pytest.raises(ValueError, match="tile_bits"):
        plan(lambda: J.Gates.RX(0.1, wires=0), 14, 14, dtype=1)
    u8 = np.linalg.qr(np.random.default_rng(1).normal(size=(8, 8)) + 1j * np.random.default_rng(2).normal(size=(8, 8)))[0]
    with pytest.raises(NotImplementedError, match="13-bit tiles"):
        plan(lambda: J.Operation(wires=[0, 5, 9], matrix=u8), 14, 14)
    assert plan(lambda: J.Gates.RX(0.1, wires=13), 14, 14).info()["tile_bits"] == 14


@pytest.mark.parametrize("bwd_rb", [4, 5])
def test_specialised_backward_program_matches_oracle(bwd_rb):
    """The backward program the plan-specialised kernels run (16 amplitudes x
    512 threads, or 32 x 256 threads sharing the CU two workgroups at a time)
    gives the oracle's adjoint gradients on the CPU interpreter."""
    from circuits import record_hea
    from jaqsi_amd import _native as nat, lowering
    from plan_interp import Program

    n, layers = 16, 3
    theta = np.random.default_rng(16).uniform(0, 2 * np.pi, (layers, n, 3))
    low = lowering.lower_pure(_record(lambda: record_hea(theta, n)), n)
    flags = nat.JQ_TUNE_BWD16 if bwd_rb == 4 else 0
    plan = nat.Plan(n, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, 0, variant_flags=flags)
    assert int(plan.table(5)[3]) == bwd_rb
    angles = np.array([float(a) for a in low.angle_values])
    ops = sv.hea_ops(theta, n)
    zs = sv.z_signs(n).sum(0)
    _, og, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
    _, _, grads, _ = Program(plan, quat=True, backward_table=5).adjoint(angles, lambda f: 2 * zs * np.conj(f), None)
    np.testing.assert_allclose(grads, sv.flat_grads(ops, og), atol=1e-10)
