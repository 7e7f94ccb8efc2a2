"""Plan-specialised pass kernels (run-time compiled, jit.cpp) against the
complex128 oracle and against the program interpreter on the same plans."""

import numpy as np
import pytest
import torch

from circuits import hea_theta, noisy_hea_reference, record_hea, zoo_params
from oracle import sv
from parity import assert_parity

pytestmark = pytest.mark.gpu



@pytest.fixture(params=["all-resident", "3-workgroups"])
def specialised(request, monkeypatch):
    """Plans compiled at creation.  With "3-workgroups" each persistent
    launch keeps at most three workgroups resident, so every workgroup walks
    several tiles (next-tile prefetch, slot and partial reuse)."""
    from jaqsi_amd import _native as nat
    from jaqsi_amd import simulation

    if request.param == "3-workgroups":
        monkeypatch.setenv("JQ_JIT_MAX_BLOCKS", "3")
    simulation.specialise(nat.JQ_JIT_SYNC)
    yield simulation
    simulation.specialise(None)


def _all_loaded(simulation):
    from jaqsi_amd import _native as nat

    plans = simulation.cached_plans()
    assert plans, "no plan was created"
    for p in plans:
        assert p.jit() == nat.JQ_JIT_LOADED, "plan ran without its specialised kernels"


def _hea_expval(n, theta, x64):
    import jaqsi_amd as J

    J.enable_x64(x64)
    th = torch.tensor(theta, dtype=torch.float64, device="cuda", requires_grad=True)
    obs = [J.PauliZ(wires=q, record=False) for q in range(n)]
    ev = J.Script(lambda t: record_hea(t, n), n_qubits=n).execute(type="expval", obs=obs, args=(th,))
    ev.sum().backward()
    return ev.detach().cpu().numpy(), th.grad.detach().cpu().numpy().reshape(-1)


@pytest.mark.parametrize("x64", [False, True])
@pytest.mark.parametrize("n,layers", [(5, 2), (14, 2), (16, 3)])
def test_hea_specialised_matches_oracle(specialised, n, layers, x64):
    print(f"specialised hea n={n} layers={layers} x64={x64}", flush=True)
    theta = hea_theta(n, layers, 1000 + n)
    ev, grad = _hea_expval(n, theta, x64)
    _all_loaded(specialised)
    ops = sv.hea_ops(theta, n)
    vals, grads, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
    assert_parity(ev, vals, x64)
    assert_parity(grad, sv.flat_grads(ops, grads), x64)


# The gate zoo spread over 16 wires: controls and diagonal factors land on
# register, lane and outside-tile bits depending on the pass.
WIRE_MAP = [0, 15, 7, 3, 12]


def _record_zoo_mapped(p, wm):
    import jaqsi_amd as J

    w = lambda *q: [wm[x] for x in q]  # noqa: E731
    J.H(wm[0])
    J.RX(p[0], wm[3])
    J.RY(p[1], wm[1])
    J.CX(w(0, 1))
    J.CRX(p[2], w(1, 2))
    J.Rot(p[3], p[4], p[5], wm[4])
    J.PauliRot(p[6], "XY", w(0, 2))
    J.ControlledPhaseShift(p[7], w(2, 1))
    J.CCX(w(3, 0, 2))
    J.CSWAP(w(2, 3, 0))
    J.SWAP(w(4, 1))
    J.CY(w(0, 3))
    J.CZ(w(4, 2))
    J.S(wm[3])
    J.RZZ(p[8], w(3, 0))
    J.CRZ(p[9], w(0, 4))
    J.ControlledPauliRot(p[10], "XY", w(3, 0, 2))
    J.ControlledPauliRot(p[11], "Y", w(2, 0, 3), n_controls=2)
    J.PauliX(wm[1])
    J.PauliY(wm[2])
    J.RZ(p[0] * 0.5, wm[4])
    for q in range(16):  # a layer touching every wire, so all passes carry gates
        J.RY(0.1 * (q + 1), q)


def _zoo_mapped_ops(p, wm):
    O = sv.Op
    m = lambda *q: tuple(wm[x] for x in q)  # noqa: E731
    ops = [
        O("H", m(0)), O("RX", m(3), (p[0],)), O("RY", m(1), (p[1],)), O("CX", m(0, 1)),
        O("CRX", m(1, 2), (p[2],)), O("Rot", m(4), (p[3], p[4], p[5])), O("PauliRot", m(0, 2), (p[6],), "XY"),
        O("CPhase", m(2, 1), (p[7],)), O("CCX", m(3, 0, 2)), O("CSWAP", m(2, 3, 0)), O("SWAP", m(4, 1)),
        O("CY", m(0, 3)), O("CZ", m(4, 2)), O("S", m(3)), O("RZZ", m(3, 0), (p[8],)), O("CRZ", m(0, 4), (p[9],)),
        O("ControlledPauliRot", m(3, 0, 2), (p[10],), ("XY", 1)),
        O("ControlledPauliRot", m(2, 0, 3), (p[11],), ("Y", 2)),
        O("PauliX", m(1)), O("PauliY", m(2)), O("RZ", m(4), (p[0] * 0.5,)),
    ]
    ops += [O("RY", (q,), (0.1 * (q + 1),)) for q in range(16)]
    return ops


def _zoo_run(p):
    import jaqsi_amd as J

    pt = torch.tensor(p, dtype=torch.float64, device="cuda", requires_grad=True)
    obs = [J.PauliZ(wires=q, record=False) for q in range(16)]
    ev = J.Script(lambda t: _record_zoo_mapped(t, WIRE_MAP), n_qubits=16).execute(type="expval", obs=obs,
                                                                                    args=(pt,))
    ev.sum().backward()
    return ev.detach().cpu().numpy(), pt.grad.cpu().numpy()


@pytest.mark.parametrize("x64", [False, True])
def test_zoo_on_16_wires_matches_oracle_and_interpreter(x64):
    import jaqsi_amd as J
    from jaqsi_amd import _native as nat
    from jaqsi_amd import simulation

    print(f"specialised zoo16 x64={x64}", flush=True)
    J.enable_x64(x64)
    p = zoo_params()
    try:
        simulation.specialise(nat.JQ_JIT_OFF)
        ev_i, g_i = _zoo_run(p)
        simulation.specialise(nat.JQ_JIT_SYNC)
        ev_j, g_j = _zoo_run(p)
        _all_loaded(simulation)
    finally:
        simulation.specialise(None)
    # same device functions in the same order: agreement far inside the oracle tolerance
    assert_parity(ev_j, ev_i, x64)
    assert_parity(g_j, g_i, x64)
    ops = _zoo_mapped_ops(p, WIRE_MAP)
    vals, grads, _, _ = sv.adjoint(ops, 16, sv.z_obs(16))
    assert_parity(ev_j, vals, x64)
    flat = sv.flat_grads(ops, grads)
    expect = flat[:12].copy()
    expect[0] += 0.5 * flat[12]  # the last RZ's angle is 0.5 * p[0]
    assert_parity(g_j, expect, x64)


def _record_noisy(theta, n, p_dep, gamma):
    import jaqsi_amd as J
    from jaqsi_amd import noise

    def after(wires):
        for w in wires:
            noise.DepolarizingChannel(p_dep, wires=w)
            noise.AmplitudeDamping(gamma, wires=w)

    for layer in range(theta.shape[0]):
        for q in range(n):
            J.Gates.RX(theta[layer, q, 0], wires=q)
            after([q])
            J.Gates.RY(theta[layer, q, 1], wires=q)
            after([q])
            J.Gates.RZ(theta[layer, q, 2], wires=q)
            after([q])
        for q in range(n):
            J.Gates.CX(wires=[q, (q + 1) % n])
            after([q, (q + 1) % n])


@pytest.mark.parametrize("x64", [False, True])
@pytest.mark.parametrize("n", [5, 7])  # 7 qubits = 14 wires: full tiles, HBM accessed under substep layouts
def test_noisy_density_matrix_specialised(specialised, n, x64):
    import jaqsi_amd as J

    layers = 1
    print(f"specialised dm n={n} x64={x64}", flush=True)
    J.enable_x64(x64)
    theta = hea_theta(n, layers, 2000 + n)
    th = torch.tensor(theta, dtype=torch.float64, device="cuda", requires_grad=True)
    script = J.Script(lambda t: _record_noisy(t, n, 0.01, 0.005), n_qubits=n)
    ev = script.execute(type="expval", obs=[J.PauliZ(wires=0, record=False)], args=(th,))
    ev.sum().backward()
    _all_loaded(specialised)
    v, g = noisy_hea_reference(n, layers, 2000 + n)
    assert_parity(ev.detach().cpu().numpy()[0], v, x64)
    assert_parity(th.grad.cpu().numpy().reshape(-1), g, x64)


def test_dense_three_wire_block_and_state_specialised(specialised):
    """A dense 3-wire unitary (LDS path) and a 4x4 block in the forward state."""
    import jaqsi_amd as J

    print("specialised dense3", flush=True)
    J.enable_x64(True)
    rng = np.random.default_rng(11)
    a = rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8))
    u8, _ = np.linalg.qr(a)
    b = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    u4, _ = np.linalg.qr(b)
    n = 14

    from jaqsi_amd.operations import Operation

    def circ():
        for q in range(n):
            J.H(q)
        Operation(wires=[2, 9, 13], matrix=torch.tensor(u8))
        Operation(wires=[12, 0], matrix=torch.tensor(u4))
        J.CRY(0.7, wires=[13, 1])

    st = J.Script(circ, n_qubits=n).execute(type="state").cpu().numpy()
    _all_loaded(specialised)
    ops = [sv.Op("H", (q,)) for q in range(n)]
    ops += [sv.Op("Matrix", (2, 9, 13), (), u8), sv.Op("Matrix", (12, 0), (), u4), sv.Op("CRY", (13, 1), (0.7,))]
    ref = sv.simulate(ops, n)
    assert_parity(st, ref, True)


def test_long_passes_stay_with_the_interpreter():
    """A deep circuit on few wires makes passes of hundreds of gates: the
    plan declines specialisation (no long compile) and runs interpreted."""
    import jaqsi_amd as J
    from jaqsi_amd import _native as nat
    from jaqsi_amd import lowering
    from jaqsi_amd.tape import recording

    print("specialised long pass", flush=True)
    J.enable_x64(False)
    theta = hea_theta(13, 40, 6)
    with recording() as tape:
        record_hea(theta, 13)
    low = lowering.lower_pure(tape, 13)
    plan = nat.Plan(13, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, 0)
    assert plan.jit(nat.JQ_JIT_SYNC) == nat.JQ_JIT_FAILED


def test_specialisation_modes_and_phases():
    import jaqsi_amd as J
    from jaqsi_amd import _native as nat
    from jaqsi_amd import lowering
    from jaqsi_amd.tape import recording

    J.enable_x64(False)
    theta = hea_theta(12, 2, 5)
    with recording() as tape:
        record_hea(theta, 12)
    low = lowering.lower_pure(tape, 12)
    plan = nat.Plan(12, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, 0)
    assert plan.jit() == nat.JQ_JIT_NONE  # below JQ_JIT_MIN_WIRES: interpreter by default
    ph = plan.jit(nat.JQ_JIT_ASYNC)
    assert ph in (nat.JQ_JIT_COMPILING, nat.JQ_JIT_COMPILED)
    assert plan.jit(nat.JQ_JIT_SYNC) == nat.JQ_JIT_LOADED
    assert plan.jit(nat.JQ_JIT_OFF) == nat.JQ_JIT_LOADED  # the module stays; launches use the interpreter
    with pytest.raises(ValueError):
        plan.jit(7)
