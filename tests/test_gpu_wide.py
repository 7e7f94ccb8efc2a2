"""GPU parity of the wide-block passes (dense blocks of 3+ target wires with
gradients, 4-5-wire dense blocks, wide diagonals, singular or badly
conditioned channels), of the gradients of probs / state / density outputs
and general observables, and of batched noisy circuits, against the
complex128 oracle."""

import math

import numpy as np
import pytest
import torch

from oracle import dm, sv
from parity import assert_parity
from test_planner import _noisy_strong, _noisy_strong_items, _wide_circuit, _wide_ops

pytestmark = pytest.mark.gpu


def _shift_grad(f, t):
    """Exact gradient of a function that is a degree-1 trigonometric
    polynomial in every angle (any linear functional of |psi><psi| or of rho
    under exp(-i t P / 2) gates): the parameter-shift rule."""
    g = np.zeros(len(t))
    for k in range(len(t)):
        tp, tm = t.copy(), t.copy()
        tp[k] += math.pi / 2
        tm[k] -= math.pi / 2
        g[k] = (f(tp) - f(tm)) / 2
    return g


@pytest.mark.parametrize("x64", [False, True])
@pytest.mark.parametrize("jit", [False, True])
def test_wide_blocks_expval_and_gradient(x64, jit):
    import jaqsi_amd as J
    from jaqsi_amd import _native as nat, simulation

    print(f"wide blocks x64={x64} jit={jit}", flush=True)
    J.enable_x64(x64)
    simulation.specialise(nat.JQ_JIT_SYNC if jit else nat.JQ_JIT_OFF)
    try:
        p = np.random.default_rng(31).uniform(0, 2 * np.pi, 13)
        pt = torch.tensor(p, device="cuda", requires_grad=True)
        obs = [J.PauliZ(wires=q, record=False) for q in range(5)]
        ev = J.Script(_wide_circuit, n_qubits=5).execute(type="expval", obs=obs, args=(pt,))
        ev.sum().backward()
        ops = _wide_ops(p)
        vals, grads, _, _ = sv.adjoint(ops, 5, sv.z_obs(5))
        assert_parity(ev.detach().cpu().numpy(), vals, x64)
        assert_parity(pt.grad.cpu().numpy(), sv.flat_grads(ops, grads), x64)
    finally:
        simulation.specialise(None)


@pytest.mark.parametrize("x64", [False, True])
def test_singular_channels_gradient(x64):
    """BitFlip(0.5), PhaseFlip(0.5) (singular), DepolarizingChannel(0.7),
    AmplitudeDamping(0.99 / 0.999) (condition 15..1000) run in wide-block
    passes with checkpointed inputs: finite gradients equal to the exact
    parameter shift of the density oracle."""
    import jaqsi_amd as J

    print(f"singular channels x64={x64}", flush=True)
    J.enable_x64(x64)
    t = np.random.default_rng(5).uniform(0, 2 * np.pi, 8)
    tt = torch.tensor(t, device="cuda", requires_grad=True)
    obs = [J.PauliZ(wires=0, record=False), J.PauliZ(wires=1, record=False)]
    ev = J.Script(_noisy_strong, n_qubits=2).execute(type="expval", obs=obs, args=(tt,))
    ev.sum().backward()
    items = _noisy_strong_items(t)
    v, g = dm.value_and_shift_grad(items, 2, [(sv.PZ, (0,)), (sv.PZ, (1,))], np.array([1.0, 1.0]))
    assert np.all(np.isfinite(tt.grad.cpu().numpy()))
    assert_parity(float(ev.detach().sum()), v, x64, scale=1.0)
    assert_parity(tt.grad.cpu().numpy(), g, x64, scale=max(1.0, np.abs(g).max()))


def _small_circuit(t, n):
    import jaqsi_amd as J

    for q in range(n):
        J.RY(t[q], wires=q)
    for q in range(n - 1):
        J.CX(wires=[q, q + 1])
    for q in range(n):
        J.RX(t[n + q], wires=q)
    J.RZZ(t[2 * n], wires=[0, n - 1])


def _small_ops(t, n):
    O = sv.Op
    ops = [O("RY", (q,), (t[q],)) for q in range(n)]
    ops += [O("CX", (q, q + 1)) for q in range(n - 1)]
    ops += [O("RX", (q,), (t[n + q],)) for q in range(n)]
    ops += [O("RZZ", (0, n - 1), (t[2 * n],))]
    return ops


@pytest.mark.parametrize("x64", [False, True])
def test_gradients_of_probs_state_and_density(x64):
    """Reverse mode through type='probs', 'state' and 'density' (the
    reference gets these from jax.grad): one forward, one adjoint sweep
    seeded with the cotangent of the final state.  Each loss is a linear
    functional of |psi><psi|, so the parameter shift of the oracle is exact."""
    import jaqsi_amd as J

    print(f"probs/state/density gradients x64={x64}", flush=True)
    J.enable_x64(x64)
    n = 4
    rng = np.random.default_rng(17)
    t = rng.uniform(0, 2 * np.pi, 2 * n + 1)
    w = rng.normal(size=1 << n)
    c = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    C = rng.normal(size=(1 << n, 1 << n)) + 1j * rng.normal(size=(1 << n, 1 << n))
    C = C + C.conj().T
    s = J.Script(lambda a: _small_circuit(a, n), n_qubits=n)
    cdt = torch.complex128 if x64 else torch.complex64

    losses = {
        "probs": (lambda out: (out * torch.tensor(w, device="cuda", dtype=out.dtype)).sum(),
                  lambda psi: float(np.sum(w * np.abs(psi) ** 2))),
        "state": (lambda out: (out * torch.tensor(c, device="cuda", dtype=cdt)).sum().abs() ** 2,
                  lambda psi: float(abs(np.sum(c * psi)) ** 2)),
        "density": (lambda out: (out * torch.tensor(C.T, device="cuda", dtype=cdt)).sum().real,
                    lambda psi: float(np.real(np.conj(psi) @ C @ psi))),
    }
    for kind, (loss_t, loss_np) in losses.items():
        tt = torch.tensor(t, device="cuda", requires_grad=True)
        out = s.execute(type=kind, args=(tt,))
        L = loss_t(out)
        L.backward()
        f = lambda a: loss_np(sv.simulate(_small_ops(a, n), n))  # noqa: E731
        assert_parity(float(L.detach()), f(t), x64, scale=max(1.0, abs(f(t))))
        g = _shift_grad(f, t)
        assert_parity(tt.grad.cpu().numpy(), g, x64, scale=max(1.0, np.abs(g).max()), err_msg=kind)


@pytest.mark.parametrize("x64", [False, True])
def test_gradients_of_noisy_outputs_and_general_observables(x64):
    """Density-matrix probs and a non-diagonal observable Tr(X_0 Y_1 rho) with
    gradients (the reference differentiates simulate_mixed by autodiff)."""
    import jaqsi_amd as J
    from jaqsi_amd import noise

    print(f"noisy outputs gradients x64={x64}", flush=True)
    J.enable_x64(x64)
    n = 3
    rng = np.random.default_rng(23)
    t = rng.uniform(0, 2 * np.pi, 2 * n + 1)
    w = rng.normal(size=1 << n)

    def circ(a):
        _small_circuit(a, n)
        noise.DepolarizingChannel(0.05, wires=1)
        noise.AmplitudeDamping(0.2, wires=2)

    def items(a):
        return _small_ops(a, n) + [("kraus", dm.kraus("DepolarizingChannel", 0.05), (1,)),
                                   ("kraus", dm.kraus("AmplitudeDamping", 0.2), (2,))]

    s = J.Script(circ, n_qubits=n)
    tt = torch.tensor(t, device="cuda", requires_grad=True)
    pr = s.execute(type="probs", args=(tt,))
    L = (pr * torch.tensor(w, device="cuda", dtype=pr.dtype)).sum()
    L.backward()
    f = lambda a: float(np.sum(w * dm.probs(dm.simulate(items(a), n))))  # noqa: E731
    assert_parity(float(L.detach()), f(t), x64, scale=1.0)
    assert_parity(tt.grad.cpu().numpy(), _shift_grad(f, t), x64, scale=1.0)

    tt = torch.tensor(t, device="cuda", requires_grad=True)
    ob = J.PauliX(0, record=False) @ J.PauliY(1, record=False)
    ev = s.execute(type="expval", obs=[ob], args=(tt,))
    ev.sum().backward()
    XY = np.kron(sv.PX, sv.PY)
    f = lambda a: dm.expval(dm.simulate(items(a), n), XY, (0, 1), n)  # noqa: E731
    assert_parity(float(ev.sum().detach()), f(t), x64, scale=1.0)
    assert_parity(tt.grad.cpu().numpy(), _shift_grad(f, t), x64, scale=1.0)


def test_batched_noisy_probs_in_axes():
    """Script.execute(type='probs', args=(thetas,), in_axes=(0,)) on a circuit
    with a channel (the reference's tests/test_jaqsi.py:1803-1825 case), with
    the gradient of a weighted sum of every sample's probabilities."""
    import jaqsi_amd as J
    from jaqsi_amd import noise

    print("batched noisy probs", flush=True)
    n = 4

    def circuit(theta):
        for i in range(n):
            J.H(wires=i)
        J.RX(theta, wires=0)
        noise.BitFlip(0.01, wires=0)

    thetas = np.linspace(0, np.pi, 5)
    th = torch.tensor(thetas, device="cuda", requires_grad=True)
    res = J.Script(circuit, n_qubits=n).execute(type="probs", args=(th,), in_axes=(0,))
    assert res.shape == (5, 2**n)
    w = np.random.default_rng(3).normal(size=(5, 2**n))
    (res * torch.tensor(w, device="cuda", dtype=res.dtype)).sum().backward()
    O = sv.Op
    for b, t0 in enumerate(thetas):
        def f(a):
            its = [O("H", (i,)) for i in range(n)] + [O("RX", (0,), (a[0],)), ("kraus", dm.kraus("BitFlip", 0.01), (0,))]
            return dm.probs(dm.simulate(its, n))
        assert_parity(res[b].detach().cpu().numpy(), f([t0]), False)
        g = _shift_grad(lambda a: float(np.sum(w[b] * f(a))), np.array([t0]))
        assert_parity(th.grad[b].item(), g[0], False, scale=1.0)


def test_backward_reuses_checkpoints_and_recovers_when_overwritten():
    """Two expectation values of one cached plan at different angles, then a
    single backward: the second forward overwrote the first one's
    checkpoints, so its backward runs the forward again -- both gradients
    stay exact."""
    import jaqsi_amd as J

    print("checkpoint token", flush=True)
    J.enable_x64(True)
    n = 5
    rng = np.random.default_rng(41)
    t1, t2 = rng.uniform(0, 2 * np.pi, 2 * n + 1), rng.uniform(0, 2 * np.pi, 2 * n + 1)
    s = J.Script(lambda a: _small_circuit(a, n), n_qubits=n)
    obs = [J.PauliZ(wires=q, record=False) for q in range(n)]
    a1 = torch.tensor(t1, device="cuda", requires_grad=True)
    a2 = torch.tensor(t2, device="cuda", requires_grad=True)
    e1 = s.execute(type="expval", obs=obs, args=(a1,))
    e2 = s.execute(type="expval", obs=obs, args=(a2,))
    (e1.sum() + 2 * e2.sum()).backward()
    for a, t, wgt in ((a1, t1, 1.0), (a2, t2, 2.0)):
        ops = _small_ops(t, n)
        _, grads, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
        assert_parity(a.grad.cpu().numpy(), wgt * sv.flat_grads(ops, grads), True)


def _noise_param_circuit(theta, p, gamma, t_g, q):
    """RX/RY on two wires, CX, then channels whose parameters are tensors."""
    import jaqsi_amd as J
    from jaqsi_amd import noise

    J.Gates.RX(theta[0], wires=0)
    J.Gates.RY(theta[1], wires=1)
    noise.BitFlip(p, wires=0)
    J.Gates.CX(wires=[0, 1], noise_params={"MultiQubitDepolarizing": q})
    noise.AmplitudeDamping(gamma, wires=1)
    noise.PhaseDamping(gamma, wires=0)
    noise.DepolarizingChannel(p, wires=1)
    noise.ThermalRelaxationError(0.1, 1.0, 0.8, t_g, wires=0)
    J.Gates.RY(theta[2], wires=0)


def _noise_param_items(v):
    """The same circuit for the density oracle."""
    t, p, g, t_g, q = v[:3], v[3], v[4], v[5], v[6]
    return [sv.Op("RX", (0,), (t[0],)), sv.Op("RY", (1,), (t[1],)), ("kraus", dm.kraus("BitFlip", p), (0,)),
            sv.Op("CX", (0, 1)), ("kraus", dm.kraus("MultiQubitDepolarizing", q, 2), (0, 1)),
            ("kraus", dm.kraus("AmplitudeDamping", g), (1,)), ("kraus", dm.kraus("PhaseDamping", g), (0,)),
            ("kraus", dm.kraus("DepolarizingChannel", p), (1,)),
            ("kraus", dm.kraus("ThermalRelaxationError", 0.1, 1.0, 0.8, t_g), (0,)), sv.Op("RY", (0,), (t[2],))]


@pytest.mark.parametrize("x64", [False, True])
def test_gradients_reach_noise_parameters(x64):
    """Channel strengths given as tensors that require grad: the adjoint's
    gradient of each channel's superoperator flows back to them (and to the
    angles), checked against central differences of the complex128 density
    oracle on the same circuit."""
    import jaqsi_amd as J

    print(f"noise-parameter gradients x64={x64}", flush=True)
    J.enable_x64(x64)
    try:
        vals = np.array([0.7, 1.3, -0.4, 0.12, 0.3, 0.25, 0.05])
        x = torch.tensor(vals, dtype=torch.float64, device="cuda", requires_grad=True)
        script = J.Script(lambda v: _noise_param_circuit(v[:3], v[3], v[4], v[5], v[6]), n_qubits=2)
        obs = [J.PauliZ(wires=0, record=False), J.PauliX(wires=1, record=False)]
        out = script.execute(type="expval", obs=obs, args=(x,))
        ct = torch.tensor([1.0, -0.6], dtype=out.dtype, device="cuda")
        (out * ct).sum().backward()

        c = ct.double().cpu().numpy()

        def f(v):
            rho = dm.simulate(_noise_param_items(v), 2)
            return c[0] * dm.expval(rho, sv.PZ, (0,), 2) + c[1] * dm.expval(rho, sv.PX, (1,), 2)

        # sixth-order central differences: truncation ~h^6, rounding ~1e-16/h
        h, w = 5e-3, np.array([-1.0, 9.0, -45.0, 0.0, 45.0, -9.0, 1.0]) / 60.0
        fd = np.array([sum(c * f(vals + k * h * e) for c, k in zip(w, range(-3, 4)) if c) / h
                       for e in np.eye(len(vals))])
        assert_parity(x.grad.cpu().numpy(), fd, x64, scale=np.abs(fd).max(), err_msg="noise-parameter gradient")
    finally:
        J.enable_x64(False)


def test_one_plan_on_two_streams_at_once():
    """Per-stream workspaces: forwards with checkpoints on two streams, then
    both backwards, with no synchronisation in between; each stream's
    gradient matches the oracle (its checkpoints were not overwritten by the
    other stream's forward)."""
    import jaqsi_amd as J
    from jaqsi_amd import _native as nat, lowering
    from jaqsi_amd.simulation import _Runner
    from jaqsi_amd.tape import recording

    print("two streams", flush=True)
    n = 6
    rng = np.random.default_rng(77)
    ts = [rng.uniform(0, 2 * np.pi, 2 * n + 1) for _ in range(2)]
    with recording() as tape:
        _small_circuit(ts[0], n)
    low = lowering.lower_pure(tape, n)
    plan = nat.Plan(n, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, 1)
    m = lowering.matrix_tensor(low.matrix_values, low.n_matrix_elems, None)
    dev = torch.device("cuda", 0)
    m = m.reshape(1, -1).to(dev) if m is not None else None
    diags = [(q, 1.0, -1.0) for q in range(n)]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    angles = [torch.tensor(t, dtype=torch.float64, device=dev).reshape(1, -1) for t in ts]
    ct = torch.ones((1, n), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    outs = []
    runs = []
    for s, a in zip(streams, angles):
        with torch.cuda.stream(s):
            r = _Runner(plan, 1, False, diags, n, False, 1)
            ev, state, tok = r.forward_ckpt(a, m, None)
            runs.append((r, ev, state, tok))
    for s, a, (r, ev, state, tok) in zip(streams, angles, runs):
        with torch.cuda.stream(s):
            ga, _, _ = r.backward(tok, a, m, None, state, ct=ct)
            outs.append((ev, ga))
    torch.cuda.synchronize()
    for t, (ev, ga) in zip(ts, outs):
        ops = _small_ops(t, n)
        vals, grads, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
        assert_parity(ev[0].cpu().numpy(), vals, True)
        assert_parity(ga[0].cpu().numpy(), sv.flat_grads(ops, grads), True)
