"""Plans whose trailing permutation / diagonal gates are measured through
the observables (fold_tail) against the same circuit simulated in full and
against the complex128 oracle."""

import ctypes

import numpy as np
import pytest
import torch

from circuits import hea_theta, record_hea
from oracle import sv
from parity import assert_parity

pytestmark = pytest.mark.gpu


def _mixed(theta, n):
    """Rotation layers, then X / SWAP / CZ / S / CX gates on every wire, with
    one parameterised gate among them (it and what precedes it on its wires
    stay in the passes)."""
    import jaqsi_amd as J

    for layer in range(theta.shape[0]):
        for q in range(n):
            J.RX(theta[layer, q, 0], q)
            J.RY(theta[layer, q, 1], q)
        for q in range(0, n - 1, 2):
            J.CZ([q, q + 1])
    J.PauliX(2)
    J.CX([2, 5])
    J.RZ(theta[0, 0, 2], 5)
    for q in range(n):
        J.CX([q, (q + 3) % n])
    J.SWAP([0, n - 1])
    J.S(1)
    J.PauliX(n - 2)
    J.CZ([3, 4])


def _mixed_ops(theta, n):
    O = sv.Op
    ops = []
    for layer in range(theta.shape[0]):
        for q in range(n):
            ops += [O("RX", (q,), (theta[layer, q, 0],)), O("RY", (q,), (theta[layer, q, 1],))]
        ops += [O("CZ", (q, q + 1)) for q in range(0, n - 1, 2)]
    ops += [O("PauliX", (2,)), O("CX", (2, 5)), O("RZ", (5,), (theta[0, 0, 2],))]
    ops += [O("CX", (q, (q + 3) % n)) for q in range(n)]
    return ops + [O("SWAP", (0, n - 1)), O("S", (1,)), O("PauliX", (n - 2,)), O("CZ", (3, 4))]


def _setup(n, layers, seed, fold, dtype=0):
    from jaqsi_amd import _native as nat, lowering
    from jaqsi_amd.tape import recording

    theta = hea_theta(n, layers, seed)
    with recording() as tape:
        _mixed(theta, n)
    low = lowering.lower_pure(tape, n)
    plan = nat.Plan(n, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, dtype, fold_tail=fold)
    mats = lowering.matrix_tensor(low.matrix_values, low.n_matrix_elems, None)
    mats = mats.reshape(1, -1).cuda() if mats is not None else None
    angles = torch.tensor([[float(a) for a in low.angle_values]], dtype=torch.float64, device="cuda")
    return plan, angles, mats, theta


def _observables(n):
    """Z on every wire, |1><1| on wire 1 (d = (0, 1)), then three Pauli-Z words."""
    from jaqsi_amd import _native as nat

    diag_w = np.arange(n, dtype=np.int32).tolist() + [1]
    diag_d = [1.0, -1.0] * n + [0.0, 1.0]
    words = [(0, 3), (1, 2, n - 1), (4,)]
    par_w, par_d = [], []
    for wd in words:
        m = sum(1 << w for w in wd)
        par_w += [m & 0xFFFFFFFF, m >> 32]
        par_d += [1.0, -1.0]
    return [(nat.JQ_OBS_SV_DIAG, np.array(diag_w, np.int32), np.array(diag_d)),
            (nat.JQ_OBS_SV_PARITY, np.array(par_w, np.int32), np.array(par_d))], words


def _adjoint(plan, angles, mats, kind, wires, d, n_obs, ct, dtype=0):
    from jaqsi_amd import _native as nat

    n = plan.n_wires
    cdt = torch.complex64 if dtype == 0 else torch.complex128
    state = torch.empty((1, 1 << n), dtype=cdt, device="cuda")
    ev = torch.empty((1, n_obs), dtype=torch.float64, device="cuda")
    ga = torch.empty_like(angles)
    ctt = torch.tensor(ct.reshape(1, -1), dtype=torch.float64, device="cuda")
    nat.check(nat.lib().jq_sv_adjoint(plan.handle, 1, nat.ptr(angles), nat.ptr(mats), None, 0, nat.ptr(state), kind,
                                      n_obs, nat.ptr(wires), nat.ptr(d), nat.ptr(ctt), nat.ptr(ev), nat.ptr(ga), None,
                                      None, torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return ev.cpu().numpy()[0], ga.cpu().numpy()[0]


@pytest.mark.parametrize("n,layers", [(9, 2), (15, 2)])
def test_folded_plan_matches_oracle_and_full_plan(n, layers):
    print(f"fold mixed n={n} layers={layers}", flush=True)
    folded, angles, mats, theta = _setup(n, layers, 900 + n, 1)
    full, _, _, _ = _setup(n, layers, 900 + n, 0)
    assert folded.fold()[0] > 0 and full.fold()[0] == 0
    ops = _mixed_ops(theta, n)
    (dk, dw, dd), (pk, pw, pd) = _observables(n)[0]
    words = _observables(n)[1]
    rng = np.random.default_rng(n)
    # single-wire observables
    ct = rng.normal(size=n + 1)
    evf, gf = _adjoint(folded, angles, mats, dk, dw, dd, n + 1, ct)
    evp, gp = _adjoint(full, angles, mats, dk, dw, dd, n + 1, ct)
    obs = [(sv.PZ, (q,)) for q in range(n)] + [(np.diag([0.0, 1.0]).astype(complex), (1,))]
    vals, grads, _, _ = sv.adjoint(ops, n, obs, ct=ct)
    assert_parity(evf, vals, False)
    assert_parity(evf, evp, False)
    assert_parity(gf, sv.flat_grads(ops, grads), False)
    assert_parity(gf, gp, False)
    # Pauli-Z words
    ct = rng.normal(size=len(words))
    evf, gf = _adjoint(folded, angles, mats, pk, pw, pd, len(words), ct)
    evp, gp = _adjoint(full, angles, mats, pk, pw, pd, len(words), ct)
    obs = [(np.diag([(-1.0) ** bin(i).count("1") for i in range(1 << len(w))]).astype(complex), tuple(w))
           for w in words]
    vals, grads, _, _ = sv.adjoint(ops, n, obs, ct=ct)
    assert_parity(evf, vals, False)
    assert_parity(evf, evp, False)
    assert_parity(gf, sv.flat_grads(ops, grads), False)
    assert_parity(gf, gp, False)


def test_folded_final_ring_at_20_qubits_matches_full_plan():
    """Plan-specialised kernels (20 wires): the HEA with its last CX ring
    folded gives the expectation values and gradients of the full plan."""
    from jaqsi_amd import _native as nat, lowering
    from jaqsi_amd.tape import recording

    n, layers = 20, 3
    print(f"fold ring n={n} layers={layers}", flush=True)
    theta = hea_theta(n, layers, 2020)
    with recording() as tape:
        record_hea(theta, n)
    low = lowering.lower_pure(tape, n)
    angles = torch.tensor(theta.reshape(1, -1), dtype=torch.float64, device="cuda")
    mats = lowering.matrix_tensor(low.matrix_values, low.n_matrix_elems, None)
    mats = mats.reshape(1, -1).cuda() if mats is not None else None
    wires = np.arange(n, dtype=np.int32)
    d = np.tile(np.array([1.0, -1.0]), n)
    ct = np.ones(n)
    out = {}
    for fold in (0, 1):
        plan = nat.Plan(n, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, 0, fold_tail=fold)
        assert plan.jit(nat.JQ_JIT_SYNC) == nat.JQ_JIT_LOADED
        out[fold] = (_adjoint(plan, angles, mats, nat.JQ_OBS_SV_DIAG, wires, d, n, ct), plan.info()["n_passes"])
        if fold:
            assert plan.fold()[0] == n
    assert out[1][1] < out[0][1]
    assert_parity(out[1][0][0], out[0][0][0], False)
    assert_parity(out[1][0][1], out[0][0][1], False)


def test_folded_plan_refuses_what_needs_the_final_vector():
    from jaqsi_amd import _native as nat

    n = 9
    plan, angles, mats, _ = _setup(n, 1, 5, 1)
    L = nat.lib()
    st = torch.cuda.current_stream().cuda_stream
    state = torch.empty((1, 1 << n), dtype=torch.complex64, device="cuda")
    ga = torch.empty_like(angles)
    lam = torch.zeros((1, 1 << n), dtype=torch.complex64, device="cuda")
    tok = ctypes.c_int64(-1)
    nat.check(L.jq_sv_forward_ckpt(plan.handle, 1, nat.ptr(angles), nat.ptr(mats), None, 0, nat.ptr(state),
                                   nat.JQ_OBS_NONE, 0, None, None, None, ctypes.byref(tok), st))
    rc = L.jq_sv_backward(plan.handle, tok.value, 1, nat.ptr(angles), nat.ptr(mats), None, 0, nat.ptr(state),
                          nat.JQ_OBS_NONE, 0, None, None, None, nat.ptr(lam), nat.ptr(ga), None, None, st)
    assert rc == nat.JQ_ERR_UNSUPPORTED
    w = np.zeros(1, np.int32)
    dd = np.array([1.0, -1.0])
    ev = torch.empty((1, 1), dtype=torch.float64, device="cuda")
    rc = L.jq_sv_forward(plan.handle, 1, nat.ptr(angles), nat.ptr(mats), None, 0, nat.ptr(state),
                         nat.JQ_OBS_DM_DIAG, 1, nat.ptr(w), nat.ptr(dd), nat.ptr(ev), st)
    assert rc == nat.JQ_ERR_UNSUPPORTED
    torch.cuda.synchronize()


def test_script_expval_folds_the_final_ring():
    import jaqsi_amd as J
    from jaqsi_amd import simulation

    n, layers = 12, 3
    print(f"script fold n={n}", flush=True)
    J.enable_x64(False)
    theta = hea_theta(n, layers, 1212)
    th = torch.tensor(theta, dtype=torch.float64, device="cuda", requires_grad=True)
    obs = [J.PauliZ(wires=q, record=False) for q in range(n)]
    ev = J.Script(lambda t: record_hea(t, n), n_qubits=n).execute(type="expval", obs=obs, args=(th,))
    ev.sum().backward()
    assert any(p.fold()[0] == n for p in simulation.cached_plans())
    ops = sv.hea_ops(theta, n)
    vals, grads, _, _ = sv.adjoint(ops, n, sv.z_obs(n))
    assert_parity(ev.detach().cpu().numpy(), vals, False)
    assert_parity(th.grad.cpu().numpy().reshape(-1), sv.flat_grads(ops, grads), False)
    # a state output of the same circuit simulates the ring
    st = J.Script(lambda t: record_hea(t, n), n_qubits=n).execute(type="state", args=(theta,))
    assert_parity(st.cpu().numpy().astype(np.complex128), sv.simulate(ops, n), False)
