"""jq_sv_forward_noalloc (the forward that allocates nothing: its
workspaces live in a fixed per-device scratch array) against jq_sv_forward
on the same plan, on two calls in a row, the second with a smaller batch;
and the clear error for a call whose workspace exceeds the scratch."""

import numpy as np
import pytest
import torch

from circuits import record_hea
from parity import assert_parity

pytestmark = pytest.mark.gpu


def _plan_and_inputs(n, layers, batch, seed):
    from jaqsi_amd import _native as nat, lowering
    from jaqsi_amd.tape import recording

    theta = np.zeros((layers, n, 3))
    with recording() as tape:
        record_hea(theta, n)
    low = lowering.lower_pure(tape, n)
    plan = nat.Plan(n, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, 0)
    angles = torch.tensor(np.random.default_rng(seed).uniform(0, 2 * np.pi, (batch, len(low.angle_values))),
                          device="cuda")
    mats = lowering.matrix_tensor(low.matrix_values, low.n_matrix_elems, None)
    plan.mats = mats.to("cuda").reshape(1, -1) if mats is not None else None
    return plan, angles


def _run(entry, plan, angles, n):
    from jaqsi_amd import _native as nat

    B = angles.shape[0]
    state = torch.empty((B, 1 << n), dtype=torch.complex64, device="cuda")
    ev = torch.empty((B, n), dtype=torch.float64, device="cuda")
    wires = np.arange(n, dtype=np.int32)
    d = np.tile(np.array([1.0, -1.0]), n)
    stream = torch.cuda.current_stream().cuda_stream
    mats = plan.mats.expand(B, -1).contiguous() if plan.mats is not None else None
    nat.check(getattr(nat.lib(), entry)(plan.handle, B, nat.ptr(angles), nat.ptr(mats), None, 0, nat.ptr(state),
                                        nat.JQ_OBS_SV_DIAG, n, nat.ptr(wires), nat.ptr(d), nat.ptr(ev), stream))
    torch.cuda.synchronize()
    return state.cpu().numpy(), ev.cpu().numpy()


def test_noalloc_forward_matches_forward_on_two_calls():
    print("no-alloc forward", flush=True)
    n = 12
    plan, angles = _plan_and_inputs(n, 3, 8, 5)
    ref1 = _run("jq_sv_forward", plan, angles, n)  # also makes the program resident
    got1 = _run("jq_sv_forward_noalloc", plan, angles, n)
    small = angles[:3].flip(0).contiguous()
    ref2 = _run("jq_sv_forward", plan, small, n)
    got2 = _run("jq_sv_forward_noalloc", plan, small, n)
    for (gs, ge), (rs, re) in ((got1, ref1), (got2, ref2)):
        assert_parity(gs, rs, False)
        assert_parity(ge, re, False)
        assert np.array_equal(gs, rs) and np.array_equal(ge, re)  # the same kernels on the same inputs


def test_noalloc_forward_rejects_oversized_workspaces():
    from jaqsi_amd import _native as nat

    print("no-alloc limits", flush=True)
    n = 6
    plan, angles = _plan_and_inputs(n, 40, 2, 9)
    _run("jq_sv_forward", plan, angles, n)
    pool_elems = plan.info()["n_blocks"] * 4
    big = (nat.JQ_NOALLOC_SCRATCH_BYTES // (pool_elems * 8)) + 1
    assert big <= 65535
    a = torch.zeros((big, angles.shape[1]), dtype=torch.float64, device="cuda")
    state = torch.empty((1, 1 << n), dtype=torch.complex64, device="cuda")
    mats = plan.mats.expand(big, -1).contiguous()
    rc = nat.lib().jq_sv_forward_noalloc(plan.handle, big, nat.ptr(a), nat.ptr(mats), None, 0, nat.ptr(state),
                                         nat.JQ_OBS_NONE, 0, None, None, None,
                                         torch.cuda.current_stream().cuda_stream)
    assert rc == nat.JQ_ERR_ARG
    assert b"JQ_NOALLOC_SCRATCH_BYTES" in nat.lib().jq_last_error()
