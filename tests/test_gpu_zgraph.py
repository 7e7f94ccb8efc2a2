"""The adjoint entry points recorded into a HIP graph (torch.cuda.CUDAGraph)
and replayed: every replay, with eager work and synchronisations between
replays and new angles copied into the captured input, matches an eager call
on the same angles.  Both kernel paths (prebuilt and plan-specialised) and
the ragged-batch call (side streams forked and joined by events)."""

import numpy as np
import pytest
import torch

from circuits import hea_theta, record_hea

pytestmark = pytest.mark.gpu

N, LAYERS = 14, 2


def _plan(jit_mode):
    from jaqsi_amd import _native as nat
    from jaqsi_amd import lowering
    from jaqsi_amd.tape import recording

    with recording() as tape:
        record_hea(hea_theta(N, LAYERS, 0), N)
    low = lowering.lower_pure(tape, N)
    plan = nat.Plan(N, low.ops, low.members, len(low.angle_values), low.n_matrix_elems, 0)
    plan.jit(jit_mode)
    xm = lowering.matrix_tensor(low.matrix_values, low.n_matrix_elems, None)
    mats = xm.reshape(1, -1).to("cuda").contiguous() if xm is not None else None
    return plan, len(low.angle_values), mats


@pytest.mark.parametrize("specialised", [False, True])
def test_sv_adjoint_replays_from_a_graph(specialised):
    from jaqsi_amd import _native as nat

    print(f"graph sv adjoint specialised={specialised}", flush=True)
    L = nat.lib()
    plan, P, mats = _plan(nat.JQ_JIT_SYNC if specialised else nat.JQ_JIT_OFF)
    if specialised:
        assert plan.jit() == nat.JQ_JIT_LOADED
    dev = torch.device("cuda")
    thetas = [torch.tensor(hea_theta(N, LAYERS, 300 + k).reshape(1, -1), dtype=torch.float64, device=dev)
              for k in range(3)]
    angles = thetas[0].clone()
    ct = torch.ones((1, N), dtype=torch.float64, device=dev)
    state = torch.empty((1, 1 << N), dtype=torch.complex64, device=dev)
    ev = torch.empty((1, N), dtype=torch.float64, device=dev)
    ga = torch.empty((1, P), dtype=torch.float64, device=dev)
    wires = np.arange(N, dtype=np.int32)
    d = np.tile(np.array([1.0, -1.0]), N)

    def call(stream):
        nat.check(L.jq_sv_adjoint(plan.handle, 1, nat.ptr(angles), nat.ptr(mats), None, 0, nat.ptr(state),
                                  nat.JQ_OBS_SV_DIAG, N, nat.ptr(wires), nat.ptr(d), nat.ptr(ct), nat.ptr(ev),
                                  nat.ptr(ga), None, None, stream))

    eager = []
    for th in thetas:
        angles.copy_(th)
        call(torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        eager.append((ev.clone(), ga.clone()))
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        call(side.cuda_stream)  # this stream's workspaces are sized before the capture
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=side):  # the stream whose workspaces the warm-up sized
        call(torch.cuda.current_stream().cuda_stream)
    for rep in range(2):
        for k, th in enumerate(thetas):
            angles.copy_(th)
            noise = torch.randn(1 << 16, device=dev).sum()  # eager work between replays
            g.replay()
            torch.cuda.synchronize()
            assert torch.isfinite(noise)
            np.testing.assert_allclose(ev.cpu().numpy(), eager[k][0].cpu().numpy(), atol=1e-6, rtol=0)
            np.testing.assert_allclose(ga.cpu().numpy(), eager[k][1].cpu().numpy(), atol=1e-5, rtol=0)


@pytest.mark.parametrize("specialised", [False, True])
def test_ragged_groups_replay_from_a_graph(specialised):
    """Prebuilt kernels: one launch per pass step over all groups (group table
    in kernel arguments); specialised: each group on a side stream."""
    from jaqsi_amd import _native as nat
    from jaqsi_amd.ragged import RaggedBatch, record

    print(f"graph ragged groups specialised={specialised}", flush=True)
    depths, per = (2, 3), (5, 3)
    tapes = [record(record_hea, np.zeros((dd, N, 3)), N) for dd in depths]
    rb = RaggedBatch(tapes, list(per), N, [(q, 1.0, -1.0) for q in range(N)])
    phases = rb.specialise(nat.JQ_JIT_SYNC if specialised else nat.JQ_JIT_OFF)
    if specialised:
        assert all(p == nat.JQ_JIT_LOADED for p in phases)
    dev = torch.device("cuda")
    rng = np.random.default_rng(9)
    sets = [[torch.tensor(rng.uniform(0, 2 * np.pi, (b, dd * N * 3)), dtype=torch.float64, device=dev)
             for dd, b in zip(depths, per)] for _ in range(2)]
    angles = [a.clone() for a in sets[0]]
    cts = [torch.ones((b, N), dtype=torch.float64, device=dev) for b in per]
    eager = []
    for s in sets:
        for a, v in zip(angles, s):
            a.copy_(v)
        evs, grads = rb.adjoint(angles, cts)
        torch.cuda.synchronize()
        eager.append(([e.clone() for e in evs], [x.clone() for x in grads]))
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        rb.adjoint(angles, cts)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):  # the stream whose workspaces the warm-up sized
        evs, grads = rb.adjoint(angles, cts)
    for rep in range(2):
        for k, s in enumerate(sets):
            for a, v in zip(angles, s):
                a.copy_(v)
            torch.randn(1 << 16, device=dev).sum()
            g.replay()
            torch.cuda.synchronize()
            for e, ref in zip(evs, eager[k][0]):
                np.testing.assert_allclose(e.cpu().numpy(), ref.cpu().numpy(), atol=1e-6, rtol=0)
            for x, ref in zip(grads, eager[k][1]):
                np.testing.assert_allclose(x.cpu().numpy(), ref.cpu().numpy(), atol=1e-5, rtol=0)
